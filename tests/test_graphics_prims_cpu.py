"""Graphics primitives (BVHRT::AddGeom_GraphicsPrim, BVH2Common_host.cpp:767-903;
IntersectGraphicPrims and rayCapsuleIntersect, BVH2Common.cpp:1184-1440) in the oracle, pinned by
the geometry of each kind: a point's hits lie on its sphere with radial normals; a segment's
lie on its capped cylinder (radius r from the axis inside the segment, on a cap plane at its
ends); an infinite line's at radius r from the line (its BLAS box is the reference's clip of
the line against the unit cube's planes); a wire box's at distance ra from its nearest edge; custom colours are packed as
(r + g / 256, b). Arrows (cylinder + cone) are checked for hits only here and bit for bit on the
GPU.
"""
import numpy as np
import pytest

from conftest import MISS, decode_normal, host_scene


def _scene(lrt, orc, ptype, pts, color=(0.25, 0.5, 0.75), res=96, eye=(0.7, 0.9, 3.2)):
    s = host_scene(lrt, lambda s: s.AddInstance(s.AddGeom_GraphicsPrim(ptype, pts, color)))
    p = lrt.default_preset()
    p["render_mode"] = 1
    s.SetPreset(p)
    pi, vi = lrt.camera_inverse(list(eye), [0, 0, 0], [0, 1, 0], 50.0, 1.0, 0.01, 100.0)
    pos, dirs = orc.eye_rays(pi, vi, res, res)
    hits, _ = orc.trace_closest(orc.Scene.from_scene(s), pos, dirs)
    hit = hits["primId"] != MISS
    P = pos[hit, :3].astype(np.float64) + hits["t"][hit, None].astype(np.float64) * dirs[hit, :3]
    return hits, hit, P, s


def test_points_on_their_spheres(lrt, orc):
    rng = np.random.default_rng(1)
    pts = np.concatenate([rng.uniform(-0.8, 0.8, (20, 3)), rng.uniform(0.05, 0.15, (20, 1))], 1).astype(np.float32)
    hits, hit, P, s = _scene(lrt, orc, lrt.GRAPH_PRIM_POINT, pts)
    assert hit.sum() > 200
    assert s.get_array("geoms")["type"][0] == 17
    c = pts[hits["primId"][hit]]
    d = np.linalg.norm(P - c[:, :3], axis=1)
    assert np.abs(d - c[:, 3]).max() < 1e-4
    n = decode_normal(hits["coords"][hit][:, 2:4])
    assert np.degrees(np.arccos(np.clip((n * (P - c[:, :3]) / d[:, None]).sum(1), -1, 1))).max() < 0.1
    assert np.allclose(hits["coords"][hit][:, 0], np.float32(0.25) + np.float32(0.5) / np.float32(256.0))
    assert np.allclose(hits["coords"][hit][:, 1], 0.75)


def test_segments_on_their_capped_cylinders_with_own_colours(lrt, orc):
    rng = np.random.default_rng(2)
    k = 12
    a = rng.uniform(-0.7, 0.7, (k, 3))
    b = a + rng.normal(size=(k, 3)) * 0.3
    r = rng.uniform(0.03, 0.08, k)
    col = rng.uniform(0, 1, (k, 3))
    pts = np.zeros((3 * k, 4), np.float32)
    pts[0::3, :3], pts[0::3, 3] = a, r
    pts[1::3, :3] = b
    pts[2::3, :3] = col
    hits, hit, P, _ = _scene(lrt, orc, lrt.GRAPH_PRIM_LINE_SEGMENT_COLOR, pts)
    assert hit.sum() > 100
    i = hits["primId"][hit]
    A, B, R = pts[0::3][i, :3].astype(np.float64), pts[1::3][i, :3].astype(np.float64), r[i]
    ax = B - A
    L = np.linalg.norm(ax, axis=1)
    u = ax / L[:, None]
    along = ((P - A) * u).sum(1)
    radial = np.linalg.norm((P - A) - along[:, None] * u, axis=1)
    body = (along > 1e-4 * L) & (along < L * (1 - 1e-4))
    assert np.abs(radial[body] - R[body]).max() < 2e-4
    cap = ~body
    assert np.all(np.minimum(np.abs(along[cap]), np.abs(along[cap] - L[cap])) < 2e-4)
    assert np.all(radial[cap] <= R[cap] + 2e-4)
    c = pts[2::3][i, :3]
    assert np.array_equal(hits["coords"][hit][:, 0], c[:, 0] + c[:, 1] / np.float32(256.0))
    assert np.array_equal(hits["coords"][hit][:, 1], c[:, 2])


def test_infinite_lines_clipped_to_the_unit_cube(lrt, orc):
    pts = np.array([[0.2, 0.1, 0.0, 0.05], [0.9, 0.8, 0.3, 0.0],
                    [-0.5, 0.3, 0.2, 0.04], [-0.1, -0.4, 0.9, 0.0]], np.float32)
    hits, hit, P, _ = _scene(lrt, orc, lrt.GRAPH_PRIM_LINE, pts, res=128, eye=(0.3, 0.6, 3.5))
    assert hit.sum() > 50
    i = hits["primId"][hit]
    A, B, R = pts[0::2][i, :3].astype(np.float64), pts[1::2][i, :3].astype(np.float64), pts[0::2][i, 3]
    u = (B - A) / np.linalg.norm(B - A, axis=1, keepdims=True)
    along = ((P - A) * u).sum(1)
    radial = np.linalg.norm((P - A) - along[:, None] * u, axis=1)
    assert np.abs(radial - R).max() < 2e-4


def test_wire_boxes_at_capsule_radius_from_edges(lrt, orc):
    pts = np.array([[-0.6, -0.5, -0.4, 0.03], [0.5, 0.4, 0.6, 0.0]], np.float32)
    hits, hit, P, _ = _scene(lrt, orc, lrt.GRAPH_PRIM_BOX, pts, res=160)
    assert hit.sum() > 100
    lo, hi = pts[0, :3].astype(np.float64), pts[1, :3].astype(np.float64)
    best = np.full(P.shape[0], np.inf)
    for axis in range(3):
        o = [a for a in range(3) if a != axis]
        for c0 in (lo[o[0]], hi[o[0]]):
            for c1 in (lo[o[1]], hi[o[1]]):
                q = np.clip(P[:, axis], lo[axis], hi[axis])
                e = np.zeros_like(P)
                e[:, axis] = q
                e[:, o[0]], e[:, o[1]] = c0, c1
                best = np.minimum(best, np.linalg.norm(P - e, axis=1))
    assert np.abs(best - 0.03).max() < 2e-4


def test_arrows_are_hit(lrt, orc):
    pts = np.array([[-0.6, -0.3, 0.0, 0.04], [0.6, 0.4, 0.1, 0.0]], np.float32)
    hits, hit, P, _ = _scene(lrt, orc, lrt.GRAPH_PRIM_LINE_SEGMENT_DIR, pts, res=160)
    assert hit.sum() > 50
    n = decode_normal(hits["coords"][hit][:, 2:4])
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5)


def test_bad_point_counts_are_rejected(lrt):
    s = lrt.SceneRT()
    with pytest.raises(lrt.LrtError):
        s.AddGeom_GraphicsPrim(lrt.GRAPH_PRIM_BOX_COLOR, np.zeros((4, 4), np.float32))
    with pytest.raises(lrt.LrtError):
        s.AddGeom_GraphicsPrim(12, np.zeros((2, 4), np.float32))
