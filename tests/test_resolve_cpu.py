"""The oracle's g-buffer resolve and tonemap (MultiRenderer_resolve.cpp:10-75, 529-1005) pinned by
closed forms: tonemap packing, Lambert shading of known normals under the default lights, the
depth modes, palettes, barycentrics, the pixel <-> g-buffer slot mapping of both ray orders, and
a planted mistake (a dropped ambient term) that the Lambert check must catch.
"""
import numpy as np
import pytest

from conftest import host_scene

MISS = 0xFFFFFFFF


def test_tonemap_closed_form(orc):
    c = np.array([[1.0, 0.5, 0.0, 1.0], [-0.2, 2.0, 0.999, 0.25], [0.5, 0.5, 0.5, 0.5], [4.0, 2.0, 1.0, 0.0]],
                 np.float32)
    out = orc.tonemap(c, 1)
    want = []
    for r in c:
        q = [int(255 * min(max(0.0, float(np.float32(x))), 1.0)) for x in r]
        want.append((q[3] << 24) | (q[2] << 16) | (q[1] << 8) | q[0])
    assert out.tolist() == want
    # two accumulated samples are averaged first
    assert orc.tonemap(2 * c[2:3], 2)[0] == want[2]


def _sdf_scene(g):
    rng = np.random.default_rng(2)
    sph = np.concatenate([rng.uniform(-0.8, 0.8, (24, 3)), rng.uniform(0.1, 0.3, (24, 1))], 1).astype(np.float32)
    return host_scene(g, lambda s: s.AddInstance(s.AddGeom_SdfFrameOctree(g.sdf_spheres_frame_octree(sph, 6))))


def _frame(orc, g, s, w=64, h=64, order=1):
    pi, vi = g.camera_inverse([0, 0, 3], [0, 0, 0], [0, 1, 0], 60.0, 1.0, 0.01, 100.0)
    pos, dirs = orc.eye_rays(pi, vi, w, h, order=order)
    hits, _ = orc.trace_closest(orc.Scene.from_scene(s), pos, dirs)
    return pi, vi, pos, dirs, hits


def _pixel_hits(hits, w, h, order):
    """Hit of each row-major pixel from a g-buffer in `order`."""
    if order == 0:
        return hits
    idx = np.zeros(w * h, np.int64)
    pw = 8 * ((w + 7) // 8)
    for y in range(h):
        for x in range(w):
            local, bx, by = (y & 3) * 4 + (x & 3), x >> 2, y >> 2
            lh, wbh = (by & 1) * 2 + (bx & 1), (pw >> 2) >> 1
            idx[y * w + x] = ((bx >> 1) + (by >> 1) * wbh) * 64 + lh * 16 + local
    return hits[idx]


def _decode(e):
    from conftest import decode_normal
    return decode_normal(e)


def test_lambert_closed_form_on_sdf(lrt, orc):
    """LAMBERT_NO_TEX, default lights: colour = 2/3 max(0, n . normalize(1,1,1)) + 0 + 0.25 with
    n the decoded hit normal; misses take the background; alpha 1."""
    g = lrt
    s = _sdf_scene(g)
    pi, vi, pos, dirs, hits = _frame(orc, g, s)
    col = orc.resolve(orc.Scene.from_scene(s), s.get_array("orig_indices"), hits, pi, vi, 64, 64, order=1,
                      lights=g.default_lights(), background=(0.1, 0.2, 0.3, 1.0))
    ph = _pixel_hits(hits, 64, 64, 1)
    hit = ph["primId"] != MISS
    assert 300 < hit.sum() < 64 * 64
    n = _decode(ph["coords"][hit, 2:])
    l = np.ones(3) / np.sqrt(3.0)
    want = 2.0 / 3.0 * np.maximum(0.0, n @ l) + 0.25
    for k in range(3):
        assert np.abs(col[hit, k] - want).max() < 2e-6
    assert np.all(col[hit, 3] == 1.0)
    assert np.allclose(col[~hit], [0.1, 0.2, 0.3, 1.0])


def test_planted_missing_ambient_is_caught(lrt, orc):
    """Dropping the 0.25 ambient light (a plausible mistake) breaks the closed form above."""
    g = lrt
    s = _sdf_scene(g)
    pi, vi, pos, dirs, hits = _frame(orc, g, s)
    lights = g.default_lights()[:2]
    col = orc.resolve(orc.Scene.from_scene(s), s.get_array("orig_indices"), hits, pi, vi, 64, 64, lights=lights)
    ph = _pixel_hits(hits, 64, 64, 1)
    hit = ph["primId"] != MISS
    want = 2.0 / 3.0 * np.maximum(0.0, _decode(ph["coords"][hit, 2:]) @ (np.ones(3) / np.sqrt(3.0))) + 0.25
    assert np.abs(col[hit, 0] - want).max() > 0.2


def test_triangle_geometry_normal_closed_form(lrt, orc):
    """NORMAL_MODE_GEOMETRY on an instanced, translated and scaled icosphere: the shaded normal is
    the triangle's face normal turned to face the ray, so colour = 2/3 max(0, n . l) + 0.25 with
    n from the triangle's vertices in float64."""
    g = lrt
    v, f = g.gen_icosphere(3, 0.0, 4.0, 1)
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] *= 0.8
    m[:3, 3] = [0.1, -0.2, 0.3]
    s = host_scene(g, lambda s: s.AddInstance(s.AddGeom_Triangles3f(v, f), m))
    p = g.default_preset()
    p["normal_mode"] = 0
    s.SetPreset(p)
    pi, vi, pos, dirs, hits = _frame(orc, g, s, order=0)
    col = orc.resolve(orc.Scene.from_scene(s), s.get_array("orig_indices"), hits, pi, vi, 64, 64, order=0,
                      lights=g.default_lights())
    hit = hits["primId"] != MISS
    assert hit.sum() > 500
    tri = f[hits["primId"][hit]].astype(np.int64)
    A, B, C = (v[tri[:, k], :3].astype(np.float64) for k in range(3))
    nrm = np.cross(B - A, C - A)     # uniform scale: the world normal is parallel to the local one
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = dirs[hit, :3].astype(np.float64)
    nrm *= np.sign(np.sum(-d * nrm, axis=1))[:, None]
    want = 2.0 / 3.0 * np.maximum(0.0, nrm @ (np.ones(3) / np.sqrt(3.0))) + 0.25
    assert np.abs(col[hit, 0] - want).max() < 1e-5
    # the primitive palette names the hit triangle
    p["render_mode"] = 5
    s.SetPreset(p)
    col5 = orc.resolve(orc.Scene.from_scene(s), s.get_array("orig_indices"), hits, pi, vi, 64, 64, order=0)
    pal = np.array([0xffe6194b, 0xff3cb44b, 0xffffe119, 0xff0082c8, 0xfff58231, 0xff911eb4, 0xff46f0f0, 0xfff032e6,
                    0xffd2f53c, 0xfffabebe, 0xff008080, 0xffe6beff, 0xffaa6e28, 0xfffffac8, 0xff800000, 0xffaaffc3,
                    0xff808000, 0xffffd8b1, 0xff000080, 0xff808080], np.uint32)
    c = pal[hits["primId"][hit] % 20]
    assert np.array_equal(col5[hit, 0], (c & 0xFF).astype(np.float32) * np.float32(1.0 / 255.0))
    assert np.array_equal(col5[hit, 2], ((c >> 16) & 0xFF).astype(np.float32) * np.float32(1.0 / 255.0))


@pytest.mark.parametrize("mode", [2, 3, 4, 9])
def test_depth_and_barycentric_modes(lrt, orc, mode):
    g = lrt
    v, f = g.gen_icosphere(3, 0.05, 4.0, 1)
    s = host_scene(g, lambda s: s.AddInstance(s.AddGeom_Triangles3f(v, f)))
    p = g.default_preset()
    p["render_mode"] = mode
    s.SetPreset(p)
    pi, vi, pos, dirs, hits = _frame(orc, g, s, order=0)
    col = orc.resolve(orc.Scene.from_scene(s), s.get_array("orig_indices"), hits, pi, vi, 64, 64, order=0)
    hit = hits["primId"] != MISS
    z = hits["t"][hit].astype(np.float64)
    want = {2: (1 / z - 10.0) / (0.1 - 10.0), 3: (z - 0.1) / 9.9, 4: 1 - (z - 0.1) / 9.9,
            9: hits["coords"][hit, 0].astype(np.float64)}[mode]
    assert np.abs(col[hit, 0] - want).max() < 1e-5
    if mode == 9:
        c = hits["coords"][hit]
        assert np.abs(col[hit, 2] - (1 - c[:, 0] - c[:, 1])).max() < 1e-6


def test_orders_resolve_to_the_same_image(lrt, orc):
    """A g-buffer traced in super-block order resolves to the same row-major image as one traced
    in row order."""
    g = lrt
    s = _sdf_scene(g)
    pi, vi, _, _, h1 = _frame(orc, g, s, order=1)
    _, _, _, _, h0 = _frame(orc, g, s, order=0)
    c1 = orc.resolve(orc.Scene.from_scene(s), s.get_array("orig_indices"), h1, pi, vi, 64, 64, order=1,
                     lights=g.default_lights())
    c0 = orc.resolve(orc.Scene.from_scene(s), s.get_array("orig_indices"), h0, pi, vi, 64, 64, order=0,
                     lights=g.default_lights())
    assert np.array_equal(c0, c1)
