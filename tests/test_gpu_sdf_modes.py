"""GPU parity of every SDF leaf kind under every intersection mode, both representation modes and
the debug render modes, against the oracle on the same arrays, rule by rule (conftest.sdf_check):
no hit flips, no id changes beyond a 1e-5*t tie, t within 1e-5 relative, hit points within 1e-4
of the scene extent, decoded normals within 1e-3 rad; every mode but the analytic cubic (the one
that calls acos/cos/pow) bit-exact in every field.

Reference: LocalSurfaceIntersection (MultiRenderer/BVH2Common.cpp:345-725) for the five
intersection modes, OctreeNodeIntersect (:727-749), OctreeBrickIntersect (:884-1112),
OctreeIntersect (:2176-2363), the sparse voxel set's leaf decode (:190-215); render modes core/render_settings.h (ST_ITERATIONS 10,
MO_VOXEL_COUNT 24, MO_FLAGS 25, INTERSECTION_COUNT 26).
"""
import sys

import numpy as np
import pytest

from conftest import MISS, assert_sdf_parity, sdf_check, svs_spheres

pytestmark = pytest.mark.gpu

MODES = {0: "ST", 1: "ANALYTIC", 2: "NEWTON", 3: "BBOX", 4: "IT"}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    return torch


def sphere_set(seed=2, n=64):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(-0.8, 0.8, (n, 3)), rng.uniform(0.05, 0.2, (n, 1))], 1).astype(np.float32)


@pytest.fixture(scope="module")
def scenes(lrt, torch_dev):
    g = lrt
    sp = sphere_set()
    out = {}
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SdfFrameOctree(g.sdf_spheres_frame_octree(sp, 7)))
    s.CommitScene()
    out["frame"] = s
    h, n, v, vf, nd = g.sdf_spheres_sbs_indexed(sp, 4, 8, 1)
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SdfSBS(h, n, v, vf))
    s.CommitScene()
    out["sbs_indexed"] = s
    hd = np.zeros((), g.SBS_HEADER_DTYPE)
    hd["brick_size"], hd["brick_pad"], hd["bytes_per_value"], hd["aux_data"] = 4, 0, 1, g.SDF_SBS_NODE_LAYOUT_DX
    n, v = g.sdf_spheres_sbs(sp, 5, hd, colour_seed=7)
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SdfSBS(hd, n, v))
    s.CommitScene()
    out["sbs_dx1"] = s
    mn, mv = g.sdf_spheres_multi_octree(sp, 7)
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SdfMultiOctree(mn, mv))
    s.CommitScene()
    out["multi"] = s
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SdfSVS(svs_spheres(sp, 48)))
    s.CommitScene()
    out["svs"] = s
    return out


@pytest.fixture(scope="module")
def rays(lrt, torch_dev):
    torch = torch_dev
    w = h = 160
    pi, vi = lrt.camera_inverse([0.3, 0.4, 3.0], [0, 0, 0], [0, 1, 0], 60.0, 1.0, 0.01, 100.0)
    pos = torch.zeros((w * h, 4), dtype=torch.float32, device="cuda")
    dirs = torch.zeros_like(pos)
    jit = torch.from_numpy(np.random.default_rng(11).random((w * h, 2), dtype=np.float32)).cuda()
    lrt.eye_rays(pi, vi, w, h, pos, dirs, jitter=jit, order=1)
    torch.cuda.synchronize()
    return pos, dirs


def _trace(torch, g, s, pos, dirs):
    hits = torch.zeros((pos.shape[0], 8), dtype=torch.int32, device="cuda")
    s.trace_closest(pos, dirs, hits)
    torch.cuda.synchronize()
    return hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)


def _run(lrt, orc, torch, s, pos, dirs, preset):
    s.SetPreset(preset)
    got = _trace(torch, lrt, s, pos, dirs)
    want, _ = orc.trace_closest(orc.Scene.from_scene(s), pos.cpu().numpy(), dirs.cpu().numpy())
    return got, want


@pytest.mark.parametrize("representation", [0, 1])
@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("kind", ["frame", "sbs_indexed", "sbs_dx1", "multi", "svs"])
def test_sdf_intersection_modes(lrt, orc, torch_dev, scenes, rays, kind, mode, representation):
    _say(f"sdf modes kind={kind} mode={MODES[mode]} representation={representation}")
    pos, dirs = rays
    p = lrt.default_preset()
    p["sdf_node_intersect"], p["representation_mode"] = mode, representation
    got, want = _run(lrt, orc, torch_dev, scenes[kind], pos, dirs, p)
    r = sdf_check(got, want, pos.cpu().numpy(), dirs.cpu().numpy(), colour=kind == "sbs_indexed")
    _say(str(r))
    assert r["hits"] > 0.1 * r["n"]
    assert_sdf_parity(r, got, want)
    if mode != 1:
        # only the analytic cubic calls transcendental functions: every other mode is bit-exact
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), \
            f"{int(np.sum(np.any(got.view(np.uint8).reshape(-1, 32) != want.view(np.uint8).reshape(-1, 32), 1)))} rows"
    scenes[kind].SetPreset(lrt.default_preset())


@pytest.mark.parametrize("kind,render_mode,mode", [
    ("frame", 10, 0), ("sbs_indexed", 10, 0), ("sbs_dx1", 10, 0), ("multi", 10, 0),
    ("svs", 10, 0), ("svs", 26, 1),
    ("frame", 26, 1), ("sbs_dx1", 26, 1), ("multi", 26, 1), ("frame", 26, 0),
    ("multi", 24, 1), ("multi", 25, 1), ("multi", 24, 0), ("multi", 25, 3),
    ("multi", 23, 0), ("multi", 27, 2), ("frame", 23, 0)])
def test_sdf_debug_render_modes(lrt, orc, torch_dev, scenes, rays, kind, render_mode, mode):
    """ST_ITERATIONS (primId = sphere-tracing steps), INTERSECTION_COUNT (instId = leaf calls),
    MO_VOXEL_COUNT / MO_FLAGS (primId = the hit node's voxel count / flags), CHANNEL /
    CHANNEL_LAMBERT (multi-octree hits keep the voxel-local hit point in coords[0..1])."""
    _say(f"sdf render mode {render_mode} kind={kind} intersect={MODES[mode]}")
    pos, dirs = rays
    p = lrt.default_preset()
    p["render_mode"], p["sdf_node_intersect"] = render_mode, mode
    got, want = _run(lrt, orc, torch_dev, scenes[kind], pos, dirs, p)
    scenes[kind].SetPreset(lrt.default_preset())
    hit = want["t"] < 1e9
    assert hit.mean() > 0.1
    if mode != 1:
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    else:
        # analytic: t may differ in the last ulps; the debug fields must not
        same = (got["t"] < 1e9) == hit
        assert same.all()
        field = "instId" if render_mode == 26 else "primId"
        assert np.array_equal(got[field], want[field]), f"{int(np.sum(got[field] != want[field]))} rows"
    if render_mode == 26:
        # a miss keeps the count of the leaves it entered; a hit overwrites instId with the
        # instance (0) before the increment, as the reference does
        assert want["instId"][~hit].max() >= 2 and np.all(want["instId"][hit] == 1)
    if render_mode in (24, 25):
        assert np.all(want["primId"][hit] != MISS)
    if render_mode in (23, 27) and kind == "multi":
        assert np.any(want["coords"][hit, 0].view(np.uint32) != 0)


@pytest.fixture(scope="module")
def in_scene(lrt, torch_dev):
    g = lrt
    h, n, v, vf, nd = g.sdf_spheres_sbs_indexed(sphere_set(), 4, 8, 1)
    H, N, V, F = g.sbs_with_neighbors(h, n, v, vf)
    s = g.SceneRT()
    gid = s.AddGeom_SdfSBS(H, N, V, F)
    s.AddInstance(gid)
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float32) * 0.5
    m[:3, 3] = [1.2, 0.3, -0.5]
    s.AddInstance(gid, m)
    s.CommitScene()
    return s


@pytest.mark.parametrize("mode", [0, 1, 3, 4])
def test_sbs_smoothed_normals(lrt, orc, torch_dev, in_scene, rays, mode):
    """Brick set with neighbour indices (ID32F_IRGB32F_IN), normal_mode SDF_SMOOTHED, Lambert:
    the smoothed normals (BVH2Common.cpp:958-1028, read across bricks) against the oracle, two
    instances (one rotated: the smoothed normal stays in the brick set's frame, as in the
    reference); bit-exact except in the analytic mode, which keeps the SDF bar."""
    _say(f"sbs smoothed normals intersect={MODES[mode]}")
    pos, dirs = rays
    p = lrt.default_preset()
    p["render_mode"], p["normal_mode"], p["sdf_node_intersect"] = 1, 2, mode
    got, want = _run(lrt, orc, torch_dev, in_scene, pos, dirs, p)
    in_scene.SetPreset(lrt.default_preset())
    r = sdf_check(got, want, pos.cpu().numpy(), dirs.cpu().numpy(), colour=True)
    _say(str(r))
    assert r["hits"] > 0.1 * r["n"]
    assert_sdf_parity(r, got, want)
    if mode != 1:
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
