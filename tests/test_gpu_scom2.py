"""GPU parity of SCom2 leaves (SCom2Intersect, MultiRenderer/BVH2Common.cpp:2560-2818) against
the oracle on the same arrays: the reference's engine-scene SCom2 file (bricks shared through
rotations and shifts) in every intersection mode, with and without the compact surface offset,
and the LOD, iteration, intersection-count, channel and normal render modes; instances beside a
DAG, a compact octree and a mesh (the mixed kernel), the persistent kernel and any-hit; the
reference's fast traversal of one plain tree (encoded by tests/scom2_encode.py). Records are
bit-identical except under the analytic cubic (which the fast walk always solves, BBOX aside),
held to the SDF bar of conftest.sdf_check.
"""
import os
import sys

import numpy as np
import pytest

from conftest import MISS, assert_sdf_parity, raw_equal, sdf_check
from test_gpu_coctree import _check, _preset, _rays

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    return torch


@pytest.fixture(scope="module")
def ref_tree(lrt):
    return lrt.load_scom2(os.path.join(GOLD, "engine_scom2_0.bin"))


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("eps", [0.0, 0.05])
def test_reference_scom2_modes(lrt, orc, torch_dev, ref_tree, mode, eps):
    g = lrt
    _say(f"scom2 reference file, intersection mode {mode}, compact eps {eps}")
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SCom2(*ref_tree))
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g)
    p = _preset(g, mode)
    p["compact_sdf_eps"] = eps
    got = _check(torch_dev, g, orc, s, pos, dirs, p)
    assert (got["primId"] != MISS).mean() > 0.1


@pytest.mark.parametrize("render", [20, 10, 26, 23, 27, 8, 5])
def test_reference_scom2_render_modes(lrt, orc, torch_dev, ref_tree, render):
    g = lrt
    _say(f"scom2 reference file, render mode {render}")
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SCom2(*ref_tree))
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g)
    _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 0, render))
    for lod in (3.0, 5.0):
        _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 0, render, fixed_lod=1, lod=lod))


def test_scom2_instances_mixed_persistent_any(lrt, orc, torch_dev, ref_tree):
    g = lrt
    torch = torch_dev
    _say("scom2 instances beside a DAG, a compact octree and a mesh")
    s = g.SceneRT()
    a = s.AddGeom_SCom2(*ref_tree)
    b = s.AddGeom_SdfDAG(*g.load_sdf_dag(os.path.join(GOLD, "engine_sdf_dag_0.bin")))
    hc, wc, _ = g.load_coctree_v3(os.path.join(GOLD, "engine_sdf_coctree_v3_0.bin"))
    c = s.AddGeom_COctreeV3(hc, wc, 1)
    verts, tris = g.gen_icosphere(3, 0.05, 4.0, 3)
    v = np.array(verts, np.float32)
    v[:, :3] *= 0.3
    m = s.AddGeom_Triangles3f(v, tris)
    mats = []
    for k, (tx, ty, sc) in enumerate([(-0.8, 0.0, 0.6), (0.7, 0.2, 0.5), (0.0, 0.9, 0.4), (0.2, -0.8, 0.35)]):
        t = np.eye(4, dtype=np.float32)
        ang = 0.4 * k
        t[:3, :3] = sc * np.array([[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]], np.float32)
        t[:3, 3] = [tx, ty, -0.2]
        mats.append(t)
    s.AddInstance(a, mats[0])
    s.AddInstance(b, mats[1])
    s.AddInstance(c, mats[2])
    s.AddInstance(a, mats[3])
    s.AddInstance(m, np.eye(4, dtype=np.float32))
    s.CommitScene()
    pos, dirs = _rays(torch, g, eye=(0.2, 0.6, 3.2))
    got = _check(torch, g, orc, s, pos, dirs, _preset(g, 0, 8))
    assert len(set((got["geomId"][got["primId"] != MISS] >> 24).tolist())) >= 3
    s.set_persistent(1)
    assert raw_equal(_check(torch, g, orc, s, pos, dirs, _preset(g, 0, 8)), got)
    s.set_persistent(0)
    flags = torch.zeros(pos.shape[0], dtype=torch.int32, device="cuda")
    s.trace_any(pos, dirs, flags)
    torch.cuda.synchronize()
    want = orc.trace_any(orc.Scene.from_scene(s), pos.cpu().numpy(), dirs.cpu().numpy())
    assert np.array_equal(flags.cpu().numpy().astype(np.uint32), want)


@pytest.fixture(scope="module")
def plain_trees(lrt):
    import scom2_encode as se
    from test_coctree_cpu import SPHERES
    nodes, values = lrt.sdf_spheres_global_octree(SPHERES, 5, 2, 0)
    return {"rotated": se.encode(nodes, values, rotate=True, seed=3),
            "mixed": se.encode(nodes, values, rotate=True, shift=True, unique_every=5, volume_every=7, seed=4)}


def _check_fast(torch, g, orc, s, pos, dirs, preset):
    """The fast walk's surface step (LocalSurfaceIntersectionFast, BVH2Common.cpp:2820-2986) solves
    the analytic cubic in every intersection mode but BBOX, so its records are held to the SDF bar
    (the device's acosf/cosf/powf differ from glibc's in the last ulps); BBOX hits are exact."""
    if int(preset["sdf_node_intersect"]) == 3:
        return _check(torch, g, orc, s, pos, dirs, preset)
    s.SetPreset(preset)
    hits = torch.zeros((pos.shape[0], 8), dtype=torch.int32, device="cuda")
    s.trace_closest(pos, dirs, hits)
    torch.cuda.synchronize()
    got = hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)
    want, ov = orc.trace_closest(orc.Scene.from_scene(s), pos.cpu().numpy(), dirs.cpu().numpy())
    assert ov == 0
    assert_sdf_parity(sdf_check(got, want, pos.cpu().numpy(), dirs.cpu().numpy(), colour=True), got, want)
    return got


@pytest.mark.parametrize("key", ["rotated", "mixed"])
@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("eps", [0.0, 0.05])
def test_fast_traversal_plain_tree(lrt, orc, torch_dev, plain_trees, key, mode, eps):
    """One plain SCom2 tree (no multi-nodes, no channels, 3^3 values, one-word links) rendered
    LAMBERT_NO_TEX takes the fast traversal (SCom2IntersectFast, BVH2Common.cpp:2988-3231) in
    every intersection mode; NORMAL rendering of the same tree takes the general walk."""
    g = lrt
    _say(f"scom2 plain tree {key}, intersection mode {mode}, compact eps {eps}")
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SCom2(*plain_trees[key]))
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g)
    for render in (1, 8):
        p = _preset(g, mode, render)
        p["compact_sdf_eps"] = eps
        got = (_check_fast if render == 1 else _check)(torch_dev, g, orc, s, pos, dirs, p)
        assert (got["primId"] != MISS).mean() > 0.1


def test_fast_traversal_instances_persistent_any(lrt, orc, torch_dev, plain_trees):
    """Instances of the one plain tree beside a mesh (the mixed kernel) keep the fast walk; the
    persistent kernel and any-hit agree."""
    g = lrt
    torch = torch_dev
    _say("scom2 plain tree instances beside a mesh")
    s = g.SceneRT()
    a = s.AddGeom_SCom2(*plain_trees["mixed"])
    verts, tris = g.gen_icosphere(3, 0.05, 4.0, 3)
    v = np.array(verts, np.float32)
    v[:, :3] *= 0.3
    m = s.AddGeom_Triangles3f(v, tris)
    for k, (tx, ty, sc) in enumerate([(-0.8, 0.0, 0.6), (0.7, 0.2, 0.5), (0.1, 0.8, 0.4)]):
        t = np.eye(4, dtype=np.float32)
        ang = 0.5 * k
        t[:3, :3] = sc * np.array([[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]], np.float32)
        t[:3, 3] = [tx, ty, -0.2]
        s.AddInstance(a, t)
    s.AddInstance(m, np.eye(4, dtype=np.float32))
    s.CommitScene()
    pos, dirs = _rays(torch, g, eye=(0.2, 0.6, 3.2))
    got = _check_fast(torch, g, orc, s, pos, dirs, _preset(g, 0, 1))
    assert len(set((got["geomId"][got["primId"] != MISS] >> 24).tolist())) >= 2
    s.set_persistent(1)
    assert raw_equal(_check_fast(torch, g, orc, s, pos, dirs, _preset(g, 0, 1)), got)
    s.set_persistent(0)
    flags = torch.zeros(pos.shape[0], dtype=torch.int32, device="cuda")
    s.trace_any(pos, dirs, flags)
    torch.cuda.synchronize()
    want = orc.trace_any(orc.Scene.from_scene(s), pos.cpu().numpy(), dirs.cpu().numpy())
    assert np.array_equal(flags.cpu().numpy().astype(np.uint32), want)


@pytest.fixture(scope="module")
def channel_trees(lrt):
    import scom2_encode as se
    from test_coctree_cpu import SPHERES
    out = {}
    for bs in (1, 2):
        nodes, values = lrt.sdf_spheres_global_octree(SPHERES, 5 if bs == 1 else 4, bs, 0)
        out[bs] = se.encode(nodes, values, brick_size=bs, channels=True, rotate=True, channel_rotate=True,
                            shift=True, unique_every=6, volume_every=9, seed=7)
    return out


@pytest.mark.parametrize("bs", [1, 2])
@pytest.mark.parametrize("mode", [0, 1, 3, 4])
def test_multi_node_channel_trees(lrt, orc, torch_dev, channel_trees, bs, mode):
    """Multi-node trees whose leaves carry channel links (tests/scom2_encode.py): primId is the
    channel link, and the channel render modes turn the quotients by the composed channel rotation
    (BVH2Common.cpp:2731-2747)."""
    g = lrt
    _say(f"scom2 multi-node channel tree, brick size {bs}, intersection mode {mode}")
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SCom2(*channel_trees[bs]))
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g)
    for render in (23, 27, 8, 5, 1):
        got = _check(torch_dev, g, orc, s, pos, dirs, _preset(g, mode, render))
        assert (got["primId"] != MISS).mean() > 0.1
