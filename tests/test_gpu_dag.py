"""GPU parity of octree DAG leaves (DAGIntersect, MultiRenderer/BVH2Common.cpp:2365-2557) against
the oracle on the same arrays: the reference's engine-scene DAG file in every intersection mode and
the LOD, iteration, intersection-count and channel render modes (hit-point quotients instead of
texture coordinates) with and without the compact surface offset; DAGs encoded from sphere-union
octrees with rotated and shifted bricks, surface nodes and LOD bricks; instances beside a mesh and
a compact octree (the mixed kernel), the persistent kernel and any-hit. Records are bit-identical
except under the analytic cubic (acos/cos/pow differ from glibc's in the last ulps), held there to
the SDF bar of conftest.sdf_check.
"""
import os
import sys

import numpy as np
import pytest

import dag_encode as de
from conftest import MISS, raw_equal
from test_gpu_coctree import SPHERES, _check, _preset, _rays

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DAG = os.path.join(GOLD, "engine_sdf_dag_0.bin")


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
def ref_dag(lrt):
    return lrt.load_sdf_dag(DAG)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("eps", [0.0, 0.05])
def test_reference_dag_modes(lrt, orc, torch_dev, ref_dag, mode, eps):
    g = lrt
    _say(f"dag reference file, intersection mode {mode}, compact eps {eps}")
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SdfDAG(*ref_dag))
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g)
    p = _preset(g, mode)
    p["compact_sdf_eps"] = eps
    got = _check(torch_dev, g, orc, s, pos, dirs, p)
    assert (got["primId"] != MISS).mean() > 0.1


@pytest.mark.parametrize("render", [20, 10, 26, 23, 27, 8])
def test_reference_dag_render_modes(lrt, orc, torch_dev, ref_dag, render):
    g = lrt
    _say(f"dag reference file, render mode {render}")
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SdfDAG(*ref_dag))
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g)
    got = _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 0, render))
    if render in (23, 27):
        hit = got["primId"] != MISS
        assert np.any(got["coords"][hit, 0].view(np.uint32) != 0)


ENCODED = [dict(bs=1), dict(bs=2, rotate=True, add=True), dict(bs=2, rotate=True, surface=True),
           dict(bs=1, lod=True, surface=True), dict(bs=1, channel_rotate=True), dict(bs=2, rotate=True, channel_rotate=True)]


@pytest.fixture(scope="module")
def encoded(lrt):
    cache = {}

    def get(i):
        if i not in cache:
            c = dict(ENCODED[i])
            bs = c.pop("bs")
            nodes, values = lrt.sdf_spheres_global_octree(SPHERES, 5 if bs == 1 else 4, bs, 0)
            cache[i] = de.encode(nodes, values, bs, seed=i, **c)
        return cache[i]
    return get


@pytest.mark.parametrize("case", range(len(ENCODED)))
def test_encoded_dags(lrt, orc, torch_dev, encoded, case):
    g = lrt
    _say(f"dag encoded case {case}: {ENCODED[case]}")
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SdfDAG(*encoded(case)))
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g)
    got = _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 0))
    assert (got["primId"] != MISS).mean() > 0.1
    _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 2, 8))
    _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 4, 23))
    if ENCODED[case].get("channel_rotate"):   # channel modes: quotients turned by the channel rotation
        for mode, render in ((0, 23), (3, 27), (1, 23)):
            _check(torch_dev, g, orc, s, pos, dirs, _preset(g, mode, render))
    if ENCODED[case].get("lod"):
        for lod in (2.0, 3.0):
            _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 0, 1, fixed_lod=1, lod=lod))
        _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 0, 1, fixed_lod=0, lod=5.0))


def test_dag_instances_mixed_persistent_any(lrt, orc, torch_dev, encoded, ref_dag):
    g = lrt
    torch = torch_dev
    _say("dag instances beside a mesh and a compact octree")
    s = g.SceneRT()
    a = s.AddGeom_SdfDAG(*ref_dag)
    b = s.AddGeom_SdfDAG(*encoded(1))
    hc, wc, _ = g.load_coctree_v3(os.path.join(GOLD, "engine_sdf_coctree_v3_0.bin"))
    c = s.AddGeom_COctreeV3(hc, wc)
    verts, tris = g.gen_icosphere(3, 0.05, 4.0, 3)
    v = np.array(verts, np.float32)
    v[:, :3] *= 0.3
    m = s.AddGeom_Triangles3f(v, tris)
    t1 = np.eye(4, dtype=np.float32)
    t1[:3, 3] = [-0.8, 0.0, 0.0]
    t1[:3, :3] *= 0.6
    cs, sn = np.cos(0.7), np.sin(0.7)
    t2 = np.eye(4, dtype=np.float32)
    t2[:3, :3] = 0.5 * np.array([[cs, -sn, 0], [sn, cs, 0], [0, 0, 1]], np.float32)
    t2[:3, 3] = [0.7, 0.2, -0.3]
    t3 = np.eye(4, dtype=np.float32)
    t3[:3, :3] *= 0.4
    t3[:3, 3] = [0.0, 0.9, 0.2]
    s.AddInstance(a, t1)
    s.AddInstance(b, t2)
    s.AddInstance(c, t3)
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
