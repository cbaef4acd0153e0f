"""GPU parity of SCom2 leaves (SCom2Intersect, MultiRenderer/BVH2Common.cpp:2560-2818) against
the oracle on the same arrays: the reference's engine-scene SCom2 file (bricks shared through
rotations and shifts) in every intersection mode, with and without the compact surface offset,
and the LOD, iteration, intersection-count, channel and normal render modes; instances beside a
DAG, a compact octree and a mesh (the mixed kernel), the persistent kernel and any-hit; the
queries the library refuses (the reference's fast traversal of one plain tree). Records are
bit-identical except under the analytic cubic, held to the SDF bar of conftest.sdf_check.
"""
import os
import sys

import numpy as np
import pytest

from conftest import MISS, raw_equal
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


def test_fast_traversal_case_is_refused(lrt, torch_dev):
    """One plain SCom2 tree (no multi-nodes, no channels, 3^3 values, one-word links) rendered
    LAMBERT_NO_TEX is where the reference switches to SCom2IntersectFast, which is not built: the
    call fails with LRT_ERR_UNSUPPORTED instead of answering differently; other render modes run."""
    g = lrt
    torch = torch_dev
    _say("scom2 fast traversal case")
    h = np.zeros((), g.SCOM2_HEADER_DTYPE)
    h["brick_size"], h["v_size"], h["bits_per_value"], h["values_per_uint"], h["value_mask"] = 2, 3, 8, 4, 0xFF
    h["dimension"], h["uints_per_link"], h["links_offset"], h["child_offset_off"] = 3, 1, 1, 0
    h["children_types_shift"], h["children_types_mask"] = 8, 0xFFFF
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SCom2(h, np.zeros(1, np.uint32), np.zeros(1, np.uint32)))
    s.CommitScene()
    pos, dirs = _rays(torch, g, res=16)
    hits = torch.zeros((pos.shape[0], 8), dtype=torch.int32, device="cuda")
    s.SetPreset(_preset(g, 0, 1))
    with pytest.raises(g.LrtError) as e:
        s.trace_closest(pos, dirs, hits)
    assert e.value.code == 4   # LRT_ERR_UNSUPPORTED
    s.SetPreset(_preset(g, 0, 8))
    s.trace_closest(pos, dirs, hits)
    torch.cuda.synchronize()
    assert np.all(hits.cpu().numpy().view(g.HIT_DTYPE)["primId"] == MISS)
