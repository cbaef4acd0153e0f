"""GPU parity of compact-octree (COctreeV3) leaves (COctreeIntersectV3 and its brick march and
loaders, MultiRenderer/BVH2Common.cpp:1637-2173) against the oracle on the same arrays: the
reference's engine-scene file at two BLAS levels, every intersection mode and the LOD, iteration,
intersection-count and smoothed-normal modes; octrees encoded with each leaf type, value width,
padding, pack mode, rotation, additive shift, texture coordinates and LOD leaves; instances beside a
mesh (the mixed kernel), the persistent kernel and the any-hit query. Records are bit-identical
except under the analytic cubic, whose acos/cos/pow differ from glibc's in the last ulps: there
the rows are held to the SDF bar of conftest.sdf_check.
"""
import os
import sys

import numpy as np
import pytest

import coct_encode as ce
from conftest import MISS, assert_sdf_parity, raw_equal, sdf_check

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
COCT = os.path.join(GOLD, "engine_sdf_coctree_v3_0.bin")
SPHERES = np.array([[0.1, 0.05, -0.05, 0.55], [-0.35, 0.2, 0.25, 0.3], [0.3, -0.3, 0.3, 0.25]], np.float32)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    return torch


def _rays(torch, g, res=160, eye=(0.9, 1.1, 2.6)):
    pi, vi = g.camera_inverse(list(eye), [0, 0, 0], [0, 1, 0], 55.0, 1.0, 0.01, 100.0)
    pos = torch.zeros((res * res, 4), dtype=torch.float32, device="cuda")
    dirs = torch.zeros_like(pos)
    jit = torch.from_numpy(np.random.default_rng(5).random((res * res, 2), dtype=np.float32)).cuda()
    g.eye_rays(pi, vi, res, res, pos, dirs, jitter=jit, order=1)
    torch.cuda.synchronize()
    return pos, dirs


def _check(torch, g, orc, s, pos, dirs, preset):
    s.SetPreset(preset)
    hits = torch.zeros((pos.shape[0], 8), dtype=torch.int32, device="cuda")
    s.trace_closest(pos, dirs, hits)
    torch.cuda.synchronize()
    got = hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)
    want, ov = orc.trace_closest(orc.Scene.from_scene(s), pos.cpu().numpy(), dirs.cpu().numpy())
    assert ov == 0
    if int(preset["sdf_node_intersect"]) == 1:   # analytic cubic: the SDF bar
        r = sdf_check(got, want, pos.cpu().numpy(), dirs.cpu().numpy(), colour=True)
        assert_sdf_parity(r, got, want)
    else:
        rows = int(np.sum(np.any(got.view(np.uint8).reshape(-1, 32) != want.view(np.uint8).reshape(-1, 32), 1)))
        assert raw_equal(got, want), f"{rows} rows differ"
    return got


def _preset(g, mode=0, render=1, normal=0, fixed_lod=0, lod=16.0):
    p = g.default_preset()
    p["sdf_node_intersect"], p["render_mode"], p["normal_mode"] = mode, render, normal
    p["fixed_lod"], p["level_of_detail"] = fixed_lod, lod
    return p


@pytest.fixture(scope="module")
def ref_file(lrt):
    return lrt.load_coctree_v3(COCT)[:2]


@pytest.mark.parametrize("level", [0, 2])
@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_reference_file_modes(lrt, orc, torch_dev, ref_file, level, mode):
    g = lrt
    _say(f"coctree reference file, BLAS level {level}, intersection mode {mode}")
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_COctreeV3(*ref_file, bvh_level=level))
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g)
    got = _check(torch_dev, g, orc, s, pos, dirs, _preset(g, mode))
    assert (got["primId"] != MISS).mean() > 0.1


@pytest.mark.parametrize("render,normal", [(20, 0), (10, 0), (26, 0), (1, 2), (8, 2)])
def test_reference_file_render_modes(lrt, orc, torch_dev, ref_file, render, normal):
    g = lrt
    _say(f"coctree reference file, render mode {render}, normal mode {normal}")
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_COctreeV3(*ref_file))
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g)
    _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 0, render, normal))


ENCODED = [
    dict(bs=1, pad=0, bits=8, leaf_type=ce.GRID),
    dict(bs=1, pad=1, bits=10, leaf_type=ce.BIT_PACK, pack_mode=1, rotate=True, sim=1),
    dict(bs=2, pad=0, bits=16, leaf_type=ce.SLICES),
    dict(bs=2, pad=1, bits=8, leaf_type=ce.SLICES, pack_mode=2, rotate=True, sim=1, uv=1),
    dict(bs=3, pad=0, bits=32, leaf_type=ce.GRID, pack_mode=1, rotate=True),
    dict(bs=2, pad=1, bits=6, leaf_type=ce.GRID, uv=1),
    dict(bs=1, pad=1, bits=16, leaf_type=ce.GRID, lods=1),
]


@pytest.fixture(scope="module")
def encoded(lrt):
    cache = {}

    def get(i):
        if i not in cache:
            c = dict(ENCODED[i])
            bs, pad = c.pop("bs"), c.pop("pad")
            nodes, values = lrt.sdf_spheres_global_octree(SPHERES, 5 if bs == 1 else 4, bs, pad)
            cache[i] = ce.encode(nodes, values, bs, pad, seed=i, **c)
        return cache[i]
    return get


@pytest.mark.parametrize("case", range(len(ENCODED)))
def test_encoded_octrees(lrt, orc, torch_dev, encoded, case):
    g = lrt
    c = ENCODED[case]
    _say(f"coctree encoded case {case}: {c}")
    h, w = encoded(case)
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_COctreeV3(h, w))
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g)
    got = _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 0))
    assert (got["primId"] != MISS).mean() > 0.1
    _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 2, 8))
    if c["pad"] > 0:
        _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 0, 1, normal=2))
    if c.get("lods"):
        for lod in (2.0, 3.0):
            _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 0, 20, fixed_lod=1, lod=lod))
        _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 0, 1, fixed_lod=0, lod=5.0))


def test_instances_mesh_persistent_and_any_hit(lrt, orc, torch_dev, encoded, ref_file):
    """Two compact octrees (the reference file and an encoded one) in four instances, one rotated
    and scaled, beside a triangle mesh: the mixed kernel, the persistent kernel and any-hit."""
    g = lrt
    torch = torch_dev
    _say("coctree instances beside a mesh")
    s = g.SceneRT()
    a = s.AddGeom_COctreeV3(*ref_file, bvh_level=1)
    b = s.AddGeom_COctreeV3(*encoded(3))
    verts, tris = g.gen_icosphere(3, 0.05, 4.0, 3)
    v = np.array(verts, np.float32).reshape(-1, verts.shape[-1])
    v[:, :3] *= 0.3
    m = s.AddGeom_Triangles3f(v, tris)
    t1 = np.eye(4, dtype=np.float32)
    t1[:3, 3] = [-0.8, 0.0, 0.0]
    t1[:3, :3] *= 0.6
    c, sn = np.cos(0.5), np.sin(0.5)
    t2 = np.eye(4, dtype=np.float32)
    t2[:3, :3] = 0.5 * np.array([[c, 0, sn], [0, 1, 0], [-sn, 0, c]], np.float32)
    t2[:3, 3] = [0.7, 0.2, -0.3]
    t3 = np.eye(4, dtype=np.float32)
    t3[:3, 3] = [0.0, 0.9, 0.2]
    s.AddInstance(a, t1)
    s.AddInstance(b, t2)
    s.AddInstance(a, t3 @ np.diag([0.4, 0.4, 0.4, 1]).astype(np.float32))
    s.AddInstance(m, np.eye(4, dtype=np.float32))
    s.CommitScene()
    pos, dirs = _rays(torch, g, eye=(0.2, 0.6, 3.2))
    got = _check(torch, g, orc, s, pos, dirs, _preset(g, 0, 8))
    assert len(set((got["geomId"][got["primId"] != MISS] >> 24).tolist())) >= 2
    s.set_persistent(1)
    got_p = _check(torch, g, orc, s, pos, dirs, _preset(g, 0, 8))
    assert raw_equal(got_p, got)
    s.set_persistent(0)
    flags = torch.zeros(pos.shape[0], dtype=torch.int32, device="cuda")
    s.trace_any(pos, dirs, flags)
    torch.cuda.synchronize()
    want = orc.trace_any(orc.Scene.from_scene(s), pos.cpu().numpy(), dirs.cpu().numpy())
    assert np.array_equal(flags.cpu().numpy().astype(np.uint32), want)
