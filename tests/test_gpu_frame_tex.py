"""GPU parity of textured frame octrees (AddGeom_SdfFrameOctreeTex, BVH2Common_host.cpp:697-731;
texture coordinates of LocalSurfaceIntersection, BVH2Common.cpp:689-710) against the oracle on the
same arrays: every intersection mode (bit-identical records; the analytic cubic held to the SDF
bar of conftest.sdf_check, texture coordinates within 1e-5), the iteration render mode, transformed instances beside a plain frame
octree and a mesh in the mixed kernel, the persistent kernel and any-hit."""
import sys

import numpy as np
import pytest

from conftest import MISS, assert_sdf_parity, raw_equal, sdf_check
from test_coctree_cpu import SPHERES
from test_frame_tex_cpu import textured, uv_of
from test_gpu_coctree import _check, _preset, _rays

pytestmark = pytest.mark.gpu


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
def tex_nodes(lrt):
    return textured(lrt, lrt.sdf_spheres_frame_octree(SPHERES, 6))


def _check_analytic(torch, g, orc, s, pos, dirs, preset):
    """Under the analytic cubic the voxel-local root can move by an ulp (device acosf/cosf/powf
    against glibc's) while the world-space t rounds alike; the texture coordinates follow the local
    root as the normal does, so they are held to 1e-5 instead of bit equality."""
    s.SetPreset(preset)
    hits = torch.zeros((pos.shape[0], 8), dtype=torch.int32, device="cuda")
    s.trace_closest(pos, dirs, hits)
    torch.cuda.synchronize()
    got = hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)
    want, ov = orc.trace_closest(orc.Scene.from_scene(s), pos.cpu().numpy(), dirs.cpu().numpy())
    assert ov == 0
    assert_sdf_parity(sdf_check(got, want, pos.cpu().numpy(), dirs.cpu().numpy()), got, want)
    both = (got["primId"] != MISS) & (want["primId"] != MISS)
    assert np.max(np.abs(got["coords"][both, :2] - want["coords"][both, :2])) < 1e-5
    return got


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_textured_octree_modes(lrt, orc, torch_dev, tex_nodes, mode):
    g = lrt
    _say(f"textured frame octree, intersection mode {mode}")
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SdfFrameOctreeTex(tex_nodes))
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g)
    if mode == 1:
        got = _check_analytic(torch_dev, g, orc, s, pos, dirs, _preset(g, mode))
    else:
        got = _check(torch_dev, g, orc, s, pos, dirs, _preset(g, mode))
    h = got["primId"] != MISS
    assert h.mean() > 0.1 and np.all((got["geomId"][h] >> 24) == g.TYPE_SDF_FRAME_OCTREE_TEX)
    if mode != 1:
        p = (pos.cpu().numpy()[h, :3].astype(np.float64) +
             got["t"][h, None].astype(np.float64) * dirs.cpu().numpy()[h, :3] + 1.0) / 2.0
        u, v = uv_of(p)
        assert np.max(np.abs(got["coords"][h, 0] - u)) < 2e-5 and np.max(np.abs(got["coords"][h, 1] - v)) < 2e-5
    if mode != 1:
        _check(torch_dev, g, orc, s, pos, dirs, _preset(g, mode, 10))


def test_textured_instances_mixed_persistent_any(lrt, orc, torch_dev, tex_nodes):
    g = lrt
    torch = torch_dev
    _say("textured frame octree instances beside a frame octree and a mesh")
    s = g.SceneRT()
    a = s.AddGeom_SdfFrameOctreeTex(tex_nodes)
    b = s.AddGeom_SdfFrameOctree(g.sdf_spheres_frame_octree(SPHERES, 5))
    verts, tris = g.gen_icosphere(3, 0.05, 4.0, 3)
    v = np.array(verts, np.float32)
    v[:, :3] *= 0.3
    m = s.AddGeom_Triangles3f(v, tris)
    for k, (geom, tx, ty, sc) in enumerate([(a, -0.8, 0.0, 0.6), (b, 0.7, 0.2, 0.5), (a, 0.1, 0.8, 0.4)]):
        t = np.eye(4, dtype=np.float32)
        ang = 0.5 * k
        t[:3, :3] = sc * np.array([[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]], np.float32)
        t[:3, 3] = [tx, ty, -0.2]
        s.AddInstance(geom, t)
    s.AddInstance(m, np.eye(4, dtype=np.float32))
    s.CommitScene()
    pos, dirs = _rays(torch, g, eye=(0.2, 0.6, 3.2))
    got = _check(torch, g, orc, s, pos, dirs, _preset(g, 0, 1))
    assert len(set((got["geomId"][got["primId"] != MISS] >> 24).tolist())) >= 3
    s.set_persistent(1)
    assert raw_equal(_check(torch, g, orc, s, pos, dirs, _preset(g, 0, 1)), got)
    s.set_persistent(0)
    flags = torch.zeros(pos.shape[0], dtype=torch.int32, device="cuda")
    s.trace_any(pos, dirs, flags)
    torch.cuda.synchronize()
    want = orc.trace_any(orc.Scene.from_scene(s), pos.cpu().numpy(), dirs.cpu().numpy())
    assert np.array_equal(flags.cpu().numpy().astype(np.uint32), want)
