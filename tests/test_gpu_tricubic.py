"""GPU parity of tricubic interpolation (BVH2Common.cpp:220-335, 790-844, 884-955, 1478-1569)
against the oracle on the same arrays: brick sets in the indexed fp32 and the quantised 8- and
16-bit layouts, every intersection mode and the iteration render mode; dense grids; a transformed
mix of brick sets, a grid and a mesh (the mixed kernel), the persistent kernel and any-hit; the
refusal of brick sets without padding. Records are bit-identical (the interpolant, its march and
its normal use no transcendental function)."""
import sys

import numpy as np
import pytest

from conftest import MISS, raw_equal
from test_coctree_cpu import SPHERES
from test_gpu_coctree import _check, _rays
from test_tricubic_cpu import C, R, quadratic_grid, quadratic_sbs

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


def _preset(g, mode=0, render=8, interp=1):
    p = g.default_preset()
    p["sdf_node_intersect"], p["render_mode"], p["interpolation_mode"] = mode, render, interp
    return p


def _quantised(g, bytes_per_value):
    h = np.zeros((), g.SBS_HEADER_DTYPE)
    h["brick_size"], h["brick_pad"], h["bytes_per_value"], h["aux_data"] = 4, 1, bytes_per_value, g.SDF_SBS_NODE_LAYOUT_DX
    n, v = g.sdf_spheres_sbs(SPHERES, 4, h)
    return h, n, v


@pytest.mark.parametrize("layout", ["indexed", "dx8", "dx16"])
@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_brick_set_tricubic(lrt, orc, torch_dev, layout, mode):
    g = lrt
    _say(f"tricubic brick set {layout}, intersection mode {mode}")
    arrays = quadratic_sbs(g) if layout == "indexed" else _quantised(g, 1 if layout == "dx8" else 2)
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SdfSBS(*arrays))
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g)
    got = _check(torch_dev, g, orc, s, pos, dirs, _preset(g, mode))
    if mode in (0, 3):
        assert (got["primId"] != MISS).mean() > 0.02
    _check(torch_dev, g, orc, s, pos, dirs, _preset(g, mode, 10))
    _check(torch_dev, g, orc, s, pos, dirs, _preset(g, mode, 1))
    p = _preset(g, mode)
    p["representation_mode"] = 0   # SURFACE: the start sample's sign flips the march and the normal
    _check(torch_dev, g, orc, s, pos, dirs, p)


def test_grid_tricubic(lrt, orc, torch_dev):
    g = lrt
    _say("tricubic dense grid")
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SdfGrid(quadratic_grid()))
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g)
    for render in (8, 10):
        got = _check(torch_dev, g, orc, s, pos, dirs, _preset(g, 0, render))
        assert (got["primId"] != MISS).mean() > 0.02


def test_tricubic_mixed_persistent_any(lrt, orc, torch_dev):
    g = lrt
    torch = torch_dev
    _say("tricubic brick sets, a grid and a mesh")
    s = g.SceneRT()
    a = s.AddGeom_SdfSBS(*quadratic_sbs(g))
    b = s.AddGeom_SdfSBS(*_quantised(g, 1))
    c = s.AddGeom_SdfGrid(quadratic_grid())
    verts, tris = g.gen_icosphere(3, 0.05, 4.0, 3)
    v = np.array(verts, np.float32)
    v[:, :3] *= 0.3
    m = s.AddGeom_Triangles3f(v, tris)
    for k, (geom, tx, ty, sc) in enumerate([(a, -0.8, 0.0, 0.6), (b, 0.7, 0.2, 0.5), (c, 0.1, 0.8, 0.4)]):
        t = np.eye(4, dtype=np.float32)
        ang = 0.5 * k
        t[:3, :3] = sc * np.array([[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]], np.float32)
        t[:3, 3] = [tx, ty, -0.2]
        s.AddInstance(geom, t)
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


def test_unpadded_brick_set_is_refused(lrt, torch_dev):
    """The 4^3 samples start one voxel before the brick: without padding the reference reads
    outside it, so the call fails with LRT_ERR_UNSUPPORTED; trilinear tracing of it runs."""
    g = lrt
    torch = torch_dev
    _say("tricubic without brick padding")
    h = np.zeros((), g.SBS_HEADER_DTYPE)
    h["brick_size"], h["brick_pad"], h["bytes_per_value"], h["aux_data"] = 4, 0, 1, g.SDF_SBS_NODE_LAYOUT_DX
    n, v = g.sdf_spheres_sbs(SPHERES, 4, h)
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SdfSBS(h, n, v))
    s.CommitScene()
    pos, dirs = _rays(torch, g, res=32)
    hits = torch.zeros((pos.shape[0], 8), dtype=torch.int32, device="cuda")
    s.SetPreset(_preset(g, 0, 8))
    with pytest.raises(g.LrtError) as e:
        s.trace_closest(pos, dirs, hits)
    assert e.value.code == 4
    s.SetPreset(_preset(g, 0, 8, interp=0))
    s.trace_closest(pos, dirs, hits)
    torch.cuda.synchronize()
    assert (hits.cpu().numpy().view(g.HIT_DTYPE)["primId"] != MISS).any()
