"""GPU parity on the reference's own SDF data files (tests/golden/, see test_sdf_io_cpu.py): the
8,489-node engine octree as a frame octree and as a multi-octree, the 49,487-brick DX set and the
49,487-voxel sparse voxel set,
each traced by the gfx950 kernels and by the oracle on the same arrays at the path's bar; plus the
frame-octree vs multi-octree cross-check on the GPU's own records.
"""
import os
import sys

import numpy as np
import pytest

from conftest import MISS, assert_sdf_parity, sdf_check

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


def _rays(torch, g, res, eye):
    pi, vi = g.camera_inverse(list(eye), [0, 0, 0], [0, 1, 0], 60.0, 1.0, 0.01, 100.0)
    pos = torch.zeros((res * res, 4), dtype=torch.float32, device="cuda")
    dirs = torch.zeros_like(pos)
    jit = torch.from_numpy(np.random.default_rng(19).random((res * res, 2), dtype=np.float32)).cuda()
    g.eye_rays(pi, vi, res, res, pos, dirs, jitter=jit, order=1)
    return pos, dirs


def _trace(torch, g, s, pos, dirs):
    hits = torch.zeros((pos.shape[0], 8), dtype=torch.int32, device="cuda")
    s.trace_closest(pos, dirs, hits)
    torch.cuda.synchronize()
    return hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)


def _scene(g, kind):
    s = g.SceneRT()
    if kind == "frame":
        s.AddInstance(s.AddGeom_SdfFrameOctree(g.load_sdf_frame_octree(os.path.join(GOLD, "engine_sdf_frame_octree_0.bin"))))
    elif kind == "multi":
        mn, mv = g.load_sdf_multi_octree(os.path.join(GOLD, "engine_sdf_multi_octree_0.bin"))
        s.AddInstance(s.AddGeom_SdfMultiOctree(mn, mv))
    elif kind == "svs":
        s.AddInstance(s.AddGeom_SdfSVS(g.load_sdf_svs(os.path.join(GOLD, "test_10_svs.bin"))))
    else:
        h, nodes, vals, vf = g.load_sdf_sbs(os.path.join(GOLD, "test_10_sbs.bin"))
        s.AddInstance(s.AddGeom_SdfSBS(h, nodes, vals, vf if vf.size else None))
    s.CommitScene()
    return s


@pytest.mark.parametrize("eye", [(0.4, 0.6, 3.0), (-2.5, 1.0, -1.5)])
@pytest.mark.parametrize("kind", ["frame", "multi", "sbs", "svs"])
def test_reference_file_parity(lrt, orc, torch_dev, kind, eye):
    g = lrt
    _say(f"reference data {kind} eye={eye}")
    s = _scene(g, kind)
    pos, dirs = _rays(torch_dev, g, 512, eye)
    got = _trace(torch_dev, g, s, pos, dirs)
    p, d = pos.cpu().numpy(), dirs.cpu().numpy()
    want, ov = orc.trace_closest(orc.Scene.from_scene(s), p, d)
    r = sdf_check(got, want, p, d)
    _say(str(r))
    assert r["hits"] > 0.05 * r["n"]
    assert_sdf_parity(r, got, want)


def test_reference_frame_vs_multi_on_gpu(lrt, torch_dev):
    """The GPU's frame-octree and multi-octree walks over the two encodings of one octree agree
    to the bar of conftest.format_cross_check (same leaf and hit/miss on every ray)."""
    g = lrt
    from conftest import format_cross_check
    _say("reference data frame vs multi")
    pos, dirs = _rays(torch_dev, g, 512, (0.4, 0.6, 3.0))
    fs = _scene(g, "frame")
    hf = _trace(torch_dev, g, fs, pos, dirs)
    hm = _trace(torch_dev, g, _scene(g, "multi"), pos, dirs)
    ob = fs.get_array("orig_nodes")
    cell = float((ob["boxMax"] - ob["boxMin"])[:, 0].min())
    r = format_cross_check(hf, hm, pos.cpu().numpy(), dirs.cpu().numpy(), cell)
    _say(str(r))
    assert r["hits"] > 0.1 * pos.shape[0]
