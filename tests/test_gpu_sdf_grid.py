"""GPU parity of dense SDF grid leaves (IntersectSDFGrid, BVH2Common.cpp:1114-1182, with
eval_distance_sdf_grid, :1442-1478) against the oracle on the same arrays: only IEEE
+,-,*,/ and sqrt are involved, so every record is bit-identical — a grid alone, two instances
(one rotated and scaled), a grid beside a triangle mesh (the mixed kernel), and the
ST_ITERATIONS and INTERSECTION_COUNT render modes.
"""
import sys

import numpy as np
import pytest

from conftest import MISS, raw_equal

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


def _union_grid(n=40, seed=2):
    rng = np.random.default_rng(seed)
    sp = np.concatenate([rng.uniform(-0.7, 0.7, (24, 3)), rng.uniform(0.08, 0.25, (24, 1))], 1)
    c = np.linspace(-1.0, 1.0, n)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    p = np.stack([x, y, z], -1).reshape(-1, 3)
    d = np.min(np.linalg.norm(p[:, None, :] - sp[None, :, :3], axis=2) - sp[None, :, 3], axis=1)
    return d.reshape(n, n, n).astype(np.float32)


def _rays(torch, g, res=160, eye=(0.3, 0.5, 3.0)):
    pi, vi = g.camera_inverse(list(eye), [0, 0, 0], [0, 1, 0], 60.0, 1.0, 0.01, 100.0)
    pos = torch.zeros((res * res, 4), dtype=torch.float32, device="cuda")
    dirs = torch.zeros_like(pos)
    jit = torch.from_numpy(np.random.default_rng(5).random((res * res, 2), dtype=np.float32)).cuda()
    g.eye_rays(pi, vi, res, res, pos, dirs, jitter=jit, order=1)
    torch.cuda.synchronize()
    return pos, dirs


def _check(torch, g, orc, s, pos, dirs, render_mode=1):
    p = g.default_preset()
    p["render_mode"] = render_mode
    s.SetPreset(p)
    hits = torch.zeros((pos.shape[0], 8), dtype=torch.int32, device="cuda")
    s.trace_closest(pos, dirs, hits)
    torch.cuda.synchronize()
    got = hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)
    want, ov = orc.trace_closest(orc.Scene.from_scene(s), pos.cpu().numpy(), dirs.cpu().numpy())
    assert ov == 0
    assert raw_equal(got, want), f"{int(np.sum(np.any(got.view(np.uint8).reshape(-1, 32) != want.view(np.uint8).reshape(-1, 32), 1)))} rows differ"
    return got


@pytest.mark.parametrize("render_mode", [1, 10, 26])
def test_grid_alone(lrt, orc, torch_dev, render_mode):
    g = lrt
    _say(f"sdf grid alone, render mode {render_mode}")
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SdfGrid(_union_grid()))
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g)
    got = _check(torch_dev, g, orc, s, pos, dirs, render_mode)
    assert (got["t"] < 1e9).mean() > 0.08


def test_grid_instances_and_mesh(lrt, orc, torch_dev):
    g = lrt
    _say("sdf grid instances beside a mesh")
    s = g.SceneRT()
    gid = s.AddGeom_SdfGrid(_union_grid(33, seed=4))
    v, f = g.gen_icosphere(4, 0.05, 4.0, 3)
    mid = s.AddGeom_Triangles3f(v * np.float32(0.4), f)
    s.AddInstance(gid)
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]], np.float32) * 0.6
    m[:3, 3] = [1.1, -0.4, 0.3]
    s.AddInstance(gid, m)
    t = np.eye(4, dtype=np.float32)
    t[:3, 3] = [-1.2, 0.6, 0.2]
    s.AddInstance(mid, t)
    s.CommitScene()
    pos, dirs = _rays(torch_dev, g, 192, eye=(0.0, 0.4, 4.0))
    got = _check(torch_dev, g, orc, s, pos, dirs)
    kinds = set((got["geomId"][got["primId"] != MISS] >> 24).tolist())
    assert kinds == {0, 1}, kinds   # triangles and grid hits both present
