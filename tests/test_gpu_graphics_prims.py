"""GPU parity of graphics-primitive leaves (IntersectGraphicPrims, BVH2Common.cpp:1225-1440) with
the oracle on the same arrays: every primitive kind, with and without per-primitive colours, in
one scene of several instances (the mixed trace kernel), bit for bit.
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


def _prims(g, rng, kind, k):
    a = rng.uniform(-0.8, 0.8, (k, 3))
    b = a + rng.normal(size=(k, 3)) * 0.35
    r = rng.uniform(0.02, 0.07, k)
    col = rng.uniform(0, 1, (k, 3))
    if kind in (g.GRAPH_PRIM_POINT, g.GRAPH_PRIM_POINT_COLOR):
        per = 1 if kind == g.GRAPH_PRIM_POINT else 2
        pts = np.zeros((per * k, 4), np.float32)
        pts[0::per, :3], pts[0::per, 3] = a, r * 2
        if per == 2:
            pts[1::2, :3] = col
        return pts
    per = 2 if kind < g.GRAPH_PRIM_LINE_COLOR else 3
    pts = np.zeros((per * k, 4), np.float32)
    pts[0::per, :3], pts[0::per, 3] = a, r
    pts[1::per, :3] = b
    if per == 3:
        pts[2::3, :3] = col
    return pts


def test_graphics_prims_parity(lrt, orc, torch_dev):
    g = lrt
    torch = torch_dev
    _say("graphics primitives")
    rng = np.random.default_rng(7)
    s = g.SceneRT()
    kinds = list(range(10))
    ids = [s.AddGeom_GraphicsPrim(k, _prims(g, rng, k, 6), color=(0.1 * k, 0.5, 0.9)) for k in kinds]
    for j, gid in enumerate(ids):
        m = np.eye(4, dtype=np.float32)
        m[:3, 3] = [2.2 * (j % 5) - 4.4, 2.2 * (j // 5) - 1.1, 0.0]
        s.AddInstance(gid, m)
    s.CommitScene()
    p = g.default_preset()
    p["render_mode"] = 1
    s.SetPreset(p)
    res = 256
    pi, vi = g.camera_inverse([0.0, 0.0, 7.0], [0, 0, 0], [0, 1, 0], 70.0, 1.0, 0.01, 100.0)
    pos = torch.zeros((res * res, 4), dtype=torch.float32, device="cuda")
    dirs = torch.zeros_like(pos)
    g.eye_rays(pi, vi, res, res, pos, dirs, order=1)
    hits = torch.zeros((res * res, 8), dtype=torch.int32, device="cuda")
    s.trace_closest(pos, dirs, hits)
    torch.cuda.synchronize()
    got = hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)
    want, ov = orc.trace_closest(orc.Scene.from_scene(s), pos.cpu().numpy(), dirs.cpu().numpy())
    assert ov == 0
    assert raw_equal(got, want), f"{int(np.sum(np.any(got.view(np.uint8).reshape(-1, 32) != want.view(np.uint8).reshape(-1, 32), 1)))} rows differ"
    hit = got["primId"] != MISS
    assert set(got["instId"][hit].tolist()) == set(range(10)), "every primitive kind is hit"
    assert np.all((got["geomId"][hit] >> 24) == 17)
