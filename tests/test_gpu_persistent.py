"""The persistent trace kernel (one launch of resident blocks, per-lane ray refill) against the
grid launch, record for record: single- and multi-instance meshes, stop-on-first-hit and any-hit,
the SDF leaf kinds, a sample of the 4,096-instance scene, ragged batch sizes, back-to-back calls
(its ray counter must be left at zero), graph replay and the allocation-free entry point.
"""
import sys

import numpy as np
import pytest

from conftest import raw_equal

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


def _trace(torch, g, s, pos, dirs, fn="trace_closest"):
    hits = torch.full((pos.shape[0], 8), -5, dtype=torch.int32, device="cuda")
    getattr(s, fn)(pos, dirs, hits)
    torch.cuda.synchronize()
    return hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)


def _eye(torch, g, w, h):
    pi, vi = g.camera_inverse([0.3, 0.2, 3.0], [0, 0, 0], [0, 1, 0], 60.0, w / h, 0.01, 100.0)
    pos = torch.zeros((w * h, 4), dtype=torch.float32, device="cuda")
    dirs = torch.zeros_like(pos)
    g.eye_rays(pi, vi, w, h, pos, dirs, order=1)
    return pos, dirs


def _scenes(g, torch):
    import bench
    out = []
    v, f = g.gen_icosphere(6, 0.05, 4.0, 1)
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_Triangles3f(v, f))
    s.CommitScene()
    pos, dirs = _eye(torch, g, 256, 256)
    out.append(("mesh", s, pos, dirs))
    d_first = dirs.clone()
    d_first[:, 3] *= -1
    out.append(("mesh first-hit", s, pos, d_first))
    rng = np.random.default_rng(2)
    sph = np.concatenate([rng.uniform(-0.8, 0.8, (32, 3)), rng.uniform(0.05, 0.2, (32, 1))], 1).astype(np.float32)
    fr = g.SceneRT()
    fr.AddInstance(fr.AddGeom_SdfFrameOctree(g.sdf_spheres_frame_octree(sph, 7)))
    fr.CommitScene()
    out.append(("frame octree", fr, pos, dirs))
    mn, mv = g.sdf_spheres_multi_octree(sph, 7)
    mo = g.SceneRT()
    mo.AddInstance(mo.AddGeom_SdfMultiOctree(mn, mv))
    mo.CommitScene()
    out.append(("multi-octree", mo, pos, dirs))
    h, nodes, vals, vf, nd = g.sdf_spheres_sbs_indexed(sph, 4, 8, 1)
    sb = g.SceneRT()
    sb.AddInstance(sb.AddGeom_SdfSBS(h, nodes, vals, vf))
    sb.CommitScene()
    out.append(("brick set", sb, pos, dirs))
    tl, meshes, mats = bench.tlas_scene()
    tp, td = bench.hemisphere_rays(meshes, mats, 1 << 17, seed=29)
    out.append(("tlas 4096x32", tl, torch.from_numpy(tp).cuda(), torch.from_numpy(td).cuda()))
    return out


@pytest.fixture(scope="module")
def scenes(lrt, torch_dev):
    return _scenes(lrt, torch_dev)


@pytest.mark.parametrize("block,lds", [(64, 8), (128, 16), (256, 32), (64, 32)])
def test_persistent_matches_grid_launch(lrt, torch_dev, scenes, block, lds):
    g, torch = lrt, torch_dev
    for name, s, pos, dirs in scenes:
        _say(f"persistent {name} block={block} lds={lds}")
        s.set_persistent(False)
        s.set_launch_config(block, lds, True)
        s.set_tail_config(0, 1024)
        base = _trace(torch, g, s, pos, dirs)
        flags0 = torch.zeros(pos.shape[0], dtype=torch.int32, device="cuda")
        s.trace_any(pos, dirs, flags0)
        s.set_persistent(True)
        for rep in range(2):   # the second call checks that the first left its counters at zero
            got = _trace(torch, g, s, pos, dirs)
            assert raw_equal(got, base), f"{name} call {rep}: {int(np.sum(got['primId'] != base['primId']))} ids"
        flags = torch.zeros(pos.shape[0], dtype=torch.int32, device="cuda")
        s.trace_any(pos, dirs, flags)
        torch.cuda.synchronize()
        assert torch.equal(flags, flags0), name
        s.set_persistent(False)


def test_persistent_ragged_sizes_and_noalloc(lrt, torch_dev, scenes):
    g, torch = lrt, torch_dev
    name, s, pos, dirs = scenes[0]
    s.set_launch_config(128, 16, True)
    for n in (1, 37, 64, 65, 1000, 64 * 97 + 5):
        _say(f"persistent ragged n={n}")
        p, d = pos[:n].contiguous(), dirs[:n].contiguous()
        s.set_persistent(False)
        base = _trace(torch, g, s, p, d)
        s.set_persistent(True)
        assert raw_equal(_trace(torch, g, s, p, d), base), f"n={n}"
        assert raw_equal(_trace(torch, g, s, p, d, "trace_closest_noalloc"), base), f"noalloc n={n}"
    s.set_persistent(False)


def test_persistent_graph_replay_and_autotune(lrt, torch_dev, scenes):
    """A captured persistent call replays with identical records; the tuner may choose it."""
    g, torch = lrt, torch_dev
    name, s, pos, dirs = scenes[-1]
    s.set_persistent(False)
    base = _trace(torch, g, s, pos, dirs)
    s.set_persistent(True)
    hits = torch.zeros((pos.shape[0], 8), dtype=torch.int32, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        s.trace_closest(pos, dirs, hits, stream=side.cuda_stream)
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=side):
            s.trace_closest(pos, dirs, hits, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for r in range(3):
        hits.fill_(-1)
        graph.replay()
        torch.cuda.synchronize()
        assert raw_equal(hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1), base), f"replay {r}"
    s.set_persistent(False)
    t = s.autotune(pos, dirs, hits)
    torch.cuda.synchronize()
    _say(f"autotune on the tlas sample: {t}")
    assert t["persistent"] in (0, 1)
    assert raw_equal(_trace(torch, g, s, pos, dirs), base)
