"""The allocation-free trace entry point against lrt_trace_closest_dev, both through the C ABI.

The no-alloc call keeps its tail queue and stack spill in the library's fixed per-GPU array; its
hit records must equal the allocating call's bit for bit, on two calls in a row (the second
smaller, so stale queue entries or spill words of the first would show), with the cooperative
tail pass forced on and off.
"""
import numpy as np
import pytest

from conftest import raw_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    return torch


def _rays(torch, g, w, h, eye=(0, 0, 3)):
    pi, vi = g.camera_inverse(list(eye), [0, 0, 0], [0, 1, 0], 60.0, w / h, 0.01, 100.0)
    pos = torch.zeros((w * h, 4), dtype=torch.float32, device="cuda")
    dirs = torch.zeros_like(pos)
    g.eye_rays(pi, vi, w, h, pos, dirs, order=1)
    return pos, dirs


def _both(torch, g, scene, pos, dirs):
    n = pos.shape[0]
    ref = torch.zeros((n, 8), dtype=torch.int32, device="cuda")
    got = torch.full((n, 8), -7, dtype=torch.int32, device="cuda")
    scene.trace_closest(pos, dirs, ref)
    torch.cuda.synchronize()
    scene.trace_closest_noalloc(pos, dirs, got)
    torch.cuda.synchronize()
    a = ref.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)
    b = got.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)
    return a, b


@pytest.mark.parametrize("budget", [0, 8])
def test_noalloc_matches_allocating_call_mesh(lrt, torch_dev, budget):
    g, torch = lrt, torch_dev
    print(f"no-alloc mesh budget={budget}", flush=True)
    v, f = g.gen_icosphere(6, 0.05, 4.0, 1)
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_Triangles3f(v, f))
    s.CommitScene()
    s.set_tail_config(budget, 256)
    pos, dirs = _rays(torch, g, 256, 256)          # 65,536 rays = the maximum
    assert pos.shape[0] == g.NOALLOC_MAX_RAYS
    a, b = _both(torch, g, s, pos, dirs)
    assert (a["primId"] != g.MISS).sum() > 1000
    assert raw_equal(a, b), "first (maximum-size) call differs"
    # second call, smaller and from another camera position: nothing of the first may leak in
    pos2, dirs2 = _rays(torch, g, 64, 64, eye=(1.5, 0.5, 2.5))
    pos2, dirs2 = pos2[:4001].contiguous(), dirs2[:4001].contiguous()
    a2, b2 = _both(torch, g, s, pos2, dirs2)
    assert (a2["primId"] != g.MISS).sum() > 100
    assert raw_equal(a2, b2), "second (smaller) call differs"


def test_noalloc_matches_allocating_call_sdf(lrt, torch_dev):
    g, torch = lrt, torch_dev
    print("no-alloc frame octree", flush=True)
    rng = np.random.default_rng(2)
    sph = np.concatenate([rng.uniform(-0.8, 0.8, (16, 3)), rng.uniform(0.1, 0.3, (16, 1))], 1).astype(np.float32)
    nodes = g.sdf_spheres_frame_octree(sph, 6)
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_SdfFrameOctree(nodes))
    s.CommitScene()
    s.set_tail_config(4, 128)
    for w, h in ((256, 256), (48, 40)):
        pos, dirs = _rays(torch, g, w, h)
        a, b = _both(torch, g, s, pos, dirs)
        assert (a["primId"] != g.MISS).sum() > 50
        assert raw_equal(a, b), f"{w}x{h} call differs"


def test_noalloc_rejects_too_many_rays(lrt, torch_dev):
    g, torch = lrt, torch_dev
    v, f = g.gen_icosphere(2, 0.0, 4.0, 1)
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_Triangles3f(v, f))
    s.CommitScene()
    n = g.NOALLOC_MAX_RAYS + 1
    pos = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    dirs = torch.zeros_like(pos)
    hits = torch.zeros((n, 8), dtype=torch.int32, device="cuda")
    with pytest.raises(g.LrtError) as e:
        s.trace_closest_noalloc(pos, dirs, hits)
    assert e.value.code == 1 and "at most" in str(e.value)
    # an empty call is accepted
    s.trace_closest_noalloc(pos[:0], dirs[:0], hits[:0])
