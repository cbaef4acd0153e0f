"""The global sparse octree built on the GPU (sdf_build_kernels.hip) against the host restatement
of sdf_to_global_octree (sdf/build/sparse_octree_builder.cpp:106-264): node records and values
byte for byte over depths, brick sizes and paddings, and the frame-octree, multi-octree, brick-set
and indexed brick-set generators that convert it, equal whichever side built the octree.
"""
import sys
import time

import numpy as np
import pytest

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


def _spheres(seed=2, n=64):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(-0.8, 0.8, (n, 3)), rng.uniform(0.05, 0.2, (n, 1))], 1).astype(np.float32)


@pytest.mark.parametrize("depth,bs,pad", [(3, 1, 0), (4, 2, 0), (7, 1, 0), (5, 4, 0), (6, 2, 1), (4, 8, 0), (9, 1, 0)])
def test_global_octree_equals_host(lrt, torch_dev, depth, bs, pad):
    g = lrt
    _say(f"global octree depth={depth} brick={bs} pad={pad}")
    sp = _spheres()
    t0 = time.perf_counter()
    hn, hv = g.sdf_spheres_global_octree(sp, depth, bs, pad, threads=0)
    t1 = time.perf_counter()
    gn, gv = g.sdf_spheres_global_octree(sp, depth, bs, pad, threads=g.GPU_BUILD)
    t2 = time.perf_counter()
    _say(f"  nodes {hn.shape[0]}, host {1e3 * (t1 - t0):.1f} ms, gpu {1e3 * (t2 - t1):.1f} ms")
    assert gn.shape == hn.shape and np.array_equal(gn, hn)
    assert gv.shape == hv.shape and np.array_equal(gv.view(np.uint32), hv.view(np.uint32))


def test_generators_equal_host(lrt, torch_dev):
    g = lrt
    _say("generators over the GPU octree")
    sp = _spheres(5, 40)
    assert np.array_equal(g.sdf_spheres_frame_octree(sp, 6, threads=g.GPU_BUILD).view(np.uint8),
                          g.sdf_spheres_frame_octree(sp, 6).view(np.uint8))
    a, b = g.sdf_spheres_multi_octree(sp, 6, threads=g.GPU_BUILD), g.sdf_spheres_multi_octree(sp, 6)
    assert np.array_equal(a[0].view(np.uint8), b[0].view(np.uint8)) and np.array_equal(a[1], b[1])
    hd = np.zeros((), g.SBS_HEADER_DTYPE)
    hd["brick_size"], hd["brick_pad"], hd["bytes_per_value"], hd["aux_data"] = 4, 1, 2, g.SDF_SBS_NODE_LAYOUT_DX
    a, b = g.sdf_spheres_sbs(sp, 4, hd, threads=g.GPU_BUILD), g.sdf_spheres_sbs(sp, 4, hd)
    assert all(np.array_equal(x.view(np.uint8), y.view(np.uint8)) for x, y in zip(a, b))
    a = g.sdf_spheres_sbs_indexed(sp, 4, 8, 1, threads=g.GPU_BUILD)
    b = g.sdf_spheres_sbs_indexed(sp, 4, 8, 1)
    for x, y in zip(a[1:4], b[1:4]):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    assert a[4] == b[4]
