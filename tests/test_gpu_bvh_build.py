"""GPU BVH construction and TLAS refit.

BLAS: the linear-BVH builder (build names "gpu..."; bvh_kernels.hip) must emit the reference's
layout contract (AlignNodes2, utils/raytracing/cbvh.cpp:143-169; CreateFatTreeArray,
cbvh_fat.cpp:128-142; leaf words cbvh.cpp:502-511, 645-655) with tight boxes, and scenes built
with it must trace exactly as the oracle on the same arrays and, closest hits being independent
of the tree, as scenes built by the host builder (t bit for bit; ids except exact ties).
TLAS: a commit after UpdateInstance only (BVH2Common_host.cpp:1303-1335) refits the TLAS on the
GPU; its boxes must equal the oracle's refit bit for bit, its topology must be unchanged, and
traces must equal the oracle's and a from-scratch build's.
"""
import sys

import numpy as np
import pytest

from conftest import MISS, raw_equal
from test_builder_cpu import LEAF_BIT, subtree_boxes_contain, walk_fat

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


def _tight_triangle_contract(g, s, f, pil):
    nodes = s.get_array("nodes")
    pic = s.get_array("prim_id_count")
    prim = s.get_array("prim_indices")
    idx = s.get_array("indices").reshape(-1, 3)
    assert np.array_equal(np.sort(prim), np.arange(f.shape[0])), "every triangle once in m_primIndices"
    assert np.array_equal(idx, f[prim])
    covered = np.zeros(f.shape[0], np.int32)
    seen = []

    def on_leaf(p, side, lo):
        assert (lo >> 28) & 7 == 1
        seen.append(lo & 0x0FFFFFFF)
        w = int(pic[lo & 0x0FFFFFFF])
        start, count = w & 0x0FFFFFFF, (w >> 28) & 7
        assert 1 <= count <= max(pil, 1)
        covered[start:start + count] += 1

    walk_fat(nodes, on_leaf)
    assert sorted(seen) == list(range(len(pic)))
    assert np.all(covered == 1)
    verts = s.get_array("vert_pos").reshape(-1, 4)[:, :3]

    def leaf_box(lo):
        w = int(pic[lo & 0x0FFFFFFF])
        start, count = w & 0x0FFFFFFF, (w >> 28) & 7
        pts = verts[idx[start:start + count].reshape(-1)]
        return pts.min(axis=0), pts.max(axis=0)

    mn, mx = subtree_boxes_contain(nodes, leaf_box)
    # tight: every stored box is exactly its subtree's bounds
    def rec(p):
        pair = nodes[p]
        for side in ("left", "right"):
            lo = int(pair[side]["leftOffset"])
            bmn, bmx = leaf_box(lo) if lo & LEAF_BIT else rec(lo)
            assert np.array_equal(pair[side]["boxMin"], bmn) and np.array_equal(pair[side]["boxMax"], bmx)
        return (np.minimum(pair["left"]["boxMin"], pair["right"]["boxMin"]),
                np.maximum(pair["left"]["boxMax"], pair["right"]["boxMax"]))
    sys.setrecursionlimit(10000)
    rec(0)
    return nodes


@pytest.mark.parametrize("subdiv,amp,pil", [(3, 0.0, 2), (5, 0.05, 2), (4, 0.05, 1), (4, 0.05, 4)])
def test_gpu_blas_contract(lrt, torch_dev, subdiv, amp, pil):
    g = lrt
    _say(f"gpu blas contract subdiv={subdiv} pil={pil}")
    v, f = g.gen_icosphere(subdiv, amp, 4.0, 1)
    s = g.SceneRT(build_name=f"gpu_lbvh{pil}")
    s.AddInstance(s.AddGeom_Triangles3f(v, f))
    s.CommitScene()
    info = s.commit_info()
    assert info["gpu_builder"] == 1 and info["kind"] == g.COMMIT_FULL
    assert info["blas_leaves"] == len(s.get_array("prim_id_count"))
    if pil == 1:
        assert info["blas_merged_leaves"] == 0
    _tight_triangle_contract(g, s, f, pil)


def test_gpu_box_blas_contract(lrt, torch_dev):
    """Box BLAS (frame-octree leaf boxes): one box per leaf, every box referenced once, stored
    boxes exactly the leaf boxes."""
    g = lrt
    _say("gpu box blas contract")
    sph = np.array([[0.1, -0.2, 0.05, 0.5], [-0.4, 0.3, 0.2, 0.3]], np.float32)
    s = g.SceneRT(build_name="gpu_lbvh2")
    s.AddInstance(s.AddGeom_SdfFrameOctree(g.sdf_spheres_frame_octree(sph, 6)))
    s.CommitScene()
    orig = s.get_array("orig_nodes")
    nodes = s.get_array("nodes")
    seen = []

    def on_leaf(p, side, lo):
        assert (lo >> 28) & 7 == 1
        k = lo & 0x0FFFFFFF
        seen.append(k)
        b = nodes[p][side]
        assert np.array_equal(b["boxMin"], orig["boxMin"][k]) and np.array_equal(b["boxMax"], orig["boxMax"][k])

    walk_fat(nodes, on_leaf)
    assert sorted(seen) == list(range(orig.size))


def _camera_rays(torch, g, res, eye=(0.4, 0.6, 3.0)):
    pi, vi = g.camera_inverse(list(eye), [0, 0, 0], [0, 1, 0], 60.0, 1.0, 0.01, 100.0)
    pos = torch.zeros((res * res, 4), dtype=torch.float32, device="cuda")
    dirs = torch.zeros_like(pos)
    g.eye_rays(pi, vi, res, res, pos, dirs, order=1)
    torch.cuda.synchronize()
    return pos, dirs


def _trace(torch, g, s, pos, dirs):
    hits = torch.zeros((pos.shape[0], 8), dtype=torch.int32, device="cuda")
    s.trace_closest(pos, dirs, hits)
    torch.cuda.synchronize()
    return hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)


def _same_closest(a, b):
    """Closest hits of two trees over the same primitives: t bit for bit everywhere; the ids may
    differ only where two primitives tie at the same t (shared edges and vertices)."""
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
    diff = (a["primId"] != b["primId"]) | (a["instId"] != b["instId"])
    assert diff.mean() <= 2e-3, f"{int(diff.sum())} id differences"
    return int(diff.sum())


def test_gpu_built_mesh_traces_like_host_built(lrt, orc, torch_dev):
    """A 81,920-triangle displaced icosphere: GPU-built BVH vs the oracle on its own arrays (every
    field bit for bit) and vs the host-built BVH (t bit for bit)."""
    g = lrt
    torch = torch_dev
    _say("gpu-built mesh trace parity")
    v, f = g.gen_icosphere(6, 0.05, 4.0, 1)
    sg = g.SceneRT(build_name="gpu_lbvh2")
    sg.AddInstance(sg.AddGeom_Triangles3f(v, f))
    sg.CommitScene()
    sh = g.SceneRT()
    sh.AddInstance(sh.AddGeom_Triangles3f(v, f))
    sh.CommitScene()
    pos, dirs = _camera_rays(torch, g, 256)
    hg = _trace(torch, g, sg, pos, dirs)
    hh = _trace(torch, g, sh, pos, dirs)
    want, ov = orc.trace_closest(orc.Scene.from_scene(sg), pos.cpu().numpy(), dirs.cpu().numpy())
    assert ov == 0
    assert raw_equal(hg, want)
    assert (hg["primId"] != MISS).mean() > 0.2
    _say(f"id differences at ties: {_same_closest(hg, hh)}")


def _two_level(g, build_name, meshes, mats):
    s = g.SceneRT(build_name=build_name)
    ids = [s.AddGeom_Triangles3f(v, f) for v, f in meshes]
    for i, m in enumerate(mats):
        s.AddInstance(ids[i % len(ids)], m)
    s.CommitScene()
    return s


def _mats(rng, n, spread=40.0):
    out = []
    for _ in range(n):
        m = np.eye(4, dtype=np.float32)
        q = rng.normal(size=4)
        a, b, c, d = q / np.linalg.norm(q)
        m[:3, :3] = np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)], [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                              [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]]) * rng.uniform(0.5, 2.0)
        m[:3, 3] = rng.uniform(0, spread, 3)
        out.append(m)
    return out


def _rays(rng, n, hi=40.0):
    org = rng.uniform(0, hi, (n, 3))
    dv = rng.normal(size=(n, 3))
    dv /= np.linalg.norm(dv, axis=1, keepdims=True)
    return (np.concatenate([org, np.full((n, 1), 1e-9)], 1).astype(np.float32),
            np.concatenate([dv, np.full((n, 1), 1e9)], 1).astype(np.float32))


def test_gpu_built_two_level_scene(lrt, orc, torch_dev):
    """Five geodesic meshes (GPU-built) under 300 instances: bit for bit with the oracle on the
    same arrays, t bit for bit with the host-built scene."""
    g = lrt
    torch = torch_dev
    _say("gpu-built two-level scene")
    rng = np.random.default_rng(21)
    meshes = [g.gen_geodesic(n, 0.05, 4.0, 200 + n) for n in (5, 9, 17, 33, 60)]
    mats = _mats(rng, 300)
    sg = _two_level(g, "gpu_lbvh2", meshes, mats)
    sh = _two_level(g, "cbvh_embree2", meshes, mats)
    p, d = _rays(rng, 1 << 16)
    pos, dirs = torch.from_numpy(p).cuda(), torch.from_numpy(d).cuda()
    hg, hh = _trace(torch, g, sg, pos, dirs), _trace(torch, g, sh, pos, dirs)
    want, ov = orc.trace_closest(orc.Scene.from_scene(sg), p, d)
    assert ov == 0 and raw_equal(hg, want)
    assert (hg["primId"] != MISS).mean() > 0.2
    _say(f"id differences at ties: {_same_closest(hg, hh)}")


def test_tlas_refit_after_update_instance(lrt, orc, torch_dev):
    """UpdateInstance on 100 of 400 instances, three rounds: each commit refits on the GPU
    (COMMIT_REFIT); the TLAS keeps its topology and its boxes equal the oracle's refit of the
    previous TLAS bit for bit; traces equal the oracle's on the same arrays and (t) a scene built
    from scratch with the final transforms. A large move triggers the host rebuild, as does
    tlas_update 0."""
    g = lrt
    torch = torch_dev
    _say("tlas refit")
    rng = np.random.default_rng(33)
    meshes = [g.gen_geodesic(n, 0.05, 4.0, 300 + n) for n in (6, 11, 20)]
    mats = _mats(rng, 400)
    s = _two_level(g, "cbvh_embree2", meshes, mats)
    p, d = _rays(rng, 1 << 15)
    pos, dirs = torch.from_numpy(p).cuda(), torch.from_numpy(d).cuda()
    for rnd in range(3):
        tlas0 = s.get_array("tlas")
        for i in rng.choice(400, 100, replace=False):
            m = mats[i].copy()
            m[:3, 3] += rng.uniform(-1.5, 1.5, 3).astype(np.float32)
            mats[i] = m
            s.UpdateInstance(int(i), m)
        s.CommitScene()
        info = s.commit_info()
        assert info["kind"] == g.COMMIT_REFIT, info
        tlas = s.get_array("tlas")
        assert np.array_equal(tlas["leftOffset"], tlas0["leftOffset"])
        assert np.array_equal(tlas["escapeIndex"], tlas0["escapeIndex"])
        ref = orc.tlas_refit(tlas0, s.get_array("instances"))
        assert tlas.tobytes() == ref.tobytes(), f"round {rnd}: refit boxes differ from the oracle"
        got = _trace(torch, g, s, pos, dirs)
        want, ov = orc.trace_closest(orc.Scene.from_scene(s), p, d)
        assert ov == 0 and raw_equal(got, want), f"round {rnd}"
        fresh = _two_level(g, "cbvh_embree2", meshes, mats)
        _same_closest(got, _trace(torch, g, fresh, pos, dirs))
        _say(f"round {rnd}: area ratio {info['tlas_area_ratio']:.3f}")
    # a large move: the refitted tree's area grows past twice the built one -> host rebuild
    for i in range(400):
        m = mats[i].copy()
        m[:3, 3] = rng.uniform(0, 400, 3).astype(np.float32)
        mats[i] = m
        s.UpdateInstance(i, m)
    s.CommitScene()
    assert s.commit_info()["kind"] == g.COMMIT_TLAS_REBUILD
    p2, d2 = _rays(rng, 1 << 15, 400.0)
    pos2, dirs2 = torch.from_numpy(p2).cuda(), torch.from_numpy(d2).cuda()
    got = _trace(torch, g, s, pos2, dirs2)
    want, _ = orc.trace_closest(orc.Scene.from_scene(s), p2, d2)
    assert raw_equal(got, want)
    # refit off: rebuild every time
    s.set_tlas_update(0)
    s.UpdateInstance(0, mats[1])
    s.CommitScene()
    assert s.commit_info()["kind"] == g.COMMIT_TLAS_REBUILD
    got = _trace(torch, g, s, pos2, dirs2)
    want, _ = orc.trace_closest(orc.Scene.from_scene(s), p2, d2)
    assert raw_equal(got, want)
    # adding an instance is a full commit
    s.AddInstance(0, mats[2])
    s.CommitScene()
    assert s.commit_info()["kind"] == g.COMMIT_FULL


def test_refit_single_instance(lrt, orc, torch_dev):
    """One instance: the TLAS is a single leaf; a refit moves its box with the instance."""
    g = lrt
    torch = torch_dev
    _say("single-instance refit")
    v, f = g.gen_icosphere(3)
    s = g.SceneRT()
    s.AddInstance(s.AddGeom_Triangles3f(v, f))
    s.CommitScene()
    m = np.eye(4, dtype=np.float32)
    m[:3, 3] = [0.5, -0.25, 0.0]
    s.UpdateInstance(0, m)
    s.CommitScene()
    assert s.commit_info()["kind"] == g.COMMIT_REFIT
    tlas = s.get_array("tlas")
    inst = s.get_array("instances")
    assert tlas.size == 1 and np.array_equal(tlas["boxMin"][0], inst["boxMin"][0][:3])
    pos, dirs = _camera_rays(torch, g, 64)
    got = _trace(torch, g, s, pos, dirs)
    want, _ = orc.trace_closest(orc.Scene.from_scene(s), pos.cpu().numpy(), dirs.cpu().numpy())
    assert raw_equal(got, want) and (got["primId"] != MISS).any()
