"""The drop-in boundary on the GPU: a compiled C caller of include/lrt.h, and the ISceneObject
calls that change a committed scene (UpdateInstance, ClearScene, ClearGeom, re-commit), each
checked against the oracle on the scene's own arrays.

Reference: ISceneObject (CrossRT/CrossRT.h:43-175); BVHRT::UpdateInstance, ClearScene, ClearGeom
(MultiRenderer/BVH2Common_host.cpp:1303-1335, 1230-1262).
"""
import ctypes
import os
import sys

import numpy as np
import pytest

from conftest import MISS, raw_equal

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    return torch


def _oracle_check(orc, s, pos, dirs, got, what):
    want, ov = orc.trace_closest(orc.Scene.from_scene(s), pos, dirs)
    assert ov == 0
    assert raw_equal(got, want), f"{what}: {int(np.sum(got['primId'] != want['primId']))} id mismatches"
    return want


def test_c_caller_of_the_header(lrt, orc, torch_dev):
    """tests/c_caller/lrt_c_caller.c, built with gcc -std=c11 against include/lrt.h: create, add
    triangles, instance, commit, query one ray, trace a batch, update the instance, re-commit,
    trace again, clear, destroy."""
    g = lrt
    _say("C caller")
    path = os.path.join(ROOT, "tests", "c_caller", "liblrt_c_caller.so")
    assert os.path.exists(path), "build it with __graft_entry__.build()"
    C = ctypes.CDLL(path)
    C.lrt_c_caller_run.argtypes = [ctypes.c_uint32] + [ctypes.c_void_p] * 3
    C.lrt_c_caller_run.restype = ctypes.c_int
    k = 16
    before = np.zeros(k * k, g.HIT_DTYPE)
    after = np.zeros(k * k, g.HIT_DTYPE)
    one = np.zeros(1, g.HIT_DTYPE)
    st = C.lrt_c_caller_run(k, before.ctypes.data, after.ctypes.data, one.ctypes.data)
    assert st == 0, g.lib().lrt_last_error().decode()
    # the same scene through the Python mirror, and the oracle on its arrays
    verts = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    tris = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    f = np.linspace(-0.9, 0.9, k).astype(np.float32)
    xs = np.float32(-0.9) + np.float32(1.8) * (np.arange(k * k) % k).astype(np.float32) / np.float32(k - 1)
    ys = np.float32(-0.9) + np.float32(1.8) * (np.arange(k * k) // k).astype(np.float32) / np.float32(k - 1)
    assert np.allclose(xs[:k], f)
    pos = np.stack([xs, ys, np.zeros(k * k, np.float32), np.full(k * k, 1e-9, np.float32)], 1)
    dirs = np.tile(np.array([[0, 0, -1, 1e9]], np.float32), (k * k, 1))
    for z, rec in ((-2.0, before), (-3.0, after)):
        m = np.eye(4, dtype=np.float32)
        m[2, 3] = z
        s = g.SceneRT()
        s.AddInstance(s.AddGeom_Triangles3f(verts, tris), m)
        s.CommitScene()
        assert raw_equal(s.trace_closest_host(pos, dirs), rec), f"C caller differs from the Python mirror at z={z}"
        _oracle_check(orc, s, pos, dirs, rec, f"z={z}")
        assert np.all(rec["t"] == np.float32(-z)) and np.all(rec["primId"] != MISS)
    assert raw_equal(one[0], before[0])


def _mats(rng, n):
    out = []
    for _ in range(n):
        m = np.eye(4, dtype=np.float32)
        q = rng.normal(size=4)
        a, b, c, d = q / np.linalg.norm(q)
        m[:3, :3] = np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)], [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                              [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]]) * rng.uniform(0.5, 2.0)
        m[:3, 3] = rng.uniform(0, 20, 3)
        out.append(m)
    return out


def _random_rays(rng, n, lo=0.0, hi=20.0):
    org = rng.uniform(lo, hi, (n, 3))
    dv = rng.normal(size=(n, 3))
    dv /= np.linalg.norm(dv, axis=1, keepdims=True)
    return (np.concatenate([org, np.full((n, 1), 1e-9)], 1).astype(np.float32),
            np.concatenate([dv, np.full((n, 1), 1e9)], 1).astype(np.float32))


def test_update_instance_and_recommit(lrt, orc, torch_dev):
    """UpdateInstance on 6 of 24 instances, re-commit, trace: records equal the oracle's on the
    re-committed arrays; t and ids equal a scene built from scratch with the new matrices (its
    normals may differ: UpdateInstance keeps the old transformInvTransposed, as the reference)."""
    g = lrt
    _say("update instance")
    rng = np.random.default_rng(31)
    meshes = [g.gen_geodesic(4 + 5 * k, 0.05, 4.0, 40 + k) for k in range(3)]
    mats = _mats(rng, 24)
    s = g.SceneRT()
    gids = [s.AddGeom_Triangles3f(v, f) for v, f in meshes]
    for i, m in enumerate(mats):
        s.AddInstance(gids[i % 3], m)
    s.CommitScene()
    pos, dirs = _random_rays(rng, 60000)
    first = s.trace_closest_host(pos, dirs)
    _oracle_check(orc, s, pos, dirs, first, "initial commit")
    new = _mats(rng, 6)
    moved = [1, 4, 9, 13, 20, 23]
    for i, m in zip(moved, new):
        s.UpdateInstance(i, m)
        mats[i] = m
    with pytest.raises(g.LrtError):       # updated but not committed: the device copy is stale
        s.trace_closest_host(pos, dirs)
    s.CommitScene()
    got = s.trace_closest_host(pos, dirs)
    _oracle_check(orc, s, pos, dirs, got, "after UpdateInstance")
    assert np.sum(got["primId"] != first["primId"]) > 100, "the moved instances changed nothing"
    fresh = g.SceneRT()
    fg = [fresh.AddGeom_Triangles3f(v, f) for v, f in meshes]
    for i, m in enumerate(mats):
        fresh.AddInstance(fg[i % 3], m)
    fresh.CommitScene()
    ref = fresh.trace_closest_host(pos, dirs)
    for k in ("t", "primId", "instId", "geomId"):
        assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), k
    assert np.array_equal(got["coords"][:, :2].view(np.uint32), ref["coords"][:, :2].view(np.uint32))
    # a device-side trace after the re-commit agrees with the host-buffer one
    torch = torch_dev
    p, d = torch.from_numpy(pos).cuda(), torch.from_numpy(dirs).cuda()
    h = torch.zeros((pos.shape[0], 8), dtype=torch.int32, device="cuda")
    s.trace_closest(p, d, h)
    torch.cuda.synchronize()
    assert raw_equal(h.cpu().numpy().view(g.HIT_DTYPE).reshape(-1), got)


def test_clear_scene_and_clear_geom(lrt, orc, torch_dev):
    """ClearScene keeps the geometries and drops the instances; ClearGeom drops both. Each is
    followed by new instances or geometries, a re-commit into the same device buffers and a
    trace against the oracle: triangles, then a single cached-BLAS mesh (its top pairs staged
    in LDS), then a brick set (occupancy masks rebuilt), then a mesh again."""
    g = lrt
    _say("clear scene / clear geom")
    rng = np.random.default_rng(37)
    s = g.SceneRT()
    v, f = g.gen_geodesic(12, 0.05, 4.0, 3)
    gid = s.AddGeom_Triangles3f(v, f)
    for m in _mats(rng, 8):
        s.AddInstance(gid, m)
    s.CommitScene()
    pos, dirs = _random_rays(rng, 40000)
    _oracle_check(orc, s, pos, dirs, s.trace_closest_host(pos, dirs), "8 instances")
    s.ClearScene()
    assert s.get_array("instances").size == 0 and s.get_array("geoms").size == 1
    for m in _mats(rng, 3):
        s.AddInstance(gid, m)
    s.CommitScene()
    got = s.trace_closest_host(pos, dirs)
    _oracle_check(orc, s, pos, dirs, got, "after ClearScene")
    assert (got["primId"] != MISS).sum() > 100
    # ClearGeom: a single mesh at the origin (one instance: the cached-BLAS launch)
    s.ClearGeom()
    assert s.get_array("geoms").size == 0 and s.get_array("instances").size == 0
    v2, f2 = g.gen_icosphere(5, 0.05, 4.0, 9)
    s.AddInstance(s.AddGeom_Triangles3f(v2, f2))
    s.CommitScene()
    pi, vi = g.camera_inverse([0, 0, 3], [0, 0, 0], [0, 1, 0], 60.0, 1.0, 0.01, 100.0)
    opos, odir = orc.eye_rays(pi, vi, 128, 128)
    got = s.trace_closest_host(opos, odir)
    _oracle_check(orc, s, opos, odir, got, "after ClearGeom (mesh)")
    assert (got["primId"] != MISS).sum() > 1000
    # a brick set in the same scene object: masks rebuilt at the re-commit
    s.ClearGeom()
    rng2 = np.random.default_rng(2)
    sph = np.concatenate([rng2.uniform(-0.8, 0.8, (16, 3)), rng2.uniform(0.1, 0.3, (16, 1))], 1).astype(np.float32)
    hd, nodes, vals, vf, nd = g.sdf_spheres_sbs_indexed(sph, 3, 8, 1)
    s.AddInstance(s.AddGeom_SdfSBS(hd, nodes, vals, vf))
    s.CommitScene()
    got = s.trace_closest_host(opos, odir)
    want, _ = orc.trace_closest(orc.Scene.from_scene(s), opos, odir)
    from conftest import assert_sdf_parity, sdf_check
    r = sdf_check(got, want, opos, odir, colour=True)
    assert r["hits"] > 1000
    assert_sdf_parity(r, got, want)
    mask = s.get_array("sbs_mask")
    wpb = 1 + (8 ** 3 + 31) // 32
    assert mask.size == nodes.size * wpb and mask[0::wpb].any()
    # and back to a mesh
    s.ClearGeom()
    s.AddInstance(s.AddGeom_Triangles3f(v, f))
    s.CommitScene()
    _oracle_check(orc, s, opos, odir, s.trace_closest_host(opos, odir), "mesh after brick set")
