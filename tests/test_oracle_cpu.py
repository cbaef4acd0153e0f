"""Pins the CPU oracle (triangle path, rays, camera) to what the mathematics fixes.

No golden vectors exist in the reference for this path (it has no tests), so each oracle
function is checked against brute force, closed forms, invariants and special cases chosen so
that a plausible mistake fails one of them; a planted mistake proves the checks bite.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, host_scene

MISS = 0xFFFFFFFF


def mesh_scene(g, subdiv=4, amp=0.05, mats=None):
    v, f = g.gen_icosphere(subdiv, amp, 4.0, 1)

    def build(s):
        gid = s.AddGeom_Triangles3f(v, f)
        for m in (mats or [None]):
            s.AddInstance(gid, m)
    return v, f, host_scene(g, build)


def camera_rays(orc, w=96, h=96, eye=(0, 0, 3)):
    pi, vi = orc.camera_inverse(list(eye), [0, 0, 0], [0, 1, 0], 60.0, w / h, 0.01, 100.0)
    return orc.eye_rays(pi, vi, w, h, order=0)


def barycentric_residual(v, f, hits, pos, dir):
    """|o + t d - (A + v(B-A) + u(C-A))| in float64 for hit rays (coords[0]=u weighs C, coords[1]=v weighs B)."""
    h = hits["primId"] != MISS
    tri = f[hits["primId"][h]]
    A, B, C = (v[tri[:, k]].astype(np.float64) for k in range(3))
    u = hits["coords"][h, 0].astype(np.float64)[:, None]
    vv = hits["coords"][h, 1].astype(np.float64)[:, None]
    p_bary = A + vv * (B - A) + u * (C - A)
    p_ray = pos[h, :3].astype(np.float64) + hits["t"][h, None].astype(np.float64) * dir[h, :3]
    return np.linalg.norm(p_bary - p_ray, axis=1)


def test_bvh_walk_equals_brute_force(lrt, orc):
    v, f, s = mesh_scene(lrt, 4, 0.05)
    pos, dir = camera_rays(orc)
    hits, ov = orc.trace_closest(orc.Scene.from_scene(s), pos, dir)
    assert ov == 0
    bt, bid, bt2 = orc.brute_force_tris(v, f, pos, dir)
    tie = (bt2 - bt) <= 1e-5 * bt
    same = hits["primId"] == bid
    assert np.all(same | tie), f"{np.sum(~same & ~tie)} rays disagree with brute force outside the tie band"
    hit = bid != MISS
    assert hit.sum() > 1000
    assert np.array_equal(hits["t"][same & hit], bt[same & hit]), "same triangle must give the same t"
    assert np.all(hits["t"][~hit] == 1e9) and np.all(hits["geomId"][~hit] == MISS)


def test_hit_point_closed_form(lrt, orc):
    v, f, s = mesh_scene(lrt, 3, 0.05)
    pos, dir = camera_rays(orc)
    hits, _ = orc.trace_closest(orc.Scene.from_scene(s), pos, dir)
    res = barycentric_residual(v, f, hits, pos, dir)
    assert res.size > 500 and res.max() < 2e-5


def test_axial_ray_distance_closed_form(lrt, orc):
    """Undisplaced icosphere: the ray along -z from (0,0,3) hits the face crossing the z axis;
    t = 3 - z where z is the plane of that face at x=y=0 (computed in float64)."""
    v, f, s = mesh_scene(lrt, 3, 0.0)
    pos = np.array([[0.0, 0.0, 3.0, 1e-9]], np.float32)
    dir = np.array([[1e-4, 2e-4, -1.0, 1e9]], np.float32)
    dir[0, :3] /= np.linalg.norm(dir[0, :3])
    hits, _ = orc.trace_closest(orc.Scene.from_scene(s), pos, dir)
    assert hits[0]["primId"] != MISS
    A, B, C = (v[f[hits[0]["primId"], k]].astype(np.float64) for k in range(3))
    n = np.cross(B - A, C - A)
    o, d = pos[0, :3].astype(np.float64), dir[0, :3].astype(np.float64)
    t_exact = np.dot(n, A - o) / np.dot(n, d)
    assert abs(hits[0]["t"] - t_exact) <= 1e-6 * t_exact
    assert 3.0 - 1.0 < t_exact < 3.0 - 0.98   # face plane lies just inside the unit sphere


def test_instance_transform_invariance(lrt, orc):
    ang = 0.7
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = np.array([[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]], np.float32) * 1.5
    m[:3, 3] = [4.0, -2.0, 1.0]
    v, f, s_id = mesh_scene(lrt, 3, 0.05)
    _, _, s_tr = mesh_scene(lrt, 3, 0.05, mats=[m])
    pos, dir = camera_rays(orc, 64, 64)
    h_id, _ = orc.trace_closest(orc.Scene.from_scene(s_id), pos, dir)
    pos_w = pos.copy()
    dir_w = dir.copy()
    pos_w[:, :3] = (pos[:, :3] @ m[:3, :3].T) + m[:3, 3]
    dir_w[:, :3] = dir[:, :3] @ m[:3, :3].T          # not renormalised: t stays comparable
    h_tr, _ = orc.trace_closest(orc.Scene.from_scene(s_tr), pos_w, dir_w)
    both = (h_id["primId"] != MISS) & (h_tr["primId"] != MISS)
    assert both.sum() > 300
    assert np.mean(h_id["primId"][both] == h_tr["primId"][both]) > 0.995
    same = both & (h_id["primId"] == h_tr["primId"])
    assert np.allclose(h_tr["t"][same], h_id["t"][same], rtol=1e-4)
    assert np.array_equal(h_id["primId"] == MISS, h_tr["primId"] == MISS) or \
        np.mean((h_id["primId"] == MISS) == (h_tr["primId"] == MISS)) > 0.995


def test_stop_on_first_hit_and_any_hit(lrt, orc):
    v, f, s = mesh_scene(lrt, 3, 0.05)
    sc = orc.Scene.from_scene(s)
    pos, dir = camera_rays(orc, 48, 48)
    closest, _ = orc.trace_closest(sc, pos, dir)
    d_any = dir.copy()
    d_any[:, 3] *= -1.0
    first, _ = orc.trace_closest(sc, pos, d_any)
    assert np.array_equal(first["primId"] != MISS, closest["primId"] != MISS)
    assert np.all(first["t"][closest["primId"] != MISS] >= closest["t"][closest["primId"] != MISS])
    flags = orc.trace_any(sc, pos, dir)
    assert np.array_equal(flags == 1, closest["primId"] != MISS)


def test_miss_record_layout(lrt, orc):
    v, f, s = mesh_scene(lrt, 2, 0.0)
    pos = np.array([[0, 0, 3, 1e-9]], np.float32)
    dir = np.array([[0, 1, 0, 123.0]], np.float32)   # points away from the sphere
    hits, _ = orc.trace_closest(orc.Scene.from_scene(s), pos, dir)
    h = hits[0]
    assert h["t"] == np.float32(123.0) and h["primId"] == MISS and h["instId"] == MISS and h["geomId"] == MISS
    assert list(h["coords"]) == [1.0, 0.0, 0.0, 0.0]


def test_eye_rays_pinhole_closed_form(orc):
    w, h = 32, 24
    pi, vi = orc.camera_inverse([0, 0, 3], [0, 0, 0], [0, 1, 0], 60.0, w / h, 0.01, 100.0)
    pos, dir = orc.eye_rays(pi, vi, w, h, order=0)
    xs, ys = np.meshgrid(np.arange(w), np.arange(h))
    pw, ph = 8 * ((w + 7) // 8), 8 * ((h + 7) // 8)
    fx, fy = (xs + 0.5) / pw, (ys + 0.5) / ph
    th = np.tan(np.radians(30.0))
    d = np.stack([(2 * fx - 1) * th * (w / h), (1 - 2 * fy) * th, -np.ones_like(fx)], -1).reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    assert np.abs(dir[:, :3] - d).max() < 2e-6
    assert np.allclose(pos[:, :3], [0, 0, 3], atol=1e-6)
    assert np.all(pos[:, 3] == np.float32(1e-9)) and np.all(dir[:, 3] == np.float32(1e9))


def test_super_block_order_is_a_permutation(orc):
    w, h = 64, 40
    pi, vi = orc.camera_inverse([0, 0, 3], [0, 0, 0], [0, 1, 0], 60.0, w / h, 0.01, 100.0)
    p0, d0 = orc.eye_rays(pi, vi, w, h, order=0)
    p1, d1 = orc.eye_rays(pi, vi, w, h, order=1)
    key0 = {tuple(r) for r in d0.view(np.uint32)}
    key1 = {tuple(r) for r in d1.view(np.uint32)}
    assert key0 == key1 and len(key1) == w * h


def test_camera_inverse_closed_form(orc):
    pi, vi = orc.camera_inverse([1, 2, 3], [0, 0.5, 0], [0, 1, 0], 45.0, 1.5, 0.1, 50.0)
    P = pi.reshape(4, 4).T.astype(np.float64)
    proj = np.linalg.inv(P)
    ymax = 0.1 * np.tan(np.radians(22.5))
    assert abs(proj[1, 1] - 0.1 / ymax) < 1e-4 * proj[1, 1]
    assert abs(proj[0, 0] - 0.1 / (ymax * 1.5)) < 1e-4 * proj[0, 0]
    assert abs(proj[3, 2] + 1) < 1e-6 and abs(proj[2, 2] + 50.1 / 49.9) < 1e-5
    V = np.linalg.inv(vi.reshape(4, 4).T.astype(np.float64))
    eye = np.array([1, 2, 3, 1.0])
    assert np.allclose(V @ eye, [0, 0, 0, 1], atol=1e-5)
    m = np.random.default_rng(1).normal(size=16).astype(np.float32)
    inv = orc.inverse4x4(m)
    assert np.allclose(inv.reshape(4, 4).T.astype(np.float64) @ m.reshape(4, 4).T, np.eye(4), atol=1e-4)


def test_planted_mistake_is_caught(lrt, orc, tmp_path):
    """An oracle built with swapped barycentrics must fail the closed-form hit-point check."""
    import ctypes
    import oracle as oracle_mod
    so = tmp_path / "liboracle_planted.so"
    src = os.path.join(ROOT, "oracle", "oracle.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
                    "-DORACLE_PLANT_SWAP_UV", "-o", str(so), src], check=True)
    lib = ctypes.CDLL(str(so))
    lib.oracle_trace_closest.argtypes = [ctypes.POINTER(oracle_mod.OracleScene), ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p]
    lib.oracle_trace_closest.restype = ctypes.c_uint64
    v, f, s = mesh_scene(lrt, 3, 0.05)
    sc = orc.Scene.from_scene(s)
    pos, dir = camera_rays(orc, 48, 48)
    hits = np.zeros(pos.shape[0], orc.HIT_DTYPE)
    lib.oracle_trace_closest(ctypes.byref(sc.view), pos.ctypes.data, dir.ctypes.data, hits.ctypes.data,
                             pos.shape[0], 1, None)
    assert barycentric_residual(v, f, hits, pos, dir).max() > 1e-3
