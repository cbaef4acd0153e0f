"""Pins the CPU oracle's SDF leaves (frame octree and sparse brick set) to what the
mathematics fixes.

* One voxel, every intersection mode: the hit must be the first root of the trilinear
  polynomial along the ray, found independently here in float64 by sampling and bisection
  (LocalSurfaceIntersection, BVH2Common.cpp:345-725). A planted mistake (the xyz term of the
  cubic dropped) must fail the check.
* A whole scene, the union of 64 spheres: hit points lie on the analytic surface within a
  fraction of a voxel, normals follow the analytic gradient, and rays disagree with the
  analytic hit/miss only when they graze a sphere (chord shorter than two voxels).
* Cross-format: the frame octree and the sparse brick set built from one octree give the
  same hits (bit-identical t for the fp32 indexed layout with 8^3 bricks), and the quantised
  layouts stay within their quantisation error.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, host_scene

MISS = 0xFFFFFFFF
EPS = 1e-6
CORNERS = np.array([[(i >> 2) & 1, (i >> 1) & 1, i & 1] for i in range(8)], np.float64)


def trilinear(v, q):
    x, y, z = q[..., 0], q[..., 1], q[..., 2]
    return ((1 - x) * (1 - y) * (1 - z) * v[0] + (1 - x) * (1 - y) * z * v[1] + (1 - x) * y * (1 - z) * v[2] +
            (1 - x) * y * z * v[3] + x * (1 - y) * (1 - z) * v[4] + x * (1 - y) * z * v[5] + x * y * (1 - z) * v[6] +
            x * y * z * v[7])


def voxel_cases(n, seed, d):
    """Corner values of a random sphere whose surface passes near the voxel (world units, voxel
    size d) and a ray entering through the x=0 face, mostly aimed at a surface point inside the
    voxel; returns (values, start_q, dir, q_far)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        r = rng.uniform(0.3, 2.0)
        u = rng.normal(size=3)
        c = 0.5 + u / np.linalg.norm(u) * (r + rng.uniform(-0.3, 0.8))
        vals = ((np.linalg.norm(CORNERS - c, axis=1) - r) * d).astype(np.float32)
        sq = np.array([0.0, rng.uniform(0, 1), rng.uniform(0, 1)], np.float32)
        q = rng.uniform(0, 1, 3)
        target = c + r * (q - c) / np.linalg.norm(q - c)
        if len(out) % 4 == 3 or not (np.all(target >= 0) and np.all(target <= 1) and target[0] > 0.05):
            target = np.array([rng.uniform(0.3, 1.0), rng.uniform(0, 1), rng.uniform(0, 1)])
        dv = (target - sq) / np.linalg.norm(target - sq)
        dv = dv.astype(np.float32)
        q_far = min((1 - sq[a]) / dv[a] if dv[a] > 0 else (0 - sq[a]) / dv[a] for a in range(3) if dv[a] != 0)
        out.append((vals, sq, dv, np.float32(q_far)))
    return out


def first_root(vals, sq, dv, q_far, sign):
    """First t in [0, q_far] with sign*f(t) <= 0 in float64 (EPS-early hit at t=0), else None."""
    v = vals.astype(np.float64)
    o, dd = sq.astype(np.float64), dv.astype(np.float64)
    ts = np.linspace(0.0, float(q_far), 4001)
    f = sign * trilinear(v, o[None] + ts[:, None] * dd[None])
    if f[0] <= EPS:
        return 0.0
    idx = np.nonzero(f <= 0)[0]
    if idx.size == 0:
        return None
    lo, hi = ts[idx[0] - 1], ts[idx[0]]
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if sign * trilinear(v, o + mid * dd) > 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def run_voxels(local_surface, lrt, mode, representation, cases, d):
    """(start value, float64 first root, oracle t) per case, t in voxel units (None = miss)."""
    p = lrt.default_preset()
    p["sdf_node_intersect"] = mode
    p["representation_mode"] = representation
    rows = []
    for vals, sq, dv, q_far in cases:
        h = local_surface(p, vals, d, 0.0, float(q_far), 0.0, sq, dv)
        got = float(h["t"]) / d if h["primId"] != MISS else None
        f0 = trilinear(vals.astype(np.float64), sq.astype(np.float64))
        sign = 1.0 if representation == 1 or f0 >= 0 else -1.0   # SDF mode flips rays that start inside
        rows.append((f0, first_root(vals, sq, dv, q_far, sign), got))
    return rows


# per mode: worst |t - root| (voxel units) allowed, and the bound 90% of hits must meet
BARS = {0: (0.02, 2e-3),    # sphere tracing stops once f <= 1e-6 (world), so it lands just short
        1: (1e-2, 1e-4),    # closed-form cubic in fp32 (ill-conditioned only near double roots)
        2: (1e-6, 1e-6),    # bracketed Newton
        4: (2e-3, 1e-3)}    # interval tracing


@pytest.mark.parametrize("representation", [0, 1])
@pytest.mark.parametrize("mode", [0, 1, 2, 4])
def test_local_surface_first_root(lrt, orc, mode, representation):
    """Rays entering from outside the surface: the hit is the first root of the trilinear."""
    d = 1.0 / 256
    cases = voxel_cases(1500, 11, d)
    rows = run_voxels(orc.local_surface, lrt, mode, representation, cases, d)
    outside = [(r, g, qf) for (f0, r, g), (_, _, _, qf) in zip(rows, cases) if f0 > 4 * EPS]
    assert sum(r is not None for r, _, _ in outside) > 150
    # marching steps may overshoot a root that lies within a step of the exit face (the
    # trilinear is not 1-Lipschitz), and the march then ends outside the voxel
    slack = 0.02 if mode in (0, 4) else 0.0
    assert all((r is None) == (g is None) or (g is None and r > qf - slack) for r, g, qf in outside), \
        "hit/miss differs from the float64 root"
    err = np.array([abs(r - g) for r, g, _ in outside if r is not None and g is not None])
    worst, p90 = BARS[mode]
    assert err.max() <= worst and np.quantile(err, 0.9) <= p90, (err.max(), np.quantile(err, 0.9))
    inside = [(r, g) for f0, r, g in rows if f0 < -4 * EPS]
    assert len(inside) > 100
    if representation == 1:
        # volume mode: a ray that starts inside is a hit at its entry point
        assert all(g == 0.0 for _, g in inside)
    elif mode in (1, 2):
        # SDF mode: the sign is flipped and the solvers find where the ray leaves the surface
        # (the marching modes may overshoot there: the flipped trilinear is not 1-Lipschitz)
        assert all((r is None) == (g is None) for r, g in inside)
        err = np.array([abs(r - g) for r, g in inside if r is not None])
        assert err.max() <= 3 * BARS[mode][0] and np.quantile(err, 0.9) <= 10 * BARS[mode][1]


def test_local_surface_bbox_mode_hits_at_entry(lrt, orc):
    d = 1.0 / 64
    p = lrt.default_preset()
    p["sdf_node_intersect"] = 3
    for vals, sq, dv, q_far in voxel_cases(50, 12, d):
        h = orc.local_surface(p, vals, d, 0.0, float(q_far), 0.25, sq, dv)
        assert h["primId"] != MISS and h["t"] == np.float32(0.25)


def test_planted_cubic_mistake_is_caught(lrt, tmp_path):
    """An oracle whose cubic drops the xyz term must fail the float64 root check."""
    so = tmp_path / "liboracle_planted_k7.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
                    "-DORACLE_PLANT_DROP_K7", "-o", str(so), os.path.join(ROOT, "oracle", "oracle.cpp")], check=True)
    lib = ctypes.CDLL(str(so))
    vp, f32 = ctypes.c_void_p, ctypes.c_float
    lib.oracle_local_surface.argtypes = [vp, vp, f32, f32, f32, f32, vp, vp, vp]
    lib.oracle_local_surface.restype = ctypes.c_int
    import oracle

    def planted(preset, values, d, qn, qf, fn, sq, dv):
        h = np.zeros((), oracle.HIT_DTYPE)
        v, s, r = (np.ascontiguousarray(a, dtype=np.float32) for a in (values, sq, dv))
        lib.oracle_local_surface(preset.ctypes.data, v.ctypes.data, d, qn, qf, fn, s.ctypes.data, r.ctypes.data,
                                 h.ctypes.data)
        return h

    d = 1.0 / 256
    rows = run_voxels(planted, lrt, 1, 1, voxel_cases(400, 13, d), d)
    bad = sum((r is None) != (g is None) or (r is not None and abs(r - g) > 1e-2) for f0, r, g in rows if f0 > 4 * EPS)
    assert bad > 5


# ---------------------------------------------------------------------------------------------
def sphere_set(seed=2, n=64):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(-0.8, 0.8, (n, 3)), rng.uniform(0.05, 0.2, (n, 1))], 1).astype(np.float32)


def union_distance(sp, p):
    dd = np.linalg.norm(p[:, None, :] - sp[None, :, :3].astype(np.float64), axis=2) - sp[None, :, 3]
    k = np.argmin(dd, axis=1)
    return dd[np.arange(len(p)), k], k


def decode_normal(e):
    """Inverse of encode_normal (octahedral, intersection_common.h:96-101)."""
    x, y = e[:, 0].astype(np.float64), e[:, 1].astype(np.float64)
    z = 1 - np.abs(x) - np.abs(y)
    neg = z < 0
    sx = np.where(x >= 0, 1.0, -1.0)
    sy = np.where(y >= 0, 1.0, -1.0)
    v = np.stack([np.where(neg, (1 - np.abs(y)) * sx, x), np.where(neg, (1 - np.abs(x)) * sy, y), z], 1)
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def analytic_first_hit(sp, pos, dir):
    """First ray/sphere entry in float64 and the chord length through that sphere."""
    o = pos[:, :3].astype(np.float64)
    dv = dir[:, :3].astype(np.float64)
    oc = o[:, None, :] - sp[None, :, :3]
    b = np.sum(oc * dv[:, None, :], 2)
    c = np.sum(oc * oc, 2) - sp[None, :, 3].astype(np.float64) ** 2
    disc = b * b - c
    t_in = np.where(disc > 0, -b - np.sqrt(np.maximum(disc, 0)), np.inf)
    t_in = np.where(t_in > 0, t_in, np.inf)
    k = np.argmin(t_in, axis=1)
    rows = np.arange(len(pos))
    return t_in[rows, k], 2 * np.sqrt(np.maximum(disc[rows, k], 0))


@pytest.fixture(scope="module")
def camera_batch(orc):
    w = 128
    pi, vi = orc.camera_inverse([0, 0, 3], [0, 0, 0], [0, 1, 0], 60.0, 1.0, 0.01, 100.0)
    jit = np.random.default_rng(3).random((w * w, 2), dtype=np.float32)
    return orc.eye_rays(pi, vi, w, w, jitter=jit, order=0)


def test_sphere_union_closed_form(lrt, orc, camera_batch):
    depth = 7
    vox = 2.0 / 2 ** depth
    sp = sphere_set()
    nodes = lrt.sdf_spheres_frame_octree(sp, depth)
    s = host_scene(lrt, lambda s: s.AddInstance(s.AddGeom_SdfFrameOctree(nodes)))
    pos, dir = camera_batch
    hits, ov = orc.trace_closest(orc.Scene.from_scene(s), pos, dir)
    assert ov == 0
    h = hits["primId"] != MISS
    p = pos[h, :3].astype(np.float64) + hits["t"][h, None].astype(np.float64) * dir[h, :3]
    f, k = union_distance(sp, p)
    assert np.abs(f).max() <= 0.35 * vox and np.quantile(np.abs(f), 0.99) <= 0.1 * vox
    grad = p - sp[k, :3]
    grad /= np.linalg.norm(grad, axis=1, keepdims=True)
    cos = np.sum(decode_normal(hits["coords"][h, 2:4]) * grad, axis=1)
    assert cos.min() > 0.5 and np.quantile(cos, 0.01) > 0.95
    t_ana, chord = analytic_first_hit(sp, pos, dir)
    ana = np.isfinite(t_ana)
    assert ana.sum() > 0.15 * len(pos)
    missed = ana & ~h
    assert np.all(chord[missed] < 2 * vox), "a ray through a sphere (chord >= 2 voxels) missed it"
    solid = ana & h & (chord >= 2 * vox)
    assert np.abs(hits["t"][solid] - t_ana[solid]).max() <= vox


def test_frame_octree_and_brick_sets_agree(lrt, orc, camera_batch):
    """One octree, several encodings: frame octree depth 7 against brick sets with the same
    finest lattice (128 points per axis)."""
    sp = sphere_set()
    pos, dir = camera_batch
    vox = 2.0 / 128

    def trace(build):
        s = host_scene(lrt, build)
        return orc.trace_closest(orc.Scene.from_scene(s), pos, dir)[0]

    def indexed(depth, bs):
        h, n, v, vf, nd = lrt.sdf_spheres_sbs_indexed(sp, depth, bs, 1)
        return lambda s: s.AddInstance(s.AddGeom_SdfSBS(h, n, v, vf))

    def quantised(depth, bs, bpv):
        h = np.zeros((), lrt.SBS_HEADER_DTYPE)
        h["brick_size"], h["brick_pad"], h["bytes_per_value"], h["aux_data"] = bs, 0, bpv, lrt.SDF_SBS_NODE_LAYOUT_DX
        n, v = lrt.sdf_spheres_sbs(sp, depth, h)
        return lambda s: s.AddInstance(s.AddGeom_SdfSBS(h, n, v))

    ref = trace(lambda s: s.AddInstance(s.AddGeom_SdfFrameOctree(lrt.sdf_spheres_frame_octree(sp, 7))))
    rh = ref["primId"] != MISS

    def compare(got):
        gh = got["primId"] != MISS
        both = rh & gh
        dt = np.abs(ref["t"][both] - got["t"][both]) / vox
        return int(np.sum(rh != gh)), dt, np.mean(ref["t"][both] == got["t"][both])

    # fp32 indexed, 8^3 bricks: the voxel march lands on the same voxels with the same values
    flips, dt, exact = compare(trace(indexed(4, 8)))
    assert flips == 0 and exact == 1.0
    # one-voxel bricks: a different BVH over the same voxels; the first leaf that hits can differ
    # when a ray crosses two surfaces (the walk stops at the first hit leaf in visit order)
    flips, dt, exact = compare(trace(indexed(7, 1)))
    assert flips == 0 and exact > 0.98
    # quantised distances (DX layout): 4-, 2- and 1-byte values
    flips, dt, exact = compare(trace(quantised(4, 8, 4)))
    assert flips == 0 and dt.max() < 1e-3
    flips, dt, exact = compare(trace(quantised(4, 8, 2)))
    assert flips <= 2 and np.quantile(dt, 0.999) < 0.05
    flips, dt, exact = compare(trace(quantised(4, 8, 1)))
    assert flips <= 0.02 * rh.sum() and np.quantile(dt, 0.99) < 1.0


def test_brick_colour_is_trilinear_albedo(lrt, orc, camera_batch):
    """ID32F layout: colour = trilinear over the brick's 8 corner albedos, 8-bit quantised and
    packed as coords[0] = r + g/256, coords[1] = b (BVH2Common.cpp:1082-1109)."""
    sp = sphere_set()
    pos, dir = camera_batch
    h, n, v, vf, nd = lrt.sdf_spheres_sbs_indexed(sp, 4, 8, 1)
    s = host_scene(lrt, lambda s: s.AddInstance(s.AddGeom_SdfSBS(h, n, v, vf)))
    hits = orc.trace_closest(orc.Scene.from_scene(s), pos, dir)[0]
    hit = hits["primId"] != MISS
    assert hit.sum() > 1000
    nodes = n[hits["primId"][hit]]
    lod = (nodes["pos_z_lod_size"] & 0xFFFF).astype(np.float64)
    bp = np.stack([nodes["pos_xy"] >> 16, nodes["pos_xy"] & 0xFFFF, nodes["pos_z_lod_size"] >> 16], 1).astype(np.float64)
    bmin = -1.0 + (2.0 / lod)[:, None] * bp
    p = pos[hit, :3].astype(np.float64) + hits["t"][hit, None].astype(np.float64) * dir[hit, :3]
    dp = (p - bmin) * (0.5 * lod)[:, None]
    v3 = 9 ** 3
    cidx = np.stack([v[nodes["data_offset"] + v3 + c] for c in range(8)], 1)
    rgb = np.stack([vf[cidx + ch] for ch in range(3)], 2).astype(np.float64)       # (hits, 8, 3)
    w = np.stack([trilinear(np.eye(8)[c], dp) for c in range(8)], 1)
    want = np.clip(np.floor(255.0 * np.einsum("hc,hck->hk", w, rgb) + 0.5), 0, 255)
    r = np.floor(hits["coords"][hit, 0])
    g = np.round((hits["coords"][hit, 0] - r) * 256.0)
    b = hits["coords"][hit, 1]
    got = np.stack([r, g, b], 1)
    assert np.abs(got - want).max() <= 1.0 and np.mean(got == want) > 0.97


@pytest.mark.parametrize("layout", ["rgb8", "uv16"])
def test_quantised_brick_colour_and_texcoords(lrt, orc, camera_batch, layout):
    """DX_RGB8: 8-bit corner colours, trilinear, rounded (coords[0] = r + g/256, coords[1] = b);
    DX_UV16: 16-bit corner texture coordinates, trilinear (coords = u, v)
    (BVH2Common.cpp:1035-1109)."""
    sp = sphere_set()
    pos, dir = camera_batch
    bs, bpv = 4, 1
    h = np.zeros((), lrt.SBS_HEADER_DTYPE)
    h["brick_size"], h["brick_pad"], h["bytes_per_value"] = bs, 0, bpv
    h["aux_data"] = lrt.SDF_SBS_NODE_LAYOUT_DX_RGB8 if layout == "rgb8" else lrt.SDF_SBS_NODE_LAYOUT_DX_UV16
    n, v = lrt.sdf_spheres_sbs(sp, 5, h, colour_seed=7)
    s = host_scene(lrt, lambda s: s.AddInstance(s.AddGeom_SdfSBS(h, n, v)))
    hits = orc.trace_closest(orc.Scene.from_scene(s), pos, dir)[0]
    hit = hits["primId"] != MISS
    assert hit.sum() > 1000
    nodes = n[hits["primId"][hit]]
    lod = (nodes["pos_z_lod_size"] & 0xFFFF).astype(np.float64)
    bp = np.stack([nodes["pos_xy"] >> 16, nodes["pos_xy"] & 0xFFFF, nodes["pos_z_lod_size"] >> 16], 1).astype(np.float64)
    bmin = -1.0 + (2.0 / lod)[:, None] * bp
    p = pos[hit, :3].astype(np.float64) + hits["t"][hit, None].astype(np.float64) * dir[hit, :3]
    w = np.stack([trilinear(np.eye(8)[c], (p - bmin) * (0.5 * lod)[:, None]) for c in range(8)], 1)
    dist_words = ((bs + 1) ** 3 + 4 // bpv - 1) // (4 // bpv)
    words = np.stack([v[nodes["data_offset"] + dist_words + c] for c in range(8)], 1).astype(np.uint64)
    if layout == "rgb8":
        rgb = np.stack([(words >> (8 * ch)) & 0xFF for ch in range(3)], 2).astype(np.float64)
        want = np.clip(np.floor(np.einsum("hc,hck->hk", w, rgb) + 0.5), 0, 255)
        r = np.floor(hits["coords"][hit, 0])
        got = np.stack([r, np.round((hits["coords"][hit, 0] - r) * 256.0), hits["coords"][hit, 1]], 1)
        assert np.abs(got - want).max() <= 1.0 and np.mean(got == want) > 0.97
    else:
        uv = np.stack([(words >> 16) & 0xFFFF, words & 0xFFFF], 2).astype(np.float64) / 65535.0
        want = np.einsum("hc,hck->hk", w, uv)
        assert np.abs(hits["coords"][hit, :2] - want).max() < 2e-3


@pytest.mark.parametrize("depth", [6, 7])
def test_multi_octree_matches_frame_octree(lrt, orc, camera_batch, depth):
    """The multi-octree walk (BVH2Common.cpp:2176-2363) over the octree of one sphere union hits
    the same surface as the frame-octree BVH of the same octree, and that surface is the
    analytic one."""
    sp = sphere_set()
    vox = 2.0 / 2 ** depth
    mn, mv = lrt.sdf_spheres_multi_octree(sp, depth)
    fr = lrt.sdf_spheres_frame_octree(sp, depth)
    assert mn.size == fr.size and mv.size == 8 * mn.size
    sm = host_scene(lrt, lambda s: s.AddInstance(s.AddGeom_SdfMultiOctree(mn, mv)))
    sf = host_scene(lrt, lambda s: s.AddInstance(s.AddGeom_SdfFrameOctree(fr)))
    pos, dir = camera_batch
    hm = orc.trace_closest(orc.Scene.from_scene(sm), pos, dir)[0]
    hf = orc.trace_closest(orc.Scene.from_scene(sf), pos, dir)[0]
    a, b = hm["primId"] != MISS, hf["primId"] != MISS
    assert a.sum() > 0.15 * a.size and np.array_equal(a, b)
    dt = np.abs(hm["t"][a] - hf["t"][a]) / vox
    assert dt.max() < 5e-3 and np.mean(hm["t"][a] == hf["t"][a]) > 0.8
    p = pos[a, :3].astype(np.float64) + hm["t"][a, None].astype(np.float64) * dir[a, :3]
    f, _ = union_distance(sp, p)
    assert np.abs(f).max() <= 0.35 * vox


def test_multi_octree_level_of_detail(lrt, orc, camera_batch):
    """With a level-of-detail cut (fixed_lod = 1, level_of_detail = 5: nodes of 32 cells per axis
    act as leaves) the walk stops at depth 5 and matches the frame octree built at depth 5."""
    sp = sphere_set()
    mn, mv = lrt.sdf_spheres_multi_octree(sp, 7)
    sm = host_scene(lrt, lambda s: s.AddInstance(s.AddGeom_SdfMultiOctree(mn, mv)))
    p = lrt.default_preset()
    p["fixed_lod"], p["level_of_detail"] = 1, 5.0
    sm.SetPreset(p)
    f5 = lrt.sdf_spheres_frame_octree(sp, 5)
    sf = host_scene(lrt, lambda s: s.AddInstance(s.AddGeom_SdfFrameOctree(f5)))
    pos, dir = camera_batch
    hm = orc.trace_closest(orc.Scene.from_scene(sm), pos, dir)[0]
    hf = orc.trace_closest(orc.Scene.from_scene(sf), pos, dir)[0]
    a, b = hm["primId"] != MISS, hf["primId"] != MISS
    assert a.sum() > 0.15 * a.size
    assert np.sum(a != b) <= 0.002 * a.size
    both = a & b
    assert np.quantile(np.abs(hm["t"][both] - hf["t"][both]), 0.99) < 1e-3 * (2.0 / 32)
