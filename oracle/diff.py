"""ORACLE - test infrastructure only. Float64 twin of the differentiable brick-set renderer.

The float32 oracle march (oracle_sbs_hit_voxels, BVH2Common.cpp:884-1112) finds each ray's hit
voxel; everything after that is float64 here:
  colour   C = sum_k w_k(q) A_k                     (trilinear albedo at the hit point q)
  adjoint  dL/dA_k = G w_k,  dL/ds_j = -(sum_c G_c grad C_c . d) w_j / (grad f . d)
  root     t* = the float64 root of f(q0 + t d) = sum_j s_j w_j(q0 + t d) near the march's t
  FD       central differences of the float64 loss, re-solving t* per perturbed parameter
How each is pinned (tests/test_diff_cpu.py): the adjoint against central differences on
random parameters (distances and albedo), the colour against the closed-form trilinear, the
root against the march's own hit parameter.
"""
import ctypes

import numpy as np

from . import HIT_DTYPE, OracleScene, _p, load

VOXEL_REC_DTYPE = np.dtype([("hit", "<u4"), ("idx", "<u4", (8,)), ("start_q", "<f4", (3,)), ("dir", "<f4", (3,)),
                            ("t_local", "<f4"), ("t", "<f4"), ("pad", "<f4")])
assert VOXEL_REC_DTYPE.itemsize == 72

BITS = np.array([[(k >> 2) & 1, (k >> 1) & 1, k & 1] for k in range(8)], np.float64)


def hit_voxels(scene, pos, dir, threads: int = 0, flavor: str = "exact"):
    """(hits, voxel records) for a ray batch."""
    L = load(flavor)
    L.oracle_sbs_hit_voxels.argtypes = [ctypes.POINTER(OracleScene), ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    L.oracle_sbs_hit_voxels.restype = None
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 4)
    dir = np.ascontiguousarray(dir, dtype=np.float32).reshape(-1, 4)
    n = pos.shape[0]
    hits = np.zeros(n, HIT_DTYPE)
    recs = np.zeros(n, VOXEL_REC_DTYPE)
    L.oracle_sbs_hit_voxels(ctypes.byref(scene.view), _p(pos), _p(dir), n, threads, _p(hits), _p(recs))
    return hits, recs


def weights(q):
    """Trilinear weights (n, 8) and their gradients (n, 8, 3) at points q (n, 3)."""
    f = np.where(BITS[None] == 1, q[:, None, :], 1.0 - q[:, None, :])       # (n, 8, 3) factors
    sgn = np.where(BITS == 1, 1.0, -1.0)                                     # (8, 3)
    w = f.prod(axis=2)
    gw = np.stack([sgn[None, :, 0] * f[:, :, 1] * f[:, :, 2], f[:, :, 0] * sgn[None, :, 1] * f[:, :, 2],
                   f[:, :, 0] * f[:, :, 1] * sgn[None, :, 2]], axis=2)
    return w, gw


class Layout:
    """Index arithmetic of the differentiable layout: albedo of value index g."""

    def __init__(self, f_offset, n_dist):
        self.f_offset, self.n_dist = int(f_offset), int(n_dist)

    def albedo(self, idx):
        return self.f_offset + self.n_dist + 3 * (idx.astype(np.int64) - self.f_offset)


def root(recs, params, t0=None, iters=30):
    """Float64 root of f(q0 + t d) near the march's hit parameter (Newton on the cubic)."""
    q0 = recs["start_q"].astype(np.float64)
    d = recs["dir"].astype(np.float64)
    s = params[recs["idx"]].astype(np.float64)                             # (n, 8)
    t = recs["t_local"].astype(np.float64) if t0 is None else t0.copy()
    for _ in range(iters):
        w, gw = weights(q0 + t[:, None] * d)
        f = (w * s).sum(1)
        df = np.einsum("nk,nkc,nc->n", s, gw, d)
        step = np.where(np.abs(df) > 1e-300, f / np.where(df == 0, 1.0, df), 0.0)
        t = t - step
    return t


def colour(recs, params, layout, t):
    """Colour (n, 3) at q = q0 + t d (float64)."""
    q = recs["start_q"].astype(np.float64) + t[:, None] * recs["dir"].astype(np.float64)
    w, _ = weights(q)
    ab = layout.albedo(recs["idx"])                                           # (n, 8)
    A = np.stack([params[ab + c] for c in range(3)], 2).astype(np.float64)    # (n, 8, 3)
    return np.einsum("nk,nkc->nc", w, A)


def adjoint(recs, params, layout, G, t):
    """dL/dparams (float64, len(params)) for upstream gradients G (n, 3) at hit parameters t."""
    grad = np.zeros(params.shape[0], np.float64)
    h = recs["hit"] == 1
    r, G, t = recs[h], G[h].astype(np.float64), t[h]
    q = r["start_q"].astype(np.float64) + t[:, None] * r["dir"].astype(np.float64)
    d = r["dir"].astype(np.float64)
    w, gw = weights(q)
    ab = layout.albedo(r["idx"])
    A = np.stack([params[ab + c] for c in range(3)], 2).astype(np.float64)
    for c in range(3):
        np.add.at(grad, ab + c, G[:, c:c + 1] * w)
    s = params[r["idx"]].astype(np.float64)
    den = np.einsum("nk,nkc,nc->n", s, gw, d)
    gcd = np.einsum("nc,nkc,nk->n", G, A, np.einsum("nkc,nc->nk", gw, d))
    ok = np.abs(den) > 1e-20
    ds = -(gcd / np.where(ok, den, 1.0))[:, None] * w
    np.add.at(grad, r["idx"][ok], ds[ok])
    return grad


def l2_loss(C, hit, target):
    """Sum over rays of |C - T|^2 with C = 0 on misses."""
    Cf = np.where(hit[:, None], C, 0.0)
    return float(((Cf - target[:, :3].astype(np.float64)) ** 2).sum())


def finite_differences(recs, params, layout, target, param_ids, eps_rel=1e-3):
    """Central differences of the float64 L2 loss for the listed parameters. Distances move the
    hit (t re-solved per perturbation); the hit voxel is kept (the perturbation is small)."""
    h = recs["hit"] == 1
    r = recs[h]
    T = target[h, :3].astype(np.float64)
    p64 = params.astype(np.float64)
    t_star = root(r, p64)
    ab = layout.albedo(r["idx"])
    # rays that read each parameter (distances through idx, albedo through ab..ab+2), as
    # (parameter, ray) pairs sorted by parameter
    ray_ids = np.repeat(np.arange(r.size, dtype=np.int64), 8)
    pairs_p = np.concatenate([col.astype(np.int64).reshape(-1) for col in (r["idx"], ab, ab + 1, ab + 2)])
    pairs_r = np.tile(ray_ids, 4)
    order = np.argsort(pairs_p, kind="stable")
    pairs_p, pairs_r = pairs_p[order], pairs_r[order]
    out = np.zeros(len(param_ids))
    for n, j in enumerate(param_ids):
        lo, hi = np.searchsorted(pairs_p, j, "left"), np.searchsorted(pairs_p, j, "right")
        rows = np.unique(pairs_r[lo:hi])
        if rows.size == 0:
            continue
        rr, TT = r[rows], T[rows]
        eps = eps_rel * max(abs(p64[j]), 1e-3)
        keep = p64[j]
        losses = []
        for sgn in (1.0, -1.0):
            p64[j] = keep + sgn * eps
            t = root(rr, p64, t0=t_star[rows])
            losses.append(((colour(rr, p64, layout, t) - TT) ** 2).sum())
        p64[j] = keep
        out[n] = (losses[0] - losses[1]) / (2 * eps)
    return out
