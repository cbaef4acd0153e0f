"""Pins the float64 twin of the differentiable brick-set renderer (oracle/diff.py), which is the
reference the GPU adjoint is checked against: the adjoint against central finite differences
of the float64 loss on random parameters (distances and albedo), the colour against the
closed-form trilinear, and the float64 root against the march's own hit parameter."""
import numpy as np
import pytest

from conftest import host_scene


def spheres(seed, n=64):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(-0.8, 0.8, (n, 3)), rng.uniform(0.05, 0.2, (n, 1))], 1).astype(np.float32)


@pytest.fixture(scope="module")
def diff_case(lrt, orc):
    import oracle.diff as od
    h, nodes, vals, vf, nd = lrt.sdf_spheres_sbs_indexed(spheres(4), 3, 8, 1)
    s = host_scene(lrt, lambda s: s.AddInstance(s.AddGeom_SdfSBS(h, nodes, vals, vf)))
    h5, n5, v5, vf5, nd5 = lrt.sdf_spheres_sbs_indexed(spheres(5), 3, 8, 2)
    s5 = host_scene(lrt, lambda s: s.AddInstance(s.AddGeom_SdfSBS(h5, n5, v5, vf5)))
    pi, vi = orc.camera_inverse([0, 0, 3], [0, 0, 0], [0, 1, 0], 60.0, 1.0, 0.01, 100.0)
    pos, dir = orc.eye_rays(pi, vi, 64, 64, order=0)
    hits, recs = od.hit_voxels(orc.Scene.from_scene(s), pos, dir)
    _, r5 = od.hit_voxels(orc.Scene.from_scene(s5), pos, dir)
    m5 = r5["hit"] == 1
    T = np.zeros((pos.shape[0], 4))
    T[m5, :3] = od.colour(r5[m5], vf5.astype(np.float64), od.Layout(0, nd5), r5["t_local"][m5].astype(np.float64))
    return dict(od=od, vf=vf, nd=nd, hits=hits, recs=recs, T=T, pos=pos, dir=dir, vals=vals, nodes=nodes)


def test_every_hit_has_a_voxel_record(diff_case):
    hits, recs = diff_case["hits"], diff_case["recs"]
    hit = hits["primId"] != 0xFFFFFFFF
    assert hit.sum() > 500
    assert np.array_equal(recs["hit"] == 1, hit)
    assert np.array_equal(recs["t"][hit], hits["t"][hit])


def test_root_and_colour_closed_form(diff_case):
    od, vf, nd, recs = diff_case["od"], diff_case["vf"], diff_case["nd"], diff_case["recs"]
    m = recs["hit"] == 1
    r = recs[m]
    p64 = vf.astype(np.float64)
    t = od.root(r, p64)
    # the float64 root is within the march's float32 error of its hit parameter, on the surface
    assert np.abs(t - r["t_local"]).max() < 1e-2 and np.median(np.abs(t - r["t_local"])) < 1e-5
    q = r["start_q"].astype(np.float64) + t[:, None] * r["dir"].astype(np.float64)
    w, _ = od.weights(q)
    assert np.abs((w * p64[r["idx"]]).sum(1)).max() < 1e-12
    assert np.all(q > -1e-4) and np.all(q < 1 + 1e-4)
    # colour = explicit trilinear sum over the 8 corner albedos
    C = od.colour(r, p64, od.Layout(0, nd), t)
    k = 17
    ab = nd + 3 * r["idx"][k].astype(np.int64)
    qq = q[k]
    want = np.zeros(3)
    for c in range(8):
        bx, by, bz = (c >> 2) & 1, (c >> 1) & 1, c & 1
        wc = (qq[0] if bx else 1 - qq[0]) * (qq[1] if by else 1 - qq[1]) * (qq[2] if bz else 1 - qq[2])
        want += wc * p64[ab[c]:ab[c] + 3]
    assert np.allclose(C[k], want, rtol=0, atol=1e-14)


def test_adjoint_matches_central_differences(diff_case):
    od, vf, nd, recs, T = diff_case["od"], diff_case["vf"], diff_case["nd"], diff_case["recs"], diff_case["T"]
    lay = od.Layout(0, nd)
    p64 = vf.astype(np.float64)
    m = recs["hit"] == 1
    t = np.zeros(recs.size)
    t[m] = od.root(recs[m], p64)
    C = np.zeros((recs.size, 3))
    C[m] = od.colour(recs[m], p64, lay, t[m])
    grad = od.adjoint(recs, p64, lay, 2 * (C - T[:, :3]), t)
    nz = np.nonzero(grad)[0]
    rng = np.random.default_rng(6)
    for ids, bar in [(rng.choice(nz[nz < nd], 150, replace=False), 1e-4),
                     (rng.choice(nz[nz >= nd], 150, replace=False), 1e-7)]:
        fd = od.finite_differences(recs, vf, lay, T, ids)
        rel = np.abs(fd - grad[ids]) / np.maximum(np.abs(grad[ids]), 1e-9)
        assert np.quantile(rel, 0.99) < bar and rel.max() < 100 * bar, (rel.max(), np.quantile(rel, 0.99))


def test_planted_adjoint_sign_error_is_caught(diff_case):
    """Flipping the sign of the distance term must fail the finite-difference check."""
    od, vf, nd, recs, T = diff_case["od"], diff_case["vf"], diff_case["nd"], diff_case["recs"], diff_case["T"]
    lay = od.Layout(0, nd)
    p64 = vf.astype(np.float64)
    m = recs["hit"] == 1
    t = np.zeros(recs.size)
    t[m] = od.root(recs[m], p64)
    C = np.zeros((recs.size, 3))
    C[m] = od.colour(recs[m], p64, lay, t[m])
    G = 2 * (C - T[:, :3])
    full = od.adjoint(recs, p64, lay, G, t)
    alb_only = od.adjoint(recs, p64, lay, G, t) - 0.0
    alb_only[:nd] = 0.0
    planted = alb_only - (full - alb_only)       # distance part with the wrong sign
    ids = np.random.default_rng(7).choice(np.nonzero(full[:nd])[0], 50, replace=False)
    fd = od.finite_differences(recs, vf, lay, T, ids)
    rel = np.abs(fd - planted[ids]) / np.maximum(np.abs(full[ids]), 1e-9)
    assert np.quantile(rel, 0.5) > 0.5
