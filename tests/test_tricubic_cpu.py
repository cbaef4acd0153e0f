"""Tricubic interpolation on the CPU: brick sets (OctreeBrickIntersect with
load_tricubic_distance_values and TricubicLocalSurfaceIntersection, MultiRenderer/BVH2Common.cpp:
790-844, 884-955, 220-335) and dense grids (eval_distance_sdf_grid, :1478-1520), with the
Catmull-Rom spline of tricubicInterpolation (:1526-1569).

The oracle is pinned by what the mathematics fixes: a Catmull-Rom spline reproduces quadratic
polynomials exactly, so with the samples of f(p) = |p - c|^2 - r^2 the interpolant is f itself and
every hit lies on the sphere |p - c| = r (to the marches' stopping distance) with the normal along
p - c, where trilinear interpolation of the same samples misses the sphere by the square of the
cell size; a wrong sample order or a dropped term of the spline breaks that. Also fixed: the
modes other than sphere tracing and BBOX hit only rays that start inside (the reference solves
nothing else under tricubic), BBOX hits lie at the entry of a voxel whose 4^3 samples reach zero.
"""
import numpy as np
import pytest

from conftest import MISS, decode_normal, host_scene

C = np.array([0.05, -0.03, 0.02])
R = 0.35


def quad(p):
    return np.sum((p - C) ** 2, axis=-1) - R * R


def quadratic_sbs(lrt, sz=8, bs=4, pad=1):
    """Indexed fp32 brick set (ID32F_IRGB32F) of f over the bricks of an sz^3 level whose padded
    lattices reach zero; every brick has its own samples and a grey colour."""
    v = bs + 2 * pad + 1
    h = 2.0 / (sz * bs)
    a = (np.arange(v) - pad) * h
    nodes, words, vals = [], [], []
    for ix in range(sz):
        for iy in range(sz):
            for iz in range(sz):
                bmin = -1.0 + 2.0 * np.array([ix, iy, iz]) / sz
                P = np.stack(np.meshgrid(bmin[0] + a, bmin[1] + a, bmin[2] + a, indexing="ij"), -1)
                f = quad(P).astype(np.float32)
                if not (f.min() <= 0.0 <= f.max()):
                    continue
                base = len(vals)
                nodes.append(((ix << 16) | iy, (iz << 16) | sz, len(words), 0))
                words.extend(range(base, base + v ** 3))     # lattice order x, y, z (z fastest)
                vals.extend(f.reshape(-1).tolist())
                words.extend([0] * 8)                        # colour indices, patched below
    grey = len(vals)
    vals.extend([0.5, 0.5, 0.5])
    words = np.array(words, np.uint32)
    for n in nodes:
        words[n[2] + v ** 3: n[2] + v ** 3 + 8] = grey
    hdr = np.zeros((), lrt.SBS_HEADER_DTYPE)
    hdr["brick_size"], hdr["brick_pad"], hdr["bytes_per_value"] = bs, pad, 4
    hdr["aux_data"] = lrt.SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F
    return hdr, np.array(nodes, np.uint32).view(lrt.SBS_NODE_DTYPE).reshape(-1), words, np.array(vals, np.float32)


# the grid's sphere tracer steps by the sampled value itself: scaled so that its gradient
# |p - c| / 1.8 stays below 1 over the cube, the quadratic never oversteps the surface
GRID_SCALE = 1.0 / (2.0 * 1.8)


def quadratic_grid(n=48):
    a = np.linspace(-1.0, 1.0, n)
    P = np.stack(np.meshgrid(a, a, a, indexing="ij"), -1)   # [x, y, z]
    return np.ascontiguousarray((GRID_SCALE * quad(P)).astype(np.float32).transpose(2, 1, 0))   # [z, y, x]


def rays(n=2500, seed=7):
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o = 2.5 * o / np.linalg.norm(o, axis=1, keepdims=True)
    tgt = C + rng.uniform(-0.3, 0.3, size=(n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = np.zeros((n, 4), np.float32)
    dir_ = np.zeros((n, 4), np.float32)
    pos[:, :3], dir_[:, :3], dir_[:, 3] = o, d, 1e9
    return pos, dir_


def preset(lrt, interp, mode=0, render=8):
    p = lrt.default_preset()
    p["interpolation_mode"], p["sdf_node_intersect"], p["render_mode"] = interp, mode, render
    return p


def trace(lrt, orc, add, p, pos, dir_):
    def build(s):
        s.AddInstance(add(s))
        s.SetPreset(p)
    return orc.trace_closest(orc.Scene.from_scene(host_scene(lrt, build)), pos, dir_)[0]


def sphere_errors(rec, pos, dir_):
    h = rec["primId"] != MISS
    p = pos[h, :3].astype(np.float64) + rec["t"][h, None].astype(np.float64) * dir_[h, :3]
    dist = np.abs(np.linalg.norm(p - C, axis=1) - R)
    n = decode_normal(rec["coords"][h, 2:])
    want = (p - C) / np.linalg.norm(p - C, axis=1, keepdims=True)
    ang = np.arccos(np.clip(np.sum(n * want, axis=1), -1, 1))
    return h, dist, ang


@pytest.fixture(scope="module")
def sbs(lrt):
    return quadratic_sbs(lrt)


def test_brick_set_tricubic_reproduces_quadratics(lrt, orc, sbs):
    pos, dir_ = rays()
    cub = trace(lrt, orc, lambda s: s.AddGeom_SdfSBS(*sbs), preset(lrt, 1), pos, dir_)
    lin = trace(lrt, orc, lambda s: s.AddGeom_SdfSBS(*sbs), preset(lrt, 0), pos, dir_)
    h, dist, ang = sphere_errors(cub, pos, dir_)
    assert h.mean() > 0.5
    assert dist.max() < 1e-5 and ang.max() < 2e-3
    hl, dl, _ = sphere_errors(lin, pos, dir_)
    assert dl.max() > 10 * dist.max() and dl.max() > 2e-5   # trilinear misses the sphere by O(h^2)
    assert (h != (lin["primId"] != MISS)).mean() < 0.01     # grazing rays aside, the same hits


def test_brick_set_tricubic_modes(lrt, orc, sbs):
    """ANALYTIC/NEWTON/IT under tricubic do not march inside a voxel: a ray is hit at the entry
    face of the first voxel whose entry sample is inside (f <= 1e-6), just past the sphere; BBOX
    hits at the entry of a voxel of the surface's neighbourhood, before the sphere-traced hit;
    ST_ITERATIONS reports the march's step count."""
    pos, dir_ = rays(1500)
    st = trace(lrt, orc, lambda s: s.AddGeom_SdfSBS(*sbs), preset(lrt, 1, 0), pos, dir_)
    hs = st["primId"] != MISS
    cell = 2.0 / (8 * 4)   # voxel size of quadratic_sbs: sz 8, brick size 4
    for mode in (1, 2, 4):
        r = trace(lrt, orc, lambda s: s.AddGeom_SdfSBS(*sbs), preset(lrt, 1, mode), pos, dir_)
        hr = r["primId"] != MISS
        assert np.array_equal(hr, hs) or (hr != hs).mean() < 0.01
        both = hr & hs
        assert np.all(r["t"][both] >= st["t"][both])
        p = pos[hr, :3].astype(np.float64) + r["t"][hr, None].astype(np.float64) * dir_[hr, :3]
        assert np.all(quad(p) <= 1e-5)                                   # inside (float32 rounding of t)
        g = (p + 1.0) / cell
        assert np.all(np.min(np.abs(g - np.round(g)), axis=1) < 1e-3)   # on a voxel face
    bb = trace(lrt, orc, lambda s: s.AddGeom_SdfSBS(*sbs), preset(lrt, 1, 3), pos, dir_)
    hb = bb["primId"] != MISS
    assert np.all(hb[hs]) and np.all(bb["t"][hs] <= st["t"][hs])
    it = trace(lrt, orc, lambda s: s.AddGeom_SdfSBS(*sbs), preset(lrt, 1, 0, 10), pos, dir_)
    assert np.array_equal(it["primId"] != MISS, hs) and it["primId"][hs].max() < 256


def test_brick_set_tricubic_from_inside(lrt, orc, sbs):
    """Rays from the centre (inside, f < 0). VOLUME representation: the first stored voxel
    entered counts as hit at its entry face (its start sample is negative), normal outward.
    SURFACE representation: the start sample's sign is flipped, so each voxel's march takes one
    step of |f| and stops, because the reference's tricubic march compares the next sample
    without the sign (BVH2Common.cpp:254-259); a step that leaves the voxel is no hit, so the hit
    comes later, still inside the sphere, with the normal turned inward."""
    rng = np.random.default_rng(2)
    d = rng.normal(size=(500, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = np.zeros((500, 4), np.float32)
    dir_ = np.zeros((500, 4), np.float32)
    pos[:, :3], dir_[:, :3], dir_[:, 3] = C, d, 1e9
    vol = trace(lrt, orc, lambda s: s.AddGeom_SdfSBS(*sbs), preset(lrt, 1), pos, dir_)   # default: VOLUME
    p = preset(lrt, 1)
    p["representation_mode"] = 0   # SURFACE
    srf = trace(lrt, orc, lambda s: s.AddGeom_SdfSBS(*sbs), p, pos, dir_)
    hv, hs = vol["primId"] != MISS, srf["primId"] != MISS
    assert hv.mean() > 0.9 and hs.mean() > 0.9
    both = hv & hs
    pv = pos[both, :3].astype(np.float64) + vol["t"][both, None].astype(np.float64) * dir_[both, :3]
    assert np.all(quad(pv) < 0)
    g = (pv + 1.0) / (2.0 / 32)
    assert np.all(np.min(np.abs(g - np.round(g)), axis=1) < 1e-3)   # on a voxel face
    assert np.all(np.sum(decode_normal(vol["coords"][both, 2:]) * (pv - C), axis=1) > 0)
    assert np.all(srf["t"][both] > vol["t"][both])
    ps = pos[both, :3].astype(np.float64) + srf["t"][both, None].astype(np.float64) * dir_[both, :3]
    assert np.all(quad(ps) <= 1e-5)
    assert np.all(np.sum(decode_normal(srf["coords"][both, 2:]) * (ps - C), axis=1) < 0)


def test_grid_tricubic_reproduces_quadratics(lrt, orc):
    pos, dir_ = rays(1500)
    g = quadratic_grid()
    cub = trace(lrt, orc, lambda s: s.AddGeom_SdfGrid(g), preset(lrt, 1), pos, dir_)
    lin = trace(lrt, orc, lambda s: s.AddGeom_SdfGrid(g), preset(lrt, 0), pos, dir_)
    h, dist, ang = sphere_errors(cub, pos, dir_)
    assert h.mean() > 0.5 and dist.max() < 1e-4 and ang.max() < 2e-3   # stop at 1e-5 of the scaled f
    _, dl, _ = sphere_errors(lin, pos, dir_)
    assert dl.max() > 5 * dist.max()


def test_interpolation_mode_values(lrt):
    s = lrt.SceneRT()
    for mode in (0, 1):
        s.SetPreset(preset(lrt, mode))
    with pytest.raises(lrt.LrtError) as e:
        s.SetPreset(preset(lrt, 2))
    assert e.value.code == 4   # LRT_ERR_UNSUPPORTED
