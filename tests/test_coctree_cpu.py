"""COctreeV3 leaves on the CPU: the oracle's restatement of COctreeIntersectV3 and its loaders
(MultiRenderer/BVH2Common.cpp:1637-2173) pinned against an independent encoder, the reference's
own engine-scene file (scenes/06_engine_scenes/scom_tree/sdf_coctree_v3_0.bin, copied under
tests/golden/), the frame octree and brick set of the same sphere union, and the host checks of
the library (AddGeom_COctreeV3, BVH2Common_host.cpp:915-966)."""
import ctypes
import os

import numpy as np
import pytest

import coct_encode as ce
from conftest import MISS, host_scene, raw_equal

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
COCT = os.path.join(GOLD, "engine_sdf_coctree_v3_0.bin")
SPHERES = np.array([[0.1, 0.05, -0.05, 0.55], [-0.35, 0.2, 0.25, 0.3], [0.3, -0.3, 0.3, 0.25]], np.float32)


def sphere_rays(n, seed=3, radius=2.5):
    """Rays from a sphere around the scene towards jittered points near the centre."""
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o = radius * o / np.linalg.norm(o, axis=1, keepdims=True)
    tgt = rng.uniform(-0.7, 0.7, size=(n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = np.zeros((n, 4), np.float32)
    dir_ = np.zeros((n, 4), np.float32)
    pos[:, :3], pos[:, 3] = o, 0.0
    dir_[:, :3], dir_[:, 3] = d, 1e9
    return pos, dir_


def oracle_trace(orc, lrt, scene, pos, dir_):
    o = orc.Scene.from_scene(scene)
    return orc.trace_closest(o, pos, dir_)[0]


def coct_scene(lrt, header, words, bvh_level=0, preset=None):
    def build(s):
        g = s.AddGeom_COctreeV3(header, words, bvh_level)
        s.AddInstance(g, np.eye(4, dtype=np.float32))
        if preset is not None:
            s.SetPreset(preset)
    return host_scene(lrt, build)


@pytest.fixture(scope="module")
def gotree(lrt):
    """Global octree of the sphere union, depth 5, for brick sizes 1 and 2 (pad 0 and 1)."""
    out = {}
    for bs, pad in [(1, 0), (1, 1), (2, 0), (2, 1), (3, 0)]:
        out[(bs, pad)] = lrt.sdf_spheres_global_octree(SPHERES, 5 if bs == 1 else 4, bs, pad)
    return out


# ---- rotation modifiers ----
def test_rot_modifiers_are_the_reference_search(lrt, gotree):
    """The library's closed form equals initialize_rot_modifiers' own search (restated in numpy)
    for the value, voxel and child lattices of several brick shapes, v = 1 included."""
    for (bs, pad) in [(1, 0), (2, 1), (3, 0)]:
        nodes, values = gotree[(bs, pad)]
        h, w = ce.encode(nodes, values, bs, pad, leaf_type=ce.GRID)
        sc = coct_scene(lrt, h, w)
        mods = sc.get_array("coct_mods").reshape(3, 48, 4)
        v = bs + 2 * pad + 1
        assert np.array_equal(mods[0], ce.rot_modifiers_search(v))
        assert np.array_equal(mods[1], ce.rot_modifiers_search(v - 1))
        assert np.array_equal(mods[2], ce.rot_modifiers_search(2))


def test_rot_modifiers_map_the_lattice():
    """Each modifier sends a lattice point to the row-major index of its rotated image, the 48
    images of one point are distinct where the lattice allows it, and a planted transposition of
    two coefficients is caught."""
    v = 4
    m = ce._closed_mods(v)
    pts = np.array([(x, y, z) for x in range(v) for y in range(v) for z in range(v)])
    for i, rot in enumerate(ce.rot_transforms(v)):
        img = np.array([ce.rotate_point(rot, v, p) for p in pts])
        assert np.array_equal(pts @ m[i, :3] + m[i, 3], img[:, 0] * v * v + img[:, 1] * v + img[:, 2])
    assert len({tuple(r) for r in m.tolist()}) == 48
    bad = m[5].copy()
    bad[[0, 1]] = bad[[1, 0]]
    img = np.array([ce.rotate_point(ce.rot_transforms(v)[5], v, p) for p in pts])
    assert not np.array_equal(pts @ bad[:3] + bad[3], img[:, 0] * v * v + img[:, 1] * v + img[:, 2])


# ---- the reference's file ----
def test_reference_file_layout(lrt, tmp_path):
    h, w, ver = lrt.load_coctree_v3(COCT)
    assert ver == 5
    assert (int(h["brick_size"]), int(h["brick_pad"]), int(h["bits_per_value"])) == (1, 0, 8)
    assert (int(h["default_leaf_type"]), int(h["fallback_leaf_type"]), int(h["node_pack_mode"])) == (2, 1, 0)
    assert w.size == 19484   # scene.xml: num_primitives
    out = tmp_path / "c.bin"
    lrt.save_coctree_v3(str(out), h, w)
    assert open(out, "rb").read() == open(COCT, "rb").read()


def border_boxes(words, upc, max_level):
    """add_border_nodes_rec (BVH2Common_host.cpp:76-117) restated in Python (float32 arithmetic)."""
    out = []

    def rec(node, p, d, level):
        if node >= words.size:
            return
        info = int(words[node])
        if (info & 0xFFFF00) or max_level == level:
            mn = np.float32(-1) + (np.float32(2) * p) * np.float32(d)
            mx = np.float32(-1) + (np.float32(2) * (p + np.float32(1))) * np.float32(d)
            out.append((node, mn, mx))
            return
        k = 0
        for i in range(8):
            if info & (1 << i):
                ch = np.float32(2) * p + np.array([(i >> 2) & 1, (i >> 1) & 1, i & 1], np.float32)
                rec(int(words[node + 1 + upc * k]), ch, np.float32(d) / np.float32(2), level + 1)
                k += 1

    rec(0, np.zeros(3, np.float32), np.float32(1), 0)
    return out


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_reference_file_blas_boxes(lrt, level):
    h, w, _ = lrt.load_coctree_v3(COCT)
    sc = coct_scene(lrt, h, w, bvh_level=level)
    boxes = sc.get_array("orig_nodes")
    want = border_boxes(w, 1, level)
    if len(want) < 2:
        want = [(0, np.full(3, -1, np.float32), np.ones(3, np.float32)),
                (0, np.full(3, -1.0001, np.float32), np.full(3, -1, np.float32))]
    assert len(boxes) == len(want)
    for b, (node, mn, mx) in zip(boxes, want):
        assert int(b["leftOffset"]) == node
        assert np.array_equal(b["boxMin"], mn) and np.array_equal(b["boxMax"], mx)


def test_reference_file_surface(lrt, orc):
    """Rays from all around hit the reference's octree where its decoded distances vanish: at each
    hit the trilinear field of the hit brick, from the oracle's own loader, is within a fraction of
    a voxel of zero, and every ray aimed at the centre hits."""
    h, w, _ = lrt.load_coctree_v3(COCT)
    sc = coct_scene(lrt, h, w)
    pos, dir_ = sphere_rays(4000, radius=2.5)
    hits = oracle_trace(orc, lrt, sc, pos, dir_)
    hit = hits["primId"] != MISS
    assert hit.mean() > 0.2
    assert np.all((hits["geomId"][hit] >> 24) == 3)   # reported as TYPE_SDF_FRAME_OCTREE
    # rays aimed at the middle of the object hit
    p0, d0 = sphere_rays(300, seed=9, radius=2.5)
    d0[:, :3] = -p0[:, :3] / np.linalg.norm(p0[:, :3], axis=1, keepdims=True)
    assert (oracle_trace(orc, lrt, sc, p0, d0)["primId"] != MISS).mean() > 0.95
    # the field at the hit point: walk to the leaf containing it and evaluate its voxel
    o = orc.Scene.from_scene(sc)
    L = orc.load("exact")
    pts = pos[hit, :3] + hits["t"][hit, None] * dir_[hit, :3]
    leaves = coct_leaves(w, h)
    checked = 0
    for q in pts[:400]:
        leaf = find_leaf(leaves, q)
        if leaf is None:
            continue
        typ, off, code, sz, bmin = leaf
        vox = np.floor(np.clip((q - bmin) * (0.5 * sz), 1e-6, 1 - 1e-6)).astype(int)
        vals = np.zeros(8, np.float32)
        r = L.oracle_coct_voxel(ctypes.byref(o.view), 0, typ, off, int(vox[0]), int(vox[1]), int(vox[2]),
                                float(sz), code, vals.ctypes.data)
        if r > 1e5:
            continue
        dp = (q - bmin) * (0.5 * sz) - vox
        x, y, z = dp
        f = ((1 - x) * (1 - y) * (1 - z) * vals[0] + (1 - x) * (1 - y) * z * vals[1] + (1 - x) * y * (1 - z) * vals[2] +
             (1 - x) * y * z * vals[3] + x * (1 - y) * (1 - z) * vals[4] + x * (1 - y) * z * vals[5] +
             x * y * (1 - z) * vals[6] + x * y * z * vals[7])
        assert abs(f) < 0.05 * (2.0 / sz), (q, f, sz)
        checked += 1
    assert checked > 200


def test_reference_file_matches_its_frame_octree(lrt, orc):
    """The reference's engine scenes hold one object in several formats: its COctreeV3 file (8-bit
    values) and its frame-octree file (tests/golden/engine_sdf_frame_octree_0.bin, float corners)
    are hit by the same rays, at t within the 8-bit quantisation."""
    h, w, _ = lrt.load_coctree_v3(COCT)
    pos, dir_ = sphere_rays(4000)
    a = oracle_trace(orc, lrt, coct_scene(lrt, h, w), pos, dir_)
    fo = lrt.load_sdf_frame_octree(os.path.join(GOLD, "engine_sdf_frame_octree_0.bin"))

    def build(s):
        s.AddInstance(s.AddGeom_SdfFrameOctree(fo))
    b = oracle_trace(orc, lrt, host_scene(lrt, build), pos, dir_)
    ha, hb = a["primId"] != MISS, b["primId"] != MISS
    assert ha.mean() > 0.5 and (ha != hb).mean() < 2e-3
    dt = np.abs(a["t"][ha & hb] - b["t"][ha & hb])
    assert np.median(dt) < 1e-3 and np.percentile(dt, 99) < 5e-3


def coct_leaves(words, header):
    """Every leaf of a default-pack COctreeV3 without LODs: (type, offset, code, level size, brick min)."""
    upc = int(header["uints_per_child_info"])
    out = []

    def rec(node, p, sz):
        info = int(words[node])
        k = 0
        for c in range(8):
            if not info & (1 << c):
                continue
            t = (info >> (8 + 2 * c)) & 3
            word = int(words[node + 1 + upc * k])
            cp = 2 * p + np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
            if t == 0:
                rec(word, cp, 2 * sz)
            else:
                out.append((t, word, word & 0x3FFFFFFF, 2 * sz, -1.0 + (2.0 / (2 * sz)) * cp))
            k += 1

    rec(0, np.zeros(3, int), 1)
    return out


def find_leaf(leaves, q):
    for lf in leaves:
        bmin, sz = lf[4], lf[3]
        if np.all(q >= bmin - 1e-6) and np.all(q <= bmin + 2.0 / sz + 1e-6):
            return lf
    return None


# ---- the loaders against the encoder's input ----
LOADER_CASES = [
    dict(bs=1, pad=0, bits=8, leaf_type=ce.GRID, pack_mode=0),
    dict(bs=1, pad=0, bits=6, leaf_type=ce.BIT_PACK, pack_mode=0),
    dict(bs=1, pad=1, bits=10, leaf_type=ce.BIT_PACK, pack_mode=1, rotate=True, sim=1),
    dict(bs=2, pad=0, bits=16, leaf_type=ce.SLICES, pack_mode=0),
    dict(bs=2, pad=1, bits=8, leaf_type=ce.SLICES, pack_mode=2, rotate=True, sim=1),
    dict(bs=3, pad=0, bits=32, leaf_type=ce.GRID, pack_mode=1, rotate=True),
    dict(bs=2, pad=1, bits=8, leaf_type=ce.GRID, pack_mode=2, rotate=True, sim=1, uv=1),
]


@pytest.mark.parametrize("case", LOADER_CASES, ids=lambda c: "-".join(f"{k}{v}" for k, v in c.items()))
def test_loaders_return_the_encoded_values(lrt, orc, gotree, case):
    """For sampled leaves and every voxel of the padded brick: the oracle's loader returns the
    source values within half a quantisation step (plus float32 rounding), reports a voxel absent
    exactly when the smallest of its corners is positive (bit-pack and slice leaves) and returns
    the smallest corner for grid leaves."""
    c = dict(case)
    bs, pad = c.pop("bs"), c.pop("pad")
    nodes, values = gotree[(bs, pad)]
    enc = ce.Encoder(nodes, values, bs, pad, seed=7, **c)
    h, w = enc.encode()
    sc = coct_scene(lrt, h, w)
    o = orc.Scene.from_scene(sc)
    L = orc.load("exact")
    v = bs + 2 * pad + 1
    rng = np.random.default_rng(1)
    picks = rng.choice(len(enc.leaves), size=min(80, len(enc.leaves)), replace=False)
    max_val = 0xFFFFFFFF if c["bits"] == 32 else (1 << c["bits"]) - 1
    checked = 0
    for li in picks:
        gi, off, code, sz = enc.leaves[li]
        B = enc._brick(gi)
        lo, step = ce.leaf_range(h, w, c["leaf_type"], off, code, sz)
        assert step * max_val <= 1.0 + 1e-9
        for vx in range(-pad, bs + pad):
            for vy in range(-pad, bs + pad):
                for vz in range(-pad, bs + pad):
                    vals = np.zeros(8, np.float32)
                    r = L.oracle_coct_voxel(ctypes.byref(o.view), 0, c["leaf_type"], off, vx, vy, vz, float(sz), code,
                                            vals.ctypes.data)
                    src = np.array([B[vx + pad + (i >> 2 & 1), vy + pad + (i >> 1 & 1), vz + pad + (i & 1)]
                                    for i in range(8)], np.float64)
                    present = src.min() <= 0
                    if c["leaf_type"] == ce.GRID:
                        assert r == vals.min()
                    elif not present:
                        assert r == np.float32(1e6)
                        continue
                    else:
                        assert r == -1.0
                    # values beyond the stored range are clamped to its ends
                    want = np.clip(src, lo, lo + step * max_val)
                    tol = 0.5 * step + 4e-6 * max(1.0, float(np.abs(B).max())) + 1e-7
                    assert np.abs(vals - want).max() <= tol, (li, vx, vy, vz, vals, src)
                    # the float64 decoder of the layout description reads the same values
                    ref = ce.decode_voxel(h, w, sc.get_array("coct_mods").reshape(-1, 4), c["leaf_type"], off, code,
                                          (vx, vy, vz), sz)
                    assert ref is not None and np.abs(np.array(ref) - vals).max() <= 1e-5 * max(1, np.abs(ref).max())
                    checked += 1
    assert checked >= 40


def test_loader_planted_error_is_caught(lrt, orc, gotree):
    """Decoding a slice leaf with the value-lattice modifiers in place of the voxel-lattice ones
    (a plausible slip) breaks the presence check the test above makes."""
    nodes, values = gotree[(2, 1)]
    enc = ce.Encoder(nodes, values, 2, 1, leaf_type=ce.SLICES, rotate=True, pack_mode=1, seed=3)
    h, w = enc.encode()
    sc = coct_scene(lrt, h, w)
    mods = sc.get_array("coct_mods").reshape(3, 48, 4).copy()
    mods[1] = mods[0]   # planted: presence flags through the wrong lattice
    wrong = 0
    for gi, off, code, sz in enc.leaves[:30]:
        B = enc._brick(gi)
        for vx in range(-1, 3):
            for vy in range(-1, 3):
                for vz in range(-1, 3):
                    src = [B[vx + 1 + (i >> 2 & 1), vy + 1 + (i >> 1 & 1), vz + 1 + (i & 1)] for i in range(8)]
                    got = ce.decode_voxel(h, w, mods.reshape(-1, 4), ce.SLICES, off, code, (vx, vy, vz), sz)
                    wrong += (got is None) != (min(src) > 0)
    assert wrong > 0


# ---- the walk ----
def presets(lrt):
    out = []
    for mode in (0, 1, 2, 4):   # ST, analytic, Newton, interval
        p = lrt.default_preset()
        p["sdf_node_intersect"] = mode
        out.append(p)
    p = lrt.default_preset()
    p["render_mode"] = 20   # LOD: primId = log2 of the level's voxel count
    out.append(p)
    return out


def test_leaf_types_trace_identically(lrt, orc, gotree):
    """Grid, bit-pack and slice encodings of one octree hold the same quantised values and present
    voxels, so their hit records agree bit for bit except primId (the leaves' word offsets)."""
    nodes, values = gotree[(1, 1)]
    pos, dir_ = sphere_rays(3000)
    for p in presets(lrt):
        res = []
        for lt in (ce.GRID, ce.BIT_PACK, ce.SLICES):
            h, w = ce.encode(nodes, values, 1, 1, bits=10, leaf_type=lt)
            res.append(oracle_trace(orc, lrt, coct_scene(lrt, h, w, preset=p), pos, dir_))
        for r in res[1:]:
            assert np.array_equal(r["t"].view(np.uint32), res[0]["t"].view(np.uint32))
            assert np.array_equal(r["coords"].view(np.uint32), res[0]["coords"].view(np.uint32))
            assert np.array_equal(r["geomId"], res[0]["geomId"])
            if p["render_mode"] == 20:
                assert np.array_equal(r["primId"], res[0]["primId"])
        assert (res[0]["primId"] != MISS).mean() > 0.2


def test_agrees_with_frame_octree(lrt, orc, gotree):
    """Brick size 1, 32-bit values: the compact octree walk finds the hits of the frame octree of
    the same global octree (BLAS over its leaves): same hit set apart from near-ties, t within 5e-5
    for all but a few rays whose two formats' values round to different roots within one voxel."""
    nodes, values = gotree[(1, 0)]
    h, w = ce.encode(nodes, values, 1, 0, bits=32, leaf_type=ce.GRID)
    pos, dir_ = sphere_rays(5000)
    a = oracle_trace(orc, lrt, coct_scene(lrt, h, w), pos, dir_)
    fo = lrt.sdf_spheres_frame_octree(SPHERES, 5)

    def build(s):
        s.AddInstance(s.AddGeom_SdfFrameOctree(fo), np.eye(4, dtype=np.float32))
    b = oracle_trace(orc, lrt, host_scene(lrt, build), pos, dir_)
    ha, hb = a["primId"] != MISS, b["primId"] != MISS
    assert ha.mean() > 0.2
    assert (ha != hb).mean() < 2e-3
    both = ha & hb
    dt = np.abs(a["t"][both] - b["t"][both])
    assert np.mean(dt < 5e-5) > 0.995 and dt.max() < 2.0 / 32   # rounding of the two formats, rare near-ties
    # a planted error in the walk's mirroring would move hits: reflected rays do not agree
    assert not np.array_equal(ha, hb[::-1])


def test_agrees_with_brick_set(lrt, orc, gotree):
    """Brick size 2: the walk reaches the bricks the brick set of the same octree holds (its own
    BLAS): hit set and t agree up to the two formats' quantisation."""
    nodes, values = gotree[(2, 0)]
    h, w = ce.encode(nodes, values, 2, 0, bits=32, leaf_type=ce.GRID)
    pos, dir_ = sphere_rays(4000)
    a = oracle_trace(orc, lrt, coct_scene(lrt, h, w), pos, dir_)
    hdr = np.zeros((), lrt.SBS_HEADER_DTYPE)
    hdr["brick_size"], hdr["brick_pad"], hdr["bytes_per_value"], hdr["aux_data"] = 2, 0, 4, 1 << 24
    sn, sv = lrt.sdf_spheres_sbs(SPHERES, 4, hdr)

    def build(s):
        s.AddInstance(s.AddGeom_SdfSBS(hdr, sn, sv), np.eye(4, dtype=np.float32))
    b = oracle_trace(orc, lrt, host_scene(lrt, build), pos, dir_)
    ha, hb = a["primId"] != MISS, b["primId"] != MISS
    assert ha.mean() > 0.2
    assert (ha != hb).mean() < 3e-3
    both = ha & hb
    assert np.median(np.abs(a["t"][both] - b["t"][both])) < 1e-5


@pytest.mark.parametrize("pack_mode", [1, 2])
def test_rotation_and_shift_are_transparent(lrt, orc, gotree, pack_mode):
    """Leaves stored rotated and shifted (similarity compression) decode to the same field: hits
    agree with the plain encoding up to quantisation."""
    nodes, values = gotree[(2, 1)]
    pos, dir_ = sphere_rays(3000)
    h0, w0 = ce.encode(nodes, values, 2, 1, bits=16, leaf_type=ce.SLICES)
    h1, w1 = ce.encode(nodes, values, 2, 1, bits=16, leaf_type=ce.SLICES, pack_mode=pack_mode, rotate=True, sim=1,
                       seed=11)
    a = oracle_trace(orc, lrt, coct_scene(lrt, h0, w0), pos, dir_)
    b = oracle_trace(orc, lrt, coct_scene(lrt, h1, w1), pos, dir_)
    ha, hb = a["primId"] != MISS, b["primId"] != MISS
    assert ha.mean() > 0.2 and (ha != hb).mean() < 3e-3
    both = ha & hb
    assert np.median(np.abs(a["t"][both] - b["t"][both])) < 1e-4


def test_lod_leaves_equal_a_shallower_octree(lrt, orc, gotree, lrt_depth=3):
    """With LOD leaves and a fixed level of detail 2^L, the walk stops at level L: the hits equal
    (bit for bit, primId aside) those of the same octree built only L levels deep; with a fine
    level of detail it equals the octree without LOD leaves."""
    bs, pad = 1, 1
    deep = lrt.sdf_spheres_global_octree(SPHERES, 5, bs, pad)
    shallow = lrt.sdf_spheres_global_octree(SPHERES, lrt_depth, bs, pad)
    pos, dir_ = sphere_rays(3000)
    p = lrt.default_preset()
    p["fixed_lod"], p["level_of_detail"] = 1, float(lrt_depth)
    hl, wl = ce.encode(*deep, bs, pad, leaf_type=ce.GRID, lods=1, bits=16)
    hs, ws = ce.encode(*shallow, bs, pad, leaf_type=ce.GRID, lods=0, bits=16)
    a = oracle_trace(orc, lrt, coct_scene(lrt, hl, wl, preset=p), pos, dir_)
    b = oracle_trace(orc, lrt, coct_scene(lrt, hs, ws, preset=p), pos, dir_)
    assert (a["primId"] != MISS).mean() > 0.2
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
    assert np.array_equal(a["coords"].view(np.uint32), b["coords"].view(np.uint32))
    p["level_of_detail"] = 16.0
    hd, wd = ce.encode(*deep, bs, pad, leaf_type=ce.GRID, lods=0, bits=16)
    c = oracle_trace(orc, lrt, coct_scene(lrt, hl, wl, preset=p), pos, dir_)
    d = oracle_trace(orc, lrt, coct_scene(lrt, hd, wd, preset=p), pos, dir_)
    assert np.array_equal(c["t"].view(np.uint32), d["t"].view(np.uint32))
    assert not np.array_equal(a["t"].view(np.uint32), c["t"].view(np.uint32))


# ---- host checks ----
def test_host_checks_reject_malformed_octrees(lrt, gotree):
    nodes, values = gotree[(1, 0)]
    h, w = ce.encode(nodes, values, 1, 0, leaf_type=ce.BIT_PACK)
    coct_scene(lrt, h, w)   # well formed

    def rejects(hh, ww, level=0):
        with pytest.raises(lrt.LrtError):
            coct_scene(lrt, hh, ww, bvh_level=level)

    bad = w.copy()
    bad[1] = w.size + 5                       # first child entry beyond the words
    rejects(h, bad)
    rejects(h, w[: w.size - 3])               # last brick cut short
    hb = h.copy()
    hb["idx_mask"] = 0xFFFF                   # header words contradict the pack mode
    rejects(hb, w)
    hs = h.copy()
    hs["sim_compression"] = 1                 # an additive shift needs a pack mode with a shift field
    rejects(hs, w)
    nodes2, values2 = gotree[(2, 1)]
    h4, w4 = ce.encode(nodes2, values2, 2, 1, leaf_type=ce.SLICES)
    h4b = h4.copy()
    h4b["brick_size"] = 3                     # same words read as larger bricks: out of range or > 64 values
    rejects(h4b, w4)
    hl, wl = ce.encode(nodes, values, 1, 0, leaf_type=ce.GRID, lods=1)
    rejects(hl, wl, level=2)                  # LOD leaves with a deeper BLAS level
    cyc = w.copy()
    info = int(cyc[0])
    first_child = (info & -info).bit_length() - 1
    if ((info >> (8 + 2 * first_child)) & 3) == 0:
        cyc[1] = 0                            # a child pointing back at the root
        rejects(h, cyc)


def test_file_round_trip_of_encoded_octree(lrt, gotree, tmp_path):
    nodes, values = gotree[(2, 1)]
    h, w = ce.encode(nodes, values, 2, 1, leaf_type=ce.SLICES, pack_mode=1, rotate=True, sim=1, uv=1)
    p = tmp_path / "x.bin"
    lrt.save_coctree_v3(str(p), h, w)
    h2, w2, ver = lrt.load_coctree_v3(str(p))
    assert ver == 5 and raw_equal(h, h2) and np.array_equal(w, w2)
    with pytest.raises(lrt.LrtError):
        lrt.load_coctree_v3(str(tmp_path / "missing.bin"))
    trunc = tmp_path / "t.bin"
    trunc.write_bytes(p.read_bytes()[:-8])
    with pytest.raises(lrt.LrtError):
        lrt.load_coctree_v3(str(trunc))
