"""SCom2 leaves on the CPU: the oracle's restatement of SCom2Intersect
(MultiRenderer/BVH2Common.cpp:2560-2818, with the unpacking of sdf/scom2/scom_shared.h:183-258)
pinned by the reference's own engine-scene SCom2 file (scenes/06_engine_scenes/scom_tree_2/
scom2_0.bin, copied under tests/golden/): it holds the object of the engine scene's DAG file
(same depth, one with bricks shared through rotations and shifts), so both trace alike, and a
planted rotation error breaks that. Host checks of AddGeom_SCom2 (BVH2Common_host.cpp:1133-1191)
and the file format (sdf/scom2/scom_utils.cpp:98-195)."""
import os

import numpy as np
import pytest

from conftest import MISS, host_scene
from test_coctree_cpu import oracle_trace, sphere_rays
from test_dag_cpu import dag_scene, no_eps

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCOM2 = os.path.join(GOLD, "engine_scom2_0.bin")


def scom2_scene(lrt, arrays, preset=None):
    def build(s):
        s.AddInstance(s.AddGeom_SCom2(*arrays))
        if preset is not None:
            s.SetPreset(preset)
    return host_scene(lrt, build)


def test_reference_file_layout(lrt, tmp_path):
    h, n, b = lrt.load_scom2(SCOM2)
    assert (int(h["brick_size"]), int(h["v_size"]), int(h["bits_per_value"]), int(h["has_multi_nodes"]),
            int(h["has_surfaces"]), int(h["has_channels"])) == (1, 2, 8, 1, 1, 0)
    assert n.size == 32510   # scene.xml: num_primitives
    out = tmp_path / "s.bin"
    lrt.save_scom2(str(out), h, n, b)
    assert open(out, "rb").read() == open(SCOM2, "rb").read()
    for bad in (b"\x00" * 8, open(SCOM2, "rb").read()[:-4]):
        p = tmp_path / "bad.bin"
        p.write_bytes(bad)
        with pytest.raises(lrt.LrtError):
            lrt.load_scom2(str(p))


def test_reference_file_matches_the_dag_file(lrt, orc):
    """The SCom2 tree and the DAG of the engine scene (both 6 levels, 8-bit values against floats)
    are hit by the same rays at t within the 8-bit quantisation; against the 5-level frame octree
    the differences are those of the coarser voxels."""
    pos, dir_ = sphere_rays(4000)
    a = oracle_trace(orc, lrt, scom2_scene(lrt, lrt.load_scom2(SCOM2), no_eps(lrt)), pos, dir_)
    b = oracle_trace(orc, lrt, dag_scene(lrt, lrt.load_sdf_dag(os.path.join(GOLD, "engine_sdf_dag_0.bin")),
                                         no_eps(lrt)), pos, dir_)
    ha, hb = a["primId"] != MISS, b["primId"] != MISS
    assert ha.mean() > 0.5 and (ha != hb).mean() < 2e-3
    dt = np.abs(a["t"][ha & hb] - b["t"][ha & hb])
    assert np.median(dt) < 1e-3 and np.percentile(dt, 99) < 5e-3
    assert np.all((a["geomId"][ha] >> 24) == 26)   # TYPE_SCOM2


def test_planted_rotation_error_is_caught(lrt, orc):
    """The tree shares bricks through rotations: decoding every brick through the next rotation
    of the table instead moves the hits away from the DAG's."""
    h, n, b = lrt.load_scom2(SCOM2)
    sc = scom2_scene(lrt, (h, n, b), no_eps(lrt))
    o = orc.Scene.from_scene(sc)
    mods = sc.get_array("scom2_mods").reshape(3, 48, 4).copy()
    mods[0] = np.roll(mods[0], 1, axis=0)
    bad = np.ascontiguousarray(mods.ravel())
    o._keep["scom2_mods"] = bad
    o.view.scom2_mods = bad.ctypes.data
    pos, dir_ = sphere_rays(4000)
    w = orc.trace_closest(o, pos, dir_)[0]
    good = oracle_trace(orc, lrt, sc, pos, dir_)
    hw, hg = w["primId"] != MISS, good["primId"] != MISS
    differ = (hw != hg).mean() + np.mean(np.abs(w["t"][hw & hg] - good["t"][hw & hg]) > 1e-3)
    assert differ > 0.02


def test_host_checks_reject_malformed_trees(lrt):
    h, n, b = lrt.load_scom2(SCOM2)
    scom2_scene(lrt, (h, n, b))

    def rejects(arrays):
        with pytest.raises(lrt.LrtError):
            scom2_scene(lrt, arrays)

    rejects((h, n, b[: b.size // 2]))                       # bricks cut short
    h2 = h.copy()
    h2["root_node_off"] = n.size + b.size + 10              # root beyond the words
    rejects((h2, n, b))
    h3 = h.copy()
    h3["values_per_uint"], h3["bits_per_value"] = 8, 8      # 64 bits per word
    rejects((h3, n, b))
    h4 = h.copy()
    h4["dimension"] = 2
    rejects((h4, n, b))
    # a child edge pointing back at the root: walk the root's first inner child link
    root = int(h["root_node_off"])
    n5 = n.copy()
    lo, upl, coo = int(h["links_offset"]), int(h["uints_per_link"]), int(h["child_offset_off"])
    node0 = int(n[root])
    types = (node0 >> int(h["children_types_shift"])) & int(h["children_types_mask"])
    base = (node0 >> int(h["base_reference_shift"])) & int(h["reference_mask"])
    for c in range(8):
        if ((types >> (2 * c)) & 3) == 2:
            ref = int(n[root + int(h["references_offset"])])
            nodeL = int(n[root + lo - 1])
            sh = (c - 1) * int(h["reference_bits"])
            start = base if c == 0 else (((nodeL >> (sh % 32)) if sh >= 32 else (ref >> sh)) & int(h["reference_mask"]))
            n5[root + lo + start * upl + coo] = (n5[root + lo + start * upl + coo] & ~np.uint32(h["node_offset_mask"])) | root
            rejects((h, n5, b))
            break


# ---- plain trees: the fast traversal ----
@pytest.fixture(scope="module")
def plain(lrt):
    import scom2_encode as se
    from test_coctree_cpu import SPHERES
    nodes, values = lrt.sdf_spheres_global_octree(SPHERES, 4, 2, 0)
    return {"go": (nodes, values),
            "plain": se.encode(nodes, values),
            "rotated": se.encode(nodes, values, rotate=True, seed=3),
            "mixed": se.encode(nodes, values, rotate=True, shift=True, unique_every=5, volume_every=7, seed=4)}


def lambert(lrt, eps=0.0, mode=1):
    p = lrt.default_preset()
    p["render_mode"], p["compact_sdf_eps"], p["sdf_node_intersect"] = 1, eps, mode   # LAMBERT_NO_TEX: fast walk
    return p


def normal_mode(lrt, eps=0.0, mode=1):
    p = lrt.default_preset()
    p["render_mode"], p["compact_sdf_eps"], p["sdf_node_intersect"] = 8, eps, mode   # NORMAL: the general walk
    return p


def test_plain_tree_rotations_are_transparent(lrt, orc, plain):
    """Subtrees and bricks stored through random rotations (composed down the tree) decode to the
    same values: both walks give bit-identical records for the rotated and the plain encoding."""
    pos, dir_ = sphere_rays(3000)
    for preset in (lambert(lrt), normal_mode(lrt), lambert(lrt, 0.05, 3), normal_mode(lrt, 0.05, 0)):
        a = oracle_trace(orc, lrt, scom2_scene(lrt, plain["plain"], preset), pos, dir_)
        b = oracle_trace(orc, lrt, scom2_scene(lrt, plain["rotated"], preset), pos, dir_)
        assert (a["primId"] != MISS).mean() > 0.2
        assert np.array_equal(np.delete(a.view(np.uint32).reshape(-1, 8), 1, 1),
                              np.delete(b.view(np.uint32).reshape(-1, 8), 1, 1))   # primId: link words differ


def test_fast_walk_agrees_with_general_walk(lrt, orc, plain):
    """The fast traversal (LAMBERT_NO_TEX) and the general walk (other render modes) find the same
    hits on plain trees, at t within rounding of their different voxel set-ups, with and without
    shared, shifted, unique and volume bricks."""
    pos, dir_ = sphere_rays(3000)
    for key in ("plain", "mixed"):
        a = oracle_trace(orc, lrt, scom2_scene(lrt, plain[key], lambert(lrt)), pos, dir_)
        b = oracle_trace(orc, lrt, scom2_scene(lrt, plain[key], normal_mode(lrt)), pos, dir_)
        ha, hb = a["primId"] != MISS, b["primId"] != MISS
        assert ha.mean() > 0.2 and (ha != hb).mean() < 2e-3
        dt = np.abs(a["t"][ha & hb] - b["t"][ha & hb])
        assert np.percentile(dt, 99) < 1e-4


def test_plain_tree_matches_brick_set(lrt, orc, plain):
    """A plain tree of the sphere-union octree (brick size 2, 8-bit values) is hit like the brick
    set of the same octree."""
    pos, dir_ = sphere_rays(3000)
    a = oracle_trace(orc, lrt, scom2_scene(lrt, plain["plain"], lambert(lrt)), pos, dir_)
    hdr = np.zeros((), lrt.SBS_HEADER_DTYPE)
    hdr["brick_size"], hdr["brick_pad"], hdr["bytes_per_value"], hdr["aux_data"] = 2, 0, 4, 1 << 24
    from test_coctree_cpu import SPHERES
    sn, sv = lrt.sdf_spheres_sbs(SPHERES, 4, hdr)

    def build(s):
        s.AddInstance(s.AddGeom_SdfSBS(hdr, sn, sv))
        s.SetPreset(lambert(lrt))
    b = oracle_trace(orc, lrt, host_scene(lrt, build), pos, dir_)
    ha, hb = a["primId"] != MISS, b["primId"] != MISS
    assert ha.mean() > 0.2 and (ha != hb).mean() < 5e-3
    assert np.median(np.abs(a["t"][ha & hb] - b["t"][ha & hb])) < 2e-3


def test_plain_tree_host_checks(lrt, plain):
    h, n, b = plain["plain"]
    h2 = h.copy()
    h2["brick_size"] = 1                      # 3^3 values but brick size 1
    with pytest.raises(lrt.LrtError):
        scom2_scene(lrt, (h2, n, b))
    n2 = n.copy()
    n2[1] = (n2[1] & np.uint32(0xFC000000)) | np.uint32(n.size + 100)   # first link points past the words
    with pytest.raises(lrt.LrtError):
        scom2_scene(lrt, (h, n2, b))


# ---- multi-node trees with channels ----
def _channel_quotients(h, hit):
    c0, c1 = h["coords"][hit, 0].view(np.uint32), h["coords"][hit, 1].view(np.uint32)
    return np.stack([c0 & 0xFFFF, c0 >> 16, c1], 1).astype(np.int64)


@pytest.mark.parametrize("bs", [1, 2])
def test_multi_node_channels_turn_the_quotients(lrt, orc, bs):
    """Multi-node trees (the reference file's field positions) whose leaves carry a channel link:
    hits name the channel link as primId, and the channel render modes turn the voxel-local hit
    point by the composed rotation rot_add[node rotation * 48 + channel rotation]
    (BVH2Common.cpp:2731-2747) with the transforms of the brick_size + 1 lattice. Pinned against
    the same tree stored without rotations (identity: the point's offset in its voxel) and the
    axis maps of the subgroup generator restated in numpy; a composition in the other order would
    turn many hits differently."""
    import dag_encode as de
    import scom2_encode as se
    from test_coctree_cpu import SPHERES
    from test_dag_cpu import _axis_maps
    nodes, values = lrt.sdf_spheres_global_octree(SPHERES, 5 if bs == 1 else 4, bs, 0)
    plain_enc = se.Scom2Encoder(nodes, values, brick_size=bs, channels=True, seed=5)
    plain = plain_enc.encode()
    rot_enc = se.Scom2Encoder(nodes, values, brick_size=bs, channels=True, rotate=True, channel_rotate=True, seed=5)
    rotated = rot_enc.encode()
    assert int(plain[0]["has_multi_nodes"]) == 1 and int(plain[0]["has_channels"]) == 1
    pos, dir_ = sphere_rays(2500)
    p = no_eps(lrt)
    p["render_mode"] = 23   # MULTI_RENDER_MODE_CHANNEL
    a = oracle_trace(orc, lrt, scom2_scene(lrt, plain, p), pos, dir_)
    b = oracle_trace(orc, lrt, scom2_scene(lrt, rotated, p), pos, dir_)
    hit = a["primId"] != MISS
    assert hit.mean() > 0.2 and np.array_equal(hit, b["primId"] != MISS)
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
    assert a["primId"][hit].max() < len(plain_enc.channel_info)
    # identity: the hit point's offset in its voxel (a point on a voxel face may sit in either)
    q0 = _channel_quotients(a, hit).astype(np.float64) / 65535.0
    pt = pos[hit, :3].astype(np.float64) + a["t"][hit, None].astype(np.float64) * dir_[hit, :3]
    cells = 2 ** (5 if bs == 1 else 4) * bs
    frac = (pt + 1.0) * cells / 2.0
    frac -= np.floor(frac)
    err = np.minimum(np.abs(frac - q0), 1.0 - np.abs(frac - q0))
    assert np.percentile(err, 99) < 1e-3
    # rotated: the composed rotation's axis map
    q0, qr = _channel_quotients(a, hit), _channel_quotients(b, hit)
    table = de.rot_add_reference(2).reshape(48, 48)
    maps = _axis_maps(bs + 1)
    n_turned = 0
    for j, ch in enumerate(b["primId"][hit]):
        q, r = rot_enc.channel_info[int(ch)]
        crot = int(table[q, r])
        for i, (o, sgn) in enumerate(maps[crot]):
            if sgn > 0:
                assert qr[j, o] == q0[j, i]
            elif bs == 1:
                assert abs(qr[j, o] - (65535 - q0[j, i])) <= 1
            else:
                assert qr[j, o] == 65535
        n_turned += crot != 0
    assert n_turned > hit.sum() // 2
    # other render modes keep the plain records (primId: the channel link)
    p["render_mode"] = 8
    c = oracle_trace(orc, lrt, scom2_scene(lrt, rotated, p), pos, dir_)
    assert np.array_equal(c["primId"][hit], b["primId"][hit]) and not np.any(c["coords"][hit, :2])
