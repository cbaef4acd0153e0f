"""Reference-format SDF scene files: the loaders and savers of include/lrt_build.h on the data files
the reference itself ships (copied under tests/golden/ as fixtures):
  engine_sdf_frame_octree_0.bin  scenes/06_engine_scenes/framed_octree/sdf_frame_octree_0.bin
  engine_sdf_multi_octree_0.bin  scenes/06_engine_scenes/multi_octree/sdf_multi_octree_0.bin
  test_10_sbs.bin                saves/test_10_sbs.bin
  test_10_svs.bin                saves/test_10_svs.bin
The two engine files hold one 8,489-node octree (scene.xml beside each names num_primitives=8489)
as a frame octree and as a multi-octree, so per ray their hits must agree: a cross-check of the
oracle's two walks on real data. The two test_10 files hold one adaptive voxel set as a DX brick
set (brick 1) and as a sparse voxel set, quantised apart by at most one byte step per value.
Reference: sdf/simple/sdf_simple.cpp:49-142,
sdf/multi/sdf_multi.cpp:8-35.
"""
import os

import numpy as np
import pytest

from conftest import host_scene

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FRAME = os.path.join(GOLD, "engine_sdf_frame_octree_0.bin")
MULTI = os.path.join(GOLD, "engine_sdf_multi_octree_0.bin")
SBS = os.path.join(GOLD, "test_10_sbs.bin")
SVS = os.path.join(GOLD, "test_10_svs.bin")


def test_fixture_counts(lrt):
    g = lrt
    fn = g.load_sdf_frame_octree(FRAME)
    assert fn.size == 8489
    mn, mv = g.load_sdf_multi_octree(MULTI)
    assert mn.size == 8489 and mv.size == 27776
    # the values are the global octree's values_f (global_octree_to_multi_octree,
    # sdf_multi.cpp:124-180): every node with voxels reads 8 values per voxel inside the array
    vox = mn["voxel_count"] > 0
    assert vox.sum() > 1000
    assert np.all(mn["data_offset"][vox].astype(np.int64) + 8 * mn["voxel_count"][vox] <= mv.size)
    h, nodes, vals, vf = g.load_sdf_sbs(SBS)
    assert (int(h["brick_size"]), int(h["brick_pad"]), int(h["bytes_per_value"]), int(h["aux_data"])) == \
        (1, 0, 1, g.SDF_SBS_NODE_LAYOUT_DX)
    assert nodes.size == 49487 and vf.size == 0
    # brick 1, pad 0: 2^3 one-byte values per brick = 2 words, data offsets in brick order
    assert vals.size == 2 * nodes.size
    assert np.array_equal(nodes["data_offset"], 2 * np.arange(nodes.size, dtype=np.uint32))
    # an adaptive set: bricks of several levels (lod sizes 16..128), each inside its level's grid
    lod = nodes["pos_z_lod_size"] & 0xFFFF
    assert set(np.unique(lod).tolist()) == {16, 32, 64, 128}
    px, py, pz = nodes["pos_xy"] >> 16, nodes["pos_xy"] & 0xFFFF, nodes["pos_z_lod_size"] >> 16
    assert np.all(px < lod) and np.all(py < lod) and np.all(pz < lod)
    sv = g.load_sdf_svs(SVS)
    assert sv.size == 49487
    # the same voxels as the brick set (other order), values within one quantisation step
    k_svs = (sv["pos_xy"].astype(np.uint64) << np.uint64(32)) | sv["pos_z_lod_size"]
    k_sbs = (nodes["pos_xy"].astype(np.uint64) << np.uint64(32)) | nodes["pos_z_lod_size"]
    i1, i2 = np.argsort(k_svs), np.argsort(k_sbs)
    assert np.array_equal(k_svs[i1], k_sbs[i2])
    b1 = sv["values"][i1].view(np.uint8).reshape(-1, 8).astype(int)
    b2 = vals.reshape(-1, 2)[i2].view(np.uint8).reshape(-1, 8).astype(int)
    assert np.abs(b1 - b2).max() <= 1 and (b1 != b2).any()


def test_round_trip_is_byte_identical(lrt, tmp_path):
    g = lrt
    fn = g.load_sdf_frame_octree(FRAME)
    g.save_sdf_frame_octree(str(tmp_path / "f.bin"), fn)
    assert (tmp_path / "f.bin").read_bytes() == open(FRAME, "rb").read()
    mn, mv = g.load_sdf_multi_octree(MULTI)
    g.save_sdf_multi_octree(str(tmp_path / "m.bin"), mn, mv)
    assert (tmp_path / "m.bin").read_bytes() == open(MULTI, "rb").read()
    h, nodes, vals, vf = g.load_sdf_sbs(SBS)
    g.save_sdf_sbs(str(tmp_path / "s.bin"), h, nodes, vals, vf)
    # the saver always writes the fp32 block (here its zero count) after the old format's end
    assert (tmp_path / "s.bin").read_bytes() == open(SBS, "rb").read() + b"\0\0\0\0"
    h2, n2, v2, f2 = g.load_sdf_sbs(str(tmp_path / "s.bin"))
    assert h2.tobytes() == h.tobytes() and np.array_equal(n2, nodes) and np.array_equal(v2, vals) and f2.size == 0
    sv = g.load_sdf_svs(SVS)
    g.save_sdf_svs(str(tmp_path / "v.bin"), sv)
    assert (tmp_path / "v.bin").read_bytes() == open(SVS, "rb").read()


def test_truncated_and_missing_files_are_errors(lrt, tmp_path):
    g = lrt
    raw = open(MULTI, "rb").read()
    (tmp_path / "cut.bin").write_bytes(raw[:-5])
    with pytest.raises(g.LrtError) as e:
        g.load_sdf_multi_octree(str(tmp_path / "cut.bin"))
    assert e.value.code == 1 and "file size" in str(e.value)
    (tmp_path / "cut2.bin").write_bytes(open(FRAME, "rb").read()[:1000])
    with pytest.raises(g.LrtError) as e:
        g.load_sdf_frame_octree(str(tmp_path / "cut2.bin"))
    assert "exceeds the file" in str(e.value)
    raw = open(SBS, "rb").read()
    (tmp_path / "cut3.bin").write_bytes(raw[:30])
    with pytest.raises(g.LrtError):
        g.load_sdf_sbs(str(tmp_path / "cut3.bin"))
    (tmp_path / "cut4.bin").write_bytes(open(SVS, "rb").read()[:-16])
    with pytest.raises(g.LrtError):
        g.load_sdf_svs(str(tmp_path / "cut4.bin"))
    with pytest.raises(g.LrtError) as e:
        g.load_sdf_frame_octree(str(tmp_path / "none.bin"))
    assert "cannot open" in str(e.value)


def _rays(orc, g, res=128):
    pi, vi = g.camera_inverse([0.4, 0.6, 3.0], [0, 0, 0], [0, 1, 0], 60.0, 1.0, 0.01, 100.0)
    jit = np.random.default_rng(19).random((res * res, 2), dtype=np.float32)
    return orc.eye_rays(pi, vi, res, res, jitter=jit)


def test_engine_frame_and_multi_octree_agree(lrt, orc):
    """The oracle's frame-octree walk (BVH over leaf cells) and its multi-octree walk (octree
    descent) over the reference's two encodings of one octree, on 512x512 jittered rays: the same
    leaf on every hit, the same hit/miss, t within the bar of conftest.format_cross_check."""
    g = lrt
    from conftest import format_cross_check
    fs = host_scene(g, lambda s: s.AddInstance(s.AddGeom_SdfFrameOctree(g.load_sdf_frame_octree(FRAME))))
    mn, mv = g.load_sdf_multi_octree(MULTI)
    ms = host_scene(g, lambda s: s.AddInstance(s.AddGeom_SdfMultiOctree(mn, mv)))
    pos, dirs = _rays(orc, g, 512)
    hf, _ = orc.trace_closest(orc.Scene.from_scene(fs), pos, dirs)
    hm, _ = orc.trace_closest(orc.Scene.from_scene(ms), pos, dirs)
    ob = fs.get_array("orig_nodes")
    cell = float((ob["boxMax"] - ob["boxMin"])[:, 0].min())
    r = format_cross_check(hf, hm, pos, dirs, cell)
    assert r["hits"] > 0.1 * pos.shape[0]


def _on_voxel_surface(hits, pos, dirs, cells, words):
    """Every hit lies on the surface its voxel's 8 decoded byte values describe: their trilinear,
    evaluated in float64 at the hit point, is zero to within 1e-4 of the voxel's distance range
    (d_max = sqrt(3) * 2 / lod_size, sdf_converter.cpp:182-187; BVH2Common.cpp:846-882 and
    :190-215), or, for a hit on the voxel's entry face, not above it. cells: (pos_xy,
    pos_z_lod_size) of each hit's voxel; words: its two value words."""
    hit = hits["primId"] != 0xFFFFFFFF
    p = pos[hit, :3].astype(np.float64) + hits["t"][hit, None].astype(np.float64) * dirs[hit, :3]
    pxy, pzl = cells
    lod = (pzl & 0xFFFF).astype(np.float64)
    cell = np.stack([pxy >> 16, pxy & 0xFFFF, pzl >> 16], 1).astype(np.float64)
    size = 2.0 / lod
    q = (p - (-1.0 + size[:, None] * cell)) / size[:, None]
    assert np.all(q > -1e-3) and np.all(q < 1 + 1e-3), "hit point outside its voxel"
    q = np.clip(q, 0.0, 1.0)
    words = words.astype(np.uint64)
    raw = np.stack([(words[:, i // 4] >> np.uint64(8 * (i % 4))) & np.uint64(255) for i in range(8)], 1)
    d_max = np.sqrt(3.0) * size
    v = -d_max[:, None] + (2 * d_max / 255.0)[:, None] * raw.astype(np.float64)
    f = np.zeros(q.shape[0])
    for i in range(8):
        wx = q[:, 0] if i & 4 else 1 - q[:, 0]
        wy = q[:, 1] if i & 2 else 1 - q[:, 1]
        wz = q[:, 2] if i & 1 else 1 - q[:, 2]
        f += wx * wy * wz * v[:, i]
    rel = f / d_max
    # volume representation (the default preset): a ray that enters a voxel already inside the
    # surface hits at the voxel face, where the value is <= 0; elsewhere the hit is a root
    face = np.any((q < 1e-4) | (q > 1 - 1e-4), axis=1)
    assert np.all(rel[face] <= 1e-4)
    assert (~face).sum() > 100
    assert np.abs(rel[~face]).max() < 1e-4, np.abs(rel[~face]).max()


def test_sbs_fixture_traces(lrt, orc):
    """The 49,487-brick DX set (brick 1, one byte per value): the oracle's march finds hits on the
    surface of the brick they name."""
    g = lrt
    h, nodes, vals, vf = g.load_sdf_sbs(SBS)
    s = host_scene(g, lambda s: s.AddInstance(s.AddGeom_SdfSBS(h, nodes, vals)))
    pos, dirs = _rays(orc, g, 96)
    hits, _ = orc.trace_closest(orc.Scene.from_scene(s), pos, dirs)
    hit = hits["primId"] != 0xFFFFFFFF
    assert hit.sum() > 300
    nd = nodes[hits["primId"][hit]]
    words = vals[nd["data_offset"][:, None] + np.arange(2)[None, :]]
    _on_voxel_surface(hits, pos, dirs, (nd["pos_xy"], nd["pos_z_lod_size"]), words)


def test_svs_fixture_traces(lrt, orc):
    """The 49,487-voxel sparse voxel set (BVHRT::AddGeom_SdfSVS, BVH2Common_host.cpp:536-597):
    one BLAS leaf per voxel, primId = the voxel's index; hits on the surface of the voxel they name."""
    g = lrt
    sv = g.load_sdf_svs(SVS)
    s = host_scene(g, lambda s: s.AddInstance(s.AddGeom_SdfSVS(sv)))
    pos, dirs = _rays(orc, g, 96)
    hits, _ = orc.trace_closest(orc.Scene.from_scene(s), pos, dirs)
    hit = hits["primId"] != 0xFFFFFFFF
    assert hit.sum() > 300
    nd = sv[hits["primId"][hit]]
    _on_voxel_surface(hits, pos, dirs, (nd["pos_xy"], nd["pos_z_lod_size"]), nd["values"])


def test_svs_and_sbs_fixtures_agree(lrt, orc):
    """The oracle's per-voxel leaf (sparse voxel set) and its brick march (brick set of 1-voxel
    bricks) over the test_10 pair: the two files quantise the same distances apart by one byte
    step, so surfaces move by up to 2 d_max / 255 and silhouettes may shift: hit/miss agree on
    all but 0.05% of rays, the same voxel is hit on at least 97% of common hits, and 99% of
    their t agree within a quarter of the finest voxel."""
    g = lrt
    sv = g.load_sdf_svs(SVS)
    h, nodes, vals, vf = g.load_sdf_sbs(SBS)
    s1 = host_scene(g, lambda s: s.AddInstance(s.AddGeom_SdfSVS(sv)))
    s2 = host_scene(g, lambda s: s.AddInstance(s.AddGeom_SdfSBS(h, nodes, vals)))
    pos, dirs = _rays(orc, g, 256)
    a, _ = orc.trace_closest(orc.Scene.from_scene(s1), pos, dirs)
    b, _ = orc.trace_closest(orc.Scene.from_scene(s2), pos, dirs)
    ha, hb = a["primId"] != 0xFFFFFFFF, b["primId"] != 0xFFFFFFFF
    assert ha.sum() > 2000
    assert (ha != hb).sum() <= 5e-4 * pos.shape[0]
    both = ha & hb
    k_svs = (sv["pos_xy"].astype(np.uint64) << np.uint64(32)) | sv["pos_z_lod_size"]
    k_sbs = (nodes["pos_xy"].astype(np.uint64) << np.uint64(32)) | nodes["pos_z_lod_size"]
    assert (k_svs[a["primId"][both]] == k_sbs[b["primId"][both]]).mean() >= 0.97
    dt = np.abs(a["t"][both].astype(np.float64) - b["t"][both])
    assert np.quantile(dt, 0.99) <= 0.25 * 2.0 / 128
