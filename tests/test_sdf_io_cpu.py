"""Reference-format SDF scene files: the loaders and savers of include/lrt_build.h on the data files
the reference itself ships (copied under tests/golden/ as fixtures):
  engine_sdf_frame_octree_0.bin  scenes/06_engine_scenes/framed_octree/sdf_frame_octree_0.bin
  engine_sdf_multi_octree_0.bin  scenes/06_engine_scenes/multi_octree/sdf_multi_octree_0.bin
  test_10_sbs.bin                saves/test_10_sbs.bin
The two engine files hold one 8,489-node octree (scene.xml beside each names num_primitives=8489)
as a frame octree and as a multi-octree, so per ray their hits must agree: a cross-check of the
oracle's two walks on real data. Reference: sdf/simple/sdf_simple.cpp:49-142,
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


def test_sbs_fixture_traces(lrt, orc):
    """The 49,487-brick DX set (brick 1, one byte per value): the oracle's march finds hits, and
    every hit point lies on the surface its brick's 8 decoded values describe: their trilinear,
    evaluated in float64 at the hit point, is zero to within 1e-4 of the brick's distance range
    (d_max = sqrt(3) * 2 / lod_size, sdf_converter.cpp:182-187; BVH2Common.cpp:846-882), or, for
    a hit on the brick's entry face, not above it."""
    g = lrt
    h, nodes, vals, vf = g.load_sdf_sbs(SBS)
    s = host_scene(g, lambda s: s.AddInstance(s.AddGeom_SdfSBS(h, nodes, vals)))
    pos, dirs = _rays(orc, g, 96)
    hits, _ = orc.trace_closest(orc.Scene.from_scene(s), pos, dirs)
    hit = hits["primId"] != 0xFFFFFFFF
    assert hit.sum() > 300
    p = pos[hit, :3].astype(np.float64) + hits["t"][hit, None].astype(np.float64) * dirs[hit, :3]
    nd = nodes[hits["primId"][hit]]
    lod = (nd["pos_z_lod_size"] & 0xFFFF).astype(np.float64)
    cell = np.stack([nd["pos_xy"] >> 16, nd["pos_xy"] & 0xFFFF, nd["pos_z_lod_size"] >> 16], 1).astype(np.float64)
    size = 2.0 / lod
    q = (p - (-1.0 + size[:, None] * cell)) / size[:, None]
    assert np.all(q > -1e-3) and np.all(q < 1 + 1e-3), "hit point outside its brick"
    q = np.clip(q, 0.0, 1.0)
    words = vals[nd["data_offset"][:, None] + np.arange(2)[None, :]].astype(np.uint64)
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
    # volume representation (the default preset): a ray that enters a brick already inside the
    # surface hits at the brick face, where the value is <= 0; elsewhere the hit is a root
    face = np.any((q < 1e-4) | (q > 1 - 1e-4), axis=1)
    assert np.all(rel[face] <= 1e-4)
    assert (~face).sum() > 100
    assert np.abs(rel[~face]).max() < 1e-4, np.abs(rel[~face]).max()
