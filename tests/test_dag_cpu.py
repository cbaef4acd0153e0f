"""Octree DAG leaves on the CPU: the oracle's restatement of DAGIntersect
(MultiRenderer/BVH2Common.cpp:2365-2557) pinned by the reference's own engine-scene DAG file
(scenes/06_engine_scenes/sdf_dag/sdf_dag_0.bin, copied under tests/golden/) against the frame
octree of the same object, by DAGs encoded from sphere-union octrees against their frame octree
and brick set, by rotation invariance, and the rotation composition table against the
reference's own subgroup generator (restated in numpy). Host checks of AddGeom_SdfDAG
(BVH2Common_host.cpp:1020-1131) and the DAG file format (sdf/scom2/scom_utils.cpp:5-73)."""
import os

import numpy as np
import pytest

import coct_encode as ce
import dag_encode as de
from conftest import MISS, host_scene
from test_coctree_cpu import SPHERES, oracle_trace, sphere_rays

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DAG = os.path.join(GOLD, "engine_sdf_dag_0.bin")


def dag_scene(lrt, arrays, preset=None):
    def build(s):
        s.AddInstance(s.AddGeom_SdfDAG(*arrays))
        if preset is not None:
            s.SetPreset(preset)
    return host_scene(lrt, build)


def no_eps(lrt):
    p = lrt.default_preset()
    p["compact_sdf_eps"] = 0.0
    p["sdf_node_intersect"] = 0
    return p


def test_rot_add_table_is_the_reference_subgroup(lrt):
    """The library's composition table equals the Cayley table of the reference's default transform
    subgroup built by its own generator (axis permutations x signs, legacy remap) for 2^3 and 3^3
    lattices; the subgroup's maps are the legacy rotations' index maps (the modifiers)."""
    arrays = lrt.load_sdf_dag(DAG)
    sc = dag_scene(lrt, arrays)
    table = sc.get_array("rot_add")
    assert np.array_equal(table, de.rot_add_reference(2))
    assert np.array_equal(table, de.rot_add_reference(3))
    for v in (2, 3):
        maps = de.universal_transpositions(v)
        m = ce._closed_mods(v)
        pts = [(x, y, z) for x in range(v) for y in range(v) for z in range(v)]
        for r in range(48):
            assert maps[r] == [int(m[r, 0] * x + m[r, 1] * y + m[r, 2] * z + m[r, 3]) for x, y, z in pts]
    # a composition read the other way round (cayley[i][j]) is a different table
    assert not np.array_equal(table.reshape(48, 48), table.reshape(48, 48).T)


def test_reference_file_layout(lrt, tmp_path):
    h, n, c, d, dist = lrt.load_sdf_dag(DAG)
    assert (int(h["brick_size"]), int(h["brick_pad"]), int(h["node_grid_size"]), int(h["dim"])) == (1, 0, 2, 3)
    assert (n.size, c.size, d.size, dist.size) == (18855, 36417, 14304, 114424)
    assert n.size == 18855   # scene.xml: num_primitives
    out = tmp_path / "d.bin"
    lrt.save_sdf_dag(str(out), h, n, c, d, dist)
    assert open(out, "rb").read() == open(DAG, "rb").read()
    # the file is an unshared octree: no rotations, one 8-value brick per data edge
    assert not np.any(d["rotation_id"]) and not np.any(c["rotation_id"]) and dist.size == 8 * (d.size - 1)


def test_reference_file_matches_its_frame_octree(lrt, orc):
    """The engine scene's DAG file and its frame-octree file hold one object, the DAG one level
    finer (6 levels against 5): with the compact surface offset switched off they are hit by the
    same rays, at t that differ by the two resolutions' trilinear approximations (a few
    thousandths, a tenth of the coarser voxel)."""
    arrays = lrt.load_sdf_dag(DAG)
    pos, dir_ = sphere_rays(4000)
    a = oracle_trace(orc, lrt, dag_scene(lrt, arrays, no_eps(lrt)), pos, dir_)
    fo = lrt.load_sdf_frame_octree(os.path.join(GOLD, "engine_sdf_frame_octree_0.bin"))

    def build(s):
        s.AddInstance(s.AddGeom_SdfFrameOctree(fo))
        s.SetPreset(no_eps(lrt))
    b = oracle_trace(orc, lrt, host_scene(lrt, build), pos, dir_)
    ha, hb = a["primId"] != MISS, b["primId"] != MISS
    assert ha.mean() > 0.5 and (ha != hb).mean() < 5e-3
    dt = np.abs(a["t"][ha & hb] - b["t"][ha & hb])
    assert np.median(dt) < 4e-3 and np.percentile(dt, 99) < 0.5 * 2 / 32   # half the coarser voxel


@pytest.fixture(scope="module")
def gotrees(lrt):
    return {bs: lrt.sdf_spheres_global_octree(SPHERES, 5 if bs == 1 else 4, bs, 0) for bs in (1, 2)}


def test_encoded_dag_matches_frame_octree(lrt, orc, gotrees):
    """Brick size 1: a DAG of the sphere-union octree (no sharing, float distances) is hit like the
    frame octree of the same octree."""
    pos, dir_ = sphere_rays(4000)
    a = oracle_trace(orc, lrt, dag_scene(lrt, de.encode(*gotrees[1], 1), no_eps(lrt)), pos, dir_)
    fo = lrt.sdf_spheres_frame_octree(SPHERES, 5)

    def build(s):
        s.AddInstance(s.AddGeom_SdfFrameOctree(fo))
        s.SetPreset(no_eps(lrt))
    b = oracle_trace(orc, lrt, host_scene(lrt, build), pos, dir_)
    ha, hb = a["primId"] != MISS, b["primId"] != MISS
    assert ha.mean() > 0.2 and (ha != hb).mean() < 2e-3
    dt = np.abs(a["t"][ha & hb] - b["t"][ha & hb])
    assert np.mean(dt < 5e-5) > 0.99


def test_encoded_dag_matches_brick_set(lrt, orc, gotrees):
    pos, dir_ = sphere_rays(4000)
    a = oracle_trace(orc, lrt, dag_scene(lrt, de.encode(*gotrees[2], 2), no_eps(lrt)), pos, dir_)
    hdr = np.zeros((), lrt.SBS_HEADER_DTYPE)
    hdr["brick_size"], hdr["brick_pad"], hdr["bytes_per_value"], hdr["aux_data"] = 2, 0, 4, 1 << 24
    sn, sv = lrt.sdf_spheres_sbs(SPHERES, 4, hdr)

    def build(s):
        s.AddInstance(s.AddGeom_SdfSBS(hdr, sn, sv))
        s.SetPreset(no_eps(lrt))
    b = oracle_trace(orc, lrt, host_scene(lrt, build), pos, dir_)
    ha, hb = a["primId"] != MISS, b["primId"] != MISS
    assert ha.mean() > 0.2 and (ha != hb).mean() < 3e-3
    assert np.median(np.abs(a["t"][ha & hb] - b["t"][ha & hb])) < 1e-5


@pytest.mark.parametrize("surface", [False, True])
def test_rotated_bricks_are_transparent(lrt, orc, gotrees, surface):
    """Bricks stored through random rotations decode to the same values: hit records equal those of
    the unrotated DAG bit for bit, in every intersection mode; additive shifts change them only by
    rounding."""
    pos, dir_ = sphere_rays(3000)
    plain = de.encode(*gotrees[2], 2, surface=surface)
    rot = de.encode(*gotrees[2], 2, rotate=True, surface=surface, seed=4)
    for mode in (0, 2, 3, 4):
        p = lrt.default_preset()
        p["sdf_node_intersect"] = mode
        a = oracle_trace(orc, lrt, dag_scene(lrt, plain, p), pos, dir_)
        b = oracle_trace(orc, lrt, dag_scene(lrt, rot, p), pos, dir_)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
        assert (a["primId"] != MISS).mean() > 0.2
    shifted = de.encode(*gotrees[2], 2, rotate=True, add=True, surface=surface, seed=5)
    c = oracle_trace(orc, lrt, dag_scene(lrt, shifted, no_eps(lrt)), pos, dir_)
    a = oracle_trace(orc, lrt, dag_scene(lrt, plain, no_eps(lrt)), pos, dir_)
    ha, hc = a["primId"] != MISS, c["primId"] != MISS
    assert (ha != hc).mean() < 2e-3
    assert np.median(np.abs(a["t"][ha & hc] - c["t"][ha & hc])) < 1e-5


def test_lod_nodes_and_surface_eps(lrt, orc, gotrees):
    """Inner nodes with their own bricks become leaves past the level of detail: at LOD 2^L the
    hits equal those of the DAG of the octree built L deep; the compact surface offset moves hits
    outwards along the ray."""
    deep = de.encode(*gotrees[1], 1, lod=True)
    shallow = de.encode(*lrt_shallow(lrt), 1)
    pos, dir_ = sphere_rays(3000)
    p = no_eps(lrt)
    p["fixed_lod"], p["level_of_detail"] = 1, 3.0
    a = oracle_trace(orc, lrt, dag_scene(lrt, deep, p), pos, dir_)
    b = oracle_trace(orc, lrt, dag_scene(lrt, shallow, p), pos, dir_)
    assert (a["primId"] != MISS).mean() > 0.2
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
    e0 = oracle_trace(orc, lrt, dag_scene(lrt, deep, no_eps(lrt)), pos, dir_)
    p2 = no_eps(lrt)
    p2["compact_sdf_eps"] = 0.1
    e1 = oracle_trace(orc, lrt, dag_scene(lrt, deep, p2), pos, dir_)
    both = (e0["primId"] != MISS) & (e1["primId"] != MISS)
    assert np.median(e1["t"][both] - e0["t"][both]) < 0


def lrt_shallow(lrt):
    return lrt.sdf_spheres_global_octree(SPHERES, 3, 1, 0)


def test_host_checks_reject_malformed_dags(lrt, gotrees):
    good = de.encode(*gotrees[1], 1)
    dag_scene(lrt, good)

    def rejects(arrays):
        with pytest.raises(lrt.LrtError):
            dag_scene(lrt, arrays)

    h, n, c, d, dist = (a.copy() for a in good)
    c2 = c.copy()
    c2["child_offset"][np.flatnonzero(c2["child_offset"])[0]] = n.size + 3     # child beyond the nodes
    rejects((h, n, c2, d, dist))
    d2 = d.copy()
    d2["data_offset"][-1] = dist.size                                         # brick past the distances
    rejects((h, n, c, d2, dist))
    d3 = d.copy()
    d3["rotation_id"][1] = 48
    rejects((h, n, c, d3, dist))
    h2 = h.copy()
    h2["transform_subgroup"] = 1
    rejects((h2, n, c, d, dist))
    n2 = n.copy()
    n2["children_edges_offset"][1] = 1                                        # a cycle through the root's edges
    rejects((h, n2, c, d, dist))


def test_channel_modes_keep_the_voxel_local_hit_point(lrt, orc, gotrees):
    """CHANNEL / CHANNEL_LAMBERT store the hit point inside its voxel, 16 bits per axis, in
    coords[0..1] (x | y << 16, z): every surface voxel of the depth-5 sphere-union DAG is a 2/32
    cell, so the stored point is the hit point's offset in that cell."""
    pos, dir_ = sphere_rays(2000)
    for render in (23, 27):
        p = no_eps(lrt)
        p["render_mode"] = render
        h = oracle_trace(orc, lrt, dag_scene(lrt, de.encode(*gotrees[1], 1), p), pos, dir_)
        hit = h["primId"] != MISS
        c0 = h["coords"][hit, 0].view(np.uint32)
        c1 = h["coords"][hit, 1].view(np.uint32)
        q = np.stack([c0 & 0xFFFF, c0 >> 16, c1], 1).astype(np.float64) / 65535.0
        pt = pos[hit, :3].astype(np.float64) + h["t"][hit, None].astype(np.float64) * dir_[hit, :3]
        cell = (pt + 1.0) * 16.0
        frac = cell - np.floor(cell)
        err = np.minimum(np.abs(frac - q), 1.0 - np.abs(frac - q))   # a point on a cell face may sit in either cell
        assert hit.mean() > 0.2 and np.percentile(err, 99) < 1e-3
    # the plain Lambert mode leaves them zero on hits
    p = no_eps(lrt)
    h = oracle_trace(orc, lrt, dag_scene(lrt, de.encode(*gotrees[1], 1), p), pos, dir_)
    assert not np.any(h["coords"][h["primId"] != MISS, :2])


def _axis_maps(v):
    """Per rotation r of the reference's subgroup (restated in numpy from its generator): for each
    input axis, the output axis it lands on and its sign, read off the index map of lattice v."""
    out = []
    for m in de.universal_transpositions(v):
        def pt(i):
            return np.array([i // (v * v), (i // v) % v, i % v])
        o = pt(m[0])
        axes = []
        for a in range(3):
            e = [0, 0, 0]
            e[a] = 1
            dvec = pt(m[e[0] * v * v + e[1] * v + e[2]]) - o
            k = int(np.nonzero(dvec)[0][0])
            axes.append((k, int(dvec[k])))
        out.append(axes)
    return out


@pytest.mark.parametrize("bs", [1, 2])
def test_channel_modes_turn_the_quotients_by_the_channel_rotation(lrt, orc, gotrees, bs):
    """A node's channel edge names the rotation its channel data is stored in; the channel modes
    turn the voxel-local hit point by it (BVH2Common.cpp:2492-2505) with the transforms of the
    brick_size + 1 lattice: input axis i lands on the rotation's output axis, mirrored axes as
    bs - x. Checked against the unrotated quotients of the same hits and the axis maps of the
    subgroup generator restated in numpy."""
    pos, dir_ = sphere_rays(2000)
    p = no_eps(lrt)
    p["render_mode"] = 23
    tree = gotrees[1] if bs == 1 else gotrees[2]
    plain = oracle_trace(orc, lrt, dag_scene(lrt, de.encode(*tree, bs), p), pos, dir_)
    arrays = de.encode(*tree, bs, channel_rotate=True, seed=11)
    rot = oracle_trace(orc, lrt, dag_scene(lrt, arrays, p), pos, dir_)
    hit = plain["primId"] != MISS
    assert hit.mean() > 0.2 and np.array_equal(hit, rot["primId"] != MISS)
    assert np.array_equal(plain["t"].view(np.uint32), rot["t"].view(np.uint32))

    def q(h):
        c0, c1 = h["coords"][hit, 0].view(np.uint32), h["coords"][hit, 1].view(np.uint32)
        return np.stack([c0 & 0xFFFF, c0 >> 16, c1], 1).astype(np.int64)
    q0, qr = q(plain), q(rot)
    rid = arrays[1]["channels_edge"]["rotation_id"][rot["primId"][hit]]
    maps = _axis_maps(bs + 1)
    n_rot = 0
    for j in range(q0.shape[0]):
        for i, (o, sgn) in enumerate(maps[int(rid[j])]):
            if sgn > 0:
                assert qr[j, o] == q0[j, i]
            elif bs == 1:
                assert abs(qr[j, o] - (65535 - q0[j, i])) <= 1
            else:
                assert qr[j, o] == 65535   # (bs - x) leaves the unit cell: clamped
        n_rot += int(rid[j]) != 0
    assert n_rot > q0.shape[0] // 2
