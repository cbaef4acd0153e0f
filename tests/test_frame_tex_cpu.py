"""Textured frame octrees on the CPU (BVHRT::AddGeom_SdfFrameOctreeTex, BVH2Common_host.cpp:697-731;
RayNodeIntersection :168-188; the texture coordinates of LocalSurfaceIntersection :689-710).

The oracle is pinned three ways: (1) the textured octree of the same nodes is hit exactly like the
plain frame octree (the reference shares every step but the node array and the type), (2) with
each corner's texture coordinates set to a linear function of the corner's position, which
trilinear interpolation reproduces, coords[0..1] at a hit equal that function of the hit point
(a closed form that a wrong corner order or a dropped term breaks, shown by permuting the
corners), and (3) the file format round trips byte for byte. No fixture of this format ships with
the reference."""
import numpy as np
import pytest

from conftest import MISS, host_scene
from test_coctree_cpu import SPHERES, oracle_trace, sphere_rays


def node_cubes(nodes):
    """(corner position p in [0,1]^3, size d) of every node of a frame octree, root = unit cube."""
    pos = np.zeros((nodes.size, 3))
    size = np.zeros(nodes.size)
    size[0] = 1.0
    stack = [0]
    while stack:
        i = stack.pop()
        ofs = int(nodes["offset"][i])
        if ofs:
            for c in range(8):
                pos[ofs + c] = pos[i] + 0.5 * size[i] * np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
                size[ofs + c] = 0.5 * size[i]
                stack.append(ofs + c)
    return pos, size


def uv_of(p):
    """The linear texture map of the closed-form check: u = x, v = y + z / 2 over the unit cube."""
    return p[..., 0], p[..., 1] + 0.5 * p[..., 2]


def textured(lrt, frame, corner_order=None):
    tex = np.zeros(frame.size, lrt.FRAME_TEX_NODE_DTYPE)
    tex["values"], tex["offset"] = frame["values"], frame["offset"]
    tex["material_id"] = np.arange(frame.size) % 7
    pos, size = node_cubes(frame)
    order = corner_order if corner_order is not None else range(8)
    for j, c in enumerate(order):
        corner = pos + size[:, None] * np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
        u, v = uv_of(corner)
        tex["tex_coords"][:, 2 * j] = u
        tex["tex_coords"][:, 2 * j + 1] = v
    return tex


@pytest.fixture(scope="module")
def frame(lrt):
    return lrt.sdf_spheres_frame_octree(SPHERES, 5)


def _scene(lrt, add, nodes, mode=0, render=1):
    def build(s):
        s.AddInstance(add(s)(nodes))
        p = lrt.default_preset()
        p["sdf_node_intersect"], p["render_mode"] = mode, render
        s.SetPreset(p)
    return host_scene(lrt, build)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_textured_octree_hits_like_the_frame_octree(lrt, orc, frame, mode):
    pos, dir_ = sphere_rays(3000)
    a = oracle_trace(orc, lrt, _scene(lrt, lambda s: s.AddGeom_SdfFrameOctree, frame, mode), pos, dir_)
    b = oracle_trace(orc, lrt, _scene(lrt, lambda s: s.AddGeom_SdfFrameOctreeTex, textured(lrt, frame), mode), pos, dir_)
    ha = a["primId"] != MISS
    assert ha.mean() > 0.2
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
    assert np.array_equal(a["primId"], b["primId"]) and np.array_equal(a["instId"], b["instId"])
    assert np.all((b["geomId"][ha] >> 24) == lrt.TYPE_SDF_FRAME_OCTREE_TEX)
    assert np.array_equal(a["geomId"][ha] & 0xFFFFFF, b["geomId"][ha] & 0xFFFFFF)
    assert np.array_equal(a["coords"][:, 2:].view(np.uint32), b["coords"][:, 2:].view(np.uint32))
    assert np.all(a["coords"][ha, :2] == 0)


def test_texture_coordinates_closed_form(lrt, orc, frame):
    """u = x, v = y + z/2 at the corners gives exactly that at every point of a leaf: compare with
    the hit point pos + t dir mapped to the unit cube (float32 rounding of the voxel-local point)."""
    pos, dir_ = sphere_rays(3000)
    b = oracle_trace(orc, lrt, _scene(lrt, lambda s: s.AddGeom_SdfFrameOctreeTex, textured(lrt, frame)), pos, dir_)
    h = b["primId"] != MISS
    p = (pos[h, :3].astype(np.float64) + b["t"][h, None].astype(np.float64) * dir_[h, :3] + 1.0) / 2.0
    u, v = uv_of(p)
    assert np.max(np.abs(b["coords"][h, 0] - u)) < 2e-5
    assert np.max(np.abs(b["coords"][h, 1] - v)) < 2e-5
    # a wrong corner order (x and z bits swapped) breaks it
    swapped = [((c & 1) << 2) | (c & 2) | ((c >> 2) & 1) for c in range(8)]
    w = oracle_trace(orc, lrt, _scene(lrt, lambda s: s.AddGeom_SdfFrameOctreeTex, textured(lrt, frame, swapped)),
                     pos, dir_)
    assert np.max(np.abs(w["coords"][h, 0] - u)) > 1e-3


def test_iterations_mode_and_instances(lrt, orc, frame):
    """ST_ITERATIONS keeps the iteration count as primId; a transformed instance leaves the
    texture coordinates (voxel-local) as they are in the untransformed octree's frame."""
    pos, dir_ = sphere_rays(2000)
    tex = textured(lrt, frame)
    a = oracle_trace(orc, lrt, _scene(lrt, lambda s: s.AddGeom_SdfFrameOctree, frame, 2, 10), pos, dir_)
    b = oracle_trace(orc, lrt, _scene(lrt, lambda s: s.AddGeom_SdfFrameOctreeTex, tex, 2, 10), pos, dir_)
    assert np.array_equal(a["primId"], b["primId"])
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] *= 0.5
    m[:3, 3] = [0.2, -0.1, 0.1]

    def build(s):
        s.AddInstance(s.AddGeom_SdfFrameOctreeTex(tex), m)
    c = oracle_trace(orc, lrt, host_scene(lrt, build), pos, dir_)
    h = c["primId"] != MISS
    assert h.mean() > 0.05
    pw = pos[h, :3].astype(np.float64) + c["t"][h, None].astype(np.float64) * dir_[h, :3]
    pl = (pw - m[:3, 3]) / 0.5
    u, v = uv_of((pl + 1.0) / 2.0)
    assert np.max(np.abs(c["coords"][h, 0] - u)) < 5e-5 and np.max(np.abs(c["coords"][h, 1] - v)) < 5e-5


def test_file_round_trip_and_host_checks(lrt, frame, tmp_path):
    tex = textured(lrt, frame)
    f = tmp_path / "t.bin"
    lrt.save_sdf_frame_octree_tex(str(f), tex)
    raw = f.read_bytes()
    assert len(raw) == 4 + 104 * tex.size and int(np.frombuffer(raw[:4], np.uint32)[0]) == tex.size
    back = lrt.load_sdf_frame_octree_tex(str(f))
    assert back.tobytes() == tex.tobytes()
    (tmp_path / "cut.bin").write_bytes(raw[:-8])
    with pytest.raises(lrt.LrtError):
        lrt.load_sdf_frame_octree_tex(str(tmp_path / "cut.bin"))
    bad = tex.copy()
    bad["offset"][0] = tex.size - 3          # children past the end
    with pytest.raises(lrt.LrtError):
        _scene(lrt, lambda s: s.AddGeom_SdfFrameOctreeTex, bad)
    empty = tex[:1].copy()
    empty["offset"], empty["values"] = 0, 1.0   # no surface leaf
    with pytest.raises(lrt.LrtError):
        _scene(lrt, lambda s: s.AddGeom_SdfFrameOctreeTex, empty)
