"""C-ABI library: loads, exports every declared symbol, validates arguments, builds host arrays."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, host_scene


def declared_symbols():
    names = set()
    for hdr in ("lrt.h", "lrt_build.h"):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(lrt_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol(lrt):
    path = lrt.library_path()
    out = subprocess.run(["nm", "-D", "--defined-only", path], stdout=subprocess.PIPE, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, f"declared but not exported: {missing}"
    assert set(lrt.exported_symbols()) <= set(declared_symbols())


def test_struct_sizes_match_reference(lrt):
    # CrossRT.h:26-33, cbvh.h:21-33, intersection_common.h:33-56, sdf_simple_shared.h:23-55,
    # sdf_multi_shared.h:8-14, render_settings.h:116-136
    assert lrt.HIT_DTYPE.itemsize == 32
    assert lrt.NODE_DTYPE.itemsize == 32
    assert lrt.NODE_PAIR_DTYPE.itemsize == 64
    assert lrt.GEOM_DTYPE.itemsize == 64
    assert lrt.INSTANCE_DTYPE.itemsize == 256
    assert lrt.FRAME_NODE_DTYPE.itemsize == 36
    assert lrt.SBS_NODE_DTYPE.itemsize == 16
    assert lrt.SBS_HEADER_DTYPE.itemsize == 16
    assert lrt.PRESET_DTYPE.itemsize == 60


def test_create_scene_rt_unknown_impl_returns_none(lrt):
    assert lrt.CreateSceneRT("NoSuchImpl") is None
    assert lrt.CreateSceneRT("BVH2Common") is not None


def test_errors_are_reported_not_thrown(lrt):
    s = lrt.SceneRT()
    v = np.zeros((3, 3), np.float32)
    with pytest.raises(lrt.LrtError) as e:
        s.AddGeom_Triangles3f(v, np.array([0, 1, 5], np.uint32))   # index out of range
    assert e.value.code == 1 and "out of range" in str(e.value)
    with pytest.raises(lrt.LrtError):
        s.AddInstance(7)                                            # invalid geometry id
    with pytest.raises(lrt.LrtError):
        s.CommitScene()                                             # no instances
    with pytest.raises(lrt.LrtError):
        s.set_leaf_postponement(2)                                  # modes are -1, 0, 1
    s.set_leaf_postponement(0)
    for bad in (3, 65, 128):
        with pytest.raises(lrt.LrtError):
            s.set_ray_interleave(bad)                               # powers of two up to 64
    s.set_ray_interleave(0)
    with pytest.raises(lrt.LrtError):
        s.set_launch_config(100, 16, True)


def test_trace_before_commit_is_an_error(lrt):
    s = lrt.SceneRT()
    with pytest.raises(lrt.LrtError) as e:
        s.trace_closest_host(np.zeros((1, 4), np.float32), np.ones((1, 4), np.float32))
    assert e.value.code == 3


def test_host_arrays_of_a_small_mesh(lrt):
    v, f = lrt.gen_icosphere(2)
    s = host_scene(lrt, lambda s: s.AddInstance(s.AddGeom_Triangles3f(v, f)))
    geoms = s.get_array("geoms")
    assert geoms.size == 1 and geoms[0]["type"] == lrt.TYPE_MESH_TRIANGLE
    assert s.get_array("indices").size == f.size
    assert np.array_equal(np.sort(s.get_array("prim_indices")), np.arange(f.shape[0]))
    inst = s.get_array("instances")
    assert np.allclose(inst[0]["transformInv"].reshape(4, 4), np.eye(4))
    tlas = s.get_array("tlas")
    assert tlas.size == 1 and tlas[0]["leftOffset"] == 0x80000000 and tlas[0]["escapeIndex"] == 0xFFFFFFFF


def test_icosphere_and_geodesic_counts(lrt):
    for k in range(0, 5):
        v, f = lrt.gen_icosphere(k)
        assert f.shape[0] == 20 * 4 ** k and v.shape[0] == 10 * 4 ** k + 2
    for n in (1, 2, 5, 7):
        v, f = lrt.gen_geodesic(n)
        assert f.shape[0] == 20 * n * n and v.shape[0] == 10 * n * n + 2
    v, _ = lrt.gen_icosphere(3)
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-6)


def test_multi_octree_validation(lrt):
    nodes = np.zeros(9, lrt.MULTI_NODE_DTYPE)
    nodes[0]["children_offset"] = 1
    values = np.zeros(8 * 9, np.float32)
    s = lrt.SceneRT()
    assert s.AddGeom_SdfMultiOctree(nodes, values) == 0
    bad = nodes.copy()
    bad[3]["children_offset"] = 2          # child stored before itself: a cycle is possible
    with pytest.raises(lrt.LrtError) as e:
        s.AddGeom_SdfMultiOctree(bad, values)
    assert e.value.code == 1
    bad = nodes.copy()
    bad[5]["data_offset"], bad[5]["voxel_count"] = 70, 1   # values past the end
    with pytest.raises(lrt.LrtError):
        s.AddGeom_SdfMultiOctree(bad, values)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message and a non-zero status."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stderr and r.stdout.strip() == ""


def test_differentiable_calls_refuse_mixed_scenes(lrt):
    """The adjoint walk intersects brick sets only: a scene that also holds a mesh is refused
    (its occluders would otherwise be skipped without a word)."""
    rng = np.random.default_rng(2)
    sph = np.concatenate([rng.uniform(-0.5, 0.5, (4, 3)), rng.uniform(0.1, 0.3, (4, 1))], 1).astype(np.float32)
    h, nodes, vals, vf, nd = lrt.sdf_spheres_sbs_indexed(sph, 2, 8, 1)
    v, f = lrt.gen_icosphere(1)

    def build(s, mesh):
        gid = s.AddGeom_SdfSBS(h, nodes, vals, vf)
        s.AddInstance(gid)
        if mesh:
            s.AddInstance(s.AddGeom_Triangles3f(v, f))

    ok = host_scene(lrt, lambda s: build(s, False))
    assert ok.sbs_param_layout(0)[1] == nd
    mixed = host_scene(lrt, lambda s: build(s, True))
    with pytest.raises(lrt.LrtError) as e:
        mixed.sbs_param_layout(0)
    assert e.value.code == 4 and "brick sets" in str(e.value)
