"""Dense SDF grids (BVHRT::AddGeom_SdfGrid, BVH2Common_host.cpp:466-499; IntersectSDFGrid and
eval_distance_sdf_grid, BVH2Common.cpp:1114-1182, 1442-1478) in the oracle, pinned by closed
forms: trilinear interpolation reproduces a linear distance exactly, so sphere tracing a plane
must stop on the plane (within the tracer's 1e-5 threshold) with the plane's normal; a sampled
sphere must be hit on the sphere within the grid's interpolation error with radial normals. The
plane is tilted on every axis, so a transposed grid index (x and z swapped) misses it.
Also the grid file format (sdf_simple.cpp:20-36) and the trilinear-only preset rule.
"""
import numpy as np
import pytest

from conftest import MISS, decode_normal, host_scene


def _grid(f, n):
    c = np.linspace(-1.0, 1.0, n)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    return f(x, y, z).astype(np.float32)   # [z, y, x]


def _rays(orc, lrt, eye, res=64):
    pi, vi = lrt.camera_inverse(list(eye), [0, 0, 0], [0, 1, 0], 45.0, 1.0, 0.01, 100.0)
    return orc.eye_rays(pi, vi, res, res)


def _trace(lrt, orc, data, pos, dirs, render_mode=1):
    s = host_scene(lrt, lambda s: s.AddInstance(s.AddGeom_SdfGrid(data)))
    p = lrt.default_preset()
    p["render_mode"] = render_mode
    s.SetPreset(p)
    hits, _ = orc.trace_closest(orc.Scene.from_scene(s), pos, dirs)
    return hits, s


def test_plane_is_hit_on_the_plane(lrt, orc):
    n = np.array([0.3, 0.8, 0.52])
    n /= np.linalg.norm(n)
    c0 = 0.1
    data = _grid(lambda x, y, z: n[0] * x + n[1] * y + n[2] * z - c0, 33)
    pos, dirs = _rays(orc, lrt, (0.5, 2.5, 1.5))
    hits, s = _trace(lrt, orc, data, pos, dirs)
    hit = hits["primId"] != MISS
    assert hit.mean() > 0.3
    assert s.get_array("geoms")["type"][0] == 1   # TYPE_SDF_GRID
    d = dirs[hit, :3].astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)   # the grid walk uses the normalised direction
    p = pos[hit, :3].astype(np.float64) + hits["t"][hit, None].astype(np.float64) * d
    inside = np.all(np.abs(p) < 1 - 1e-3, axis=1)
    assert inside.sum() > 100
    assert np.abs(p[inside] @ n - c0).max() < 2e-5
    nn = decode_normal(hits["coords"][hit][inside][:, 2:4])
    assert np.degrees(np.arccos(np.clip(nn @ n, -1, 1))).max() < 0.1
    # planted mistake: the same distances read with x and z swapped describe another plane
    bad, _ = _trace(lrt, orc, np.ascontiguousarray(np.transpose(data, (2, 1, 0))), pos, dirs)
    bh = bad["primId"] != MISS
    pb = pos[bh, :3].astype(np.float64) + bad["t"][bh, None].astype(np.float64) * (
        dirs[bh, :3] / np.linalg.norm(dirs[bh, :3], axis=1, keepdims=True))
    ib = np.all(np.abs(pb) < 1 - 1e-3, axis=1)
    assert np.abs(pb[ib] @ n - c0).max() > 0.05


def test_sphere_is_hit_on_the_sphere(lrt, orc):
    r = 0.55
    data = _grid(lambda x, y, z: np.sqrt(x * x + y * y + z * z) - r, 48)
    pos, dirs = _rays(orc, lrt, (0.4, 0.9, 2.8))
    hits, _ = _trace(lrt, orc, data, pos, dirs)
    hit = hits["primId"] != MISS
    assert hit.sum() > 400
    d = dirs[hit, :3].astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    p = pos[hit, :3].astype(np.float64) + hits["t"][hit, None].astype(np.float64) * d
    rad = np.linalg.norm(p, axis=1)
    assert np.abs(rad - r).max() < 0.01
    nn = decode_normal(hits["coords"][hit][:, 2:4])
    ang = np.degrees(np.arccos(np.clip((nn * (p / rad[:, None])).sum(1), -1, 1)))
    assert np.median(ang) < 2.0 and ang.max() < 10.0   # trilinear gradients: cell 2/47 against radius 0.55
    # sphere-tracing iterations (ST_ITERATIONS render mode): primId counts the steps
    it, _ = _trace(lrt, orc, data, pos, dirs, render_mode=10)
    assert np.array_equal(it["t"], hits["t"])
    assert it["primId"][hit].min() >= 1 and it["primId"][hit].max() < 1000


def test_grid_file_round_trip_and_errors(lrt, tmp_path):
    data = _grid(lambda x, y, z: x - 0.25 * y + z * z, 9)[:, :7, :5].copy()
    lrt.save_sdf_grid(str(tmp_path / "g.bin"), data)
    raw = (tmp_path / "g.bin").read_bytes()
    assert np.array_equal(np.frombuffer(raw[:12], np.uint32), [5, 7, 9])
    back = lrt.load_sdf_grid(str(tmp_path / "g.bin"))
    assert back.shape == data.shape and np.array_equal(back, data)
    (tmp_path / "cut.bin").write_bytes(raw[:-4])
    with pytest.raises(lrt.LrtError):
        lrt.load_sdf_grid(str(tmp_path / "cut.bin"))
