"""GPU g-buffer resolve and tonemap against the oracle's restatement, on the records the GPU trace
wrote: every supported render mode bit for bit (NORMAL, which calls powf, within 2 ulp), both
normal modes on an instanced mesh with vertex normals, SDF leaves, per-voxel colour under
LAMBERT, point lights, highlight, anti-aliasing frames, sample accumulation and tonemapping.
Reference: MultiRenderer/MultiRenderer_resolve.cpp:10-75, 529-1005.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODES = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 15, 19, 20, 24, 25, 26]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    return torch


def _mesh_scene(g):
    rng = np.random.default_rng(5)
    v, f = g.gen_icosphere(4, 0.05, 4.0, 3)
    vn = v / np.linalg.norm(v, axis=1, keepdims=True)
    s = g.SceneRT()
    gid = s.AddGeom_Triangles3f(v, f, vnorm=vn.astype(np.float32))
    v2, f2 = g.gen_geodesic(6, 0.0, 4.0, 1)
    gid2 = s.AddGeom_Triangles3f(v2, f2)
    for i in range(6):
        m = np.eye(4, dtype=np.float32)
        q = rng.normal(size=4)
        a, b, c, d = q / np.linalg.norm(q)
        m[:3, :3] = np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)], [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                              [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]]) * rng.uniform(0.2, 0.5)
        m[:3, 3] = rng.uniform(-0.9, 0.9, 3)
        s.AddInstance(gid if i % 2 == 0 else gid2, m)
    s.CommitScene()
    return s


def _sdf_scene(g, kind):
    rng = np.random.default_rng(2)
    sph = np.concatenate([rng.uniform(-0.8, 0.8, (24, 3)), rng.uniform(0.1, 0.3, (24, 1))], 1).astype(np.float32)
    s = g.SceneRT()
    if kind == "frame":
        s.AddInstance(s.AddGeom_SdfFrameOctree(g.sdf_spheres_frame_octree(sph, 6)))
    else:   # per-voxel colour (TYPE_SDF_SBS_COL)
        h = np.zeros((), g.SBS_HEADER_DTYPE)
        h["brick_size"], h["brick_pad"], h["bytes_per_value"], h["aux_data"] = 4, 0, 1, g.SDF_SBS_NODE_LAYOUT_DX_RGB8
        n, v = g.sdf_spheres_sbs(sph, 4, h, colour_seed=7)
        s.AddInstance(s.AddGeom_SdfSBS(h, n, v))
    s.CommitScene()
    return s


def _frame(torch, g, orc, s, w, h, order, eye=(0.2, 0.3, 2.6)):
    pi, vi = g.camera_inverse(list(eye), [0, 0, 0], [0, 1, 0], 60.0, w / h, 0.01, 100.0)
    pos = torch.zeros((w * h, 4), dtype=torch.float32, device="cuda")
    dirs = torch.zeros_like(pos)
    g.eye_rays(pi, vi, w, h, pos, dirs, order=order)
    hits = torch.zeros((w * h, 8), dtype=torch.int32, device="cuda")
    s.trace_closest(pos, dirs, hits)
    torch.cuda.synchronize()
    return pi, vi, hits


def _compare(g, orc, s, torch, hits, pi, vi, w, h, order, mode, **kw):
    col = torch.zeros((w * h, 4), dtype=torch.float32, device="cuda")
    s.resolve(hits, pi, vi, w, h, col, order=order, **kw)
    torch.cuda.synchronize()
    got = col.cpu().numpy()
    hn = hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)
    okw = dict(kw)
    okw["lights"] = kw.get("lights", g.default_lights())
    want = orc.resolve(orc.Scene.from_scene(s), s.get_array("orig_indices"), hn, pi, vi, w, h, order=order, **okw)
    if mode == 8:
        ulp = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
        assert ulp.max() <= 2, f"mode 8: {ulp.max()} ulp"
    else:
        bad = np.any(got.view(np.uint32) != want.view(np.uint32), axis=1)
        assert not bad.any(), f"mode {mode}: {int(bad.sum())} pixels differ, first {np.flatnonzero(bad)[:3]}: " \
                              f"{got[bad][:2]} vs {want[bad][:2]}"
    return got


@pytest.mark.parametrize("normal_mode", [0, 1])
@pytest.mark.parametrize("mode", MODES)
def test_resolve_mesh_modes(lrt, orc, torch_dev, mode, normal_mode):
    g, torch = lrt, torch_dev
    _say(f"resolve mesh mode={mode} normal_mode={normal_mode}")
    s = _mesh_scene(g)
    p = g.default_preset()
    p["render_mode"], p["normal_mode"] = mode, normal_mode
    s.SetPreset(p)
    pi, vi, hits = _frame(torch, g, orc, s, 96, 80, 1)
    got = _compare(g, orc, s, torch, hits, pi, vi, 96, 80, 1, mode, background=(0.1, 0.2, 0.3, 0.7))
    # MASK, TYPE and TEX_COORDS (no texture coordinates on meshes) paint every hit alike
    assert len(np.unique(got[:, 0])) == 2 if mode in (0, 6, 15) else len(np.unique(got[:, 0])) > 2


@pytest.mark.parametrize("mode", MODES)
def test_resolve_sdf_modes(lrt, orc, torch_dev, mode):
    g, torch = lrt, torch_dev
    _say(f"resolve sdf mode={mode}")
    s = _sdf_scene(g, "frame")
    p = g.default_preset()
    p["render_mode"] = mode
    if mode == 10:
        p["sdf_node_intersect"] = 0
    s.SetPreset(p)
    pi, vi, hits = _frame(torch, g, orc, s, 64, 64, 0)
    _compare(g, orc, s, torch, hits, pi, vi, 64, 64, 0, mode)


def test_resolve_lights_highlight_aa_and_accumulation(lrt, orc, torch_dev):
    """Point + direct + ambient lights, a highlighted triangle, an anti-aliasing frame (alpha =
    t, Hamilton offset), two accumulated samples and the tonemap of their average."""
    g, torch = lrt, torch_dev
    _say("resolve lights / highlight / accumulation")
    s = _mesh_scene(g)
    pi, vi, hits = _frame(torch, g, orc, s, 64, 64, 1)
    hn = hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)
    first = hn[np.flatnonzero(hn["primId"] != 0xFFFFFFFF)[0]]
    L = np.zeros(3, g.LIGHT_DTYPE)
    L[0]["space"], L[0]["type"], L[0]["color"] = [0.5, 2.0, 1.5], g.LIGHT_TYPE_POINT, [3.0, 2.5, 2.0]
    L[1]["space"], L[1]["type"], L[1]["color"] = np.array([0, 0.6, 0.8], np.float32), g.LIGHT_TYPE_DIRECT, 0.5
    L[2]["type"], L[2]["color"] = g.LIGHT_TYPE_AMBIENT, [0.05, 0.1, 0.15]
    hl = (int(first["primId"]), int(first["geomId"]), int(first["instId"]))
    col = torch.zeros((64 * 64, 4), dtype=torch.float32, device="cuda")
    want = np.zeros((64 * 64, 4), np.float32)
    for sample, aa in ((0, 0), (1, 3)):
        s.resolve(hits, pi, vi, 64, 64, col, order=1, sample_id=sample, lights=L, aa_frame=aa, highlight=hl)
        orc.resolve(orc.Scene.from_scene(s), s.get_array("orig_indices"), hn, pi, vi, 64, 64, order=1,
                    sample_id=sample, aa_frame=aa, lights=L, highlight=hl, colour=want)
    torch.cuda.synchronize()
    got = col.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    rgba = torch.zeros(64 * 64, dtype=torch.int32, device="cuda")
    g.tonemap(col, rgba, 2)
    torch.cuda.synchronize()
    assert np.array_equal(rgba.cpu().numpy().view(np.uint32), orc.tonemap(want, 2))


def test_resolve_per_voxel_colour_lambert_and_errors(lrt, orc, torch_dev):
    """LAMBERT (17) shades per-voxel colour brick sets; on a mesh it needs material tables and
    fails with LRT_ERR_UNSUPPORTED, as do the PBR and channel modes."""
    g, torch = lrt, torch_dev
    _say("resolve per-voxel colour")
    s = _sdf_scene(g, "col")
    p = g.default_preset()
    p["render_mode"] = 17
    s.SetPreset(p)
    pi, vi, hits = _frame(torch, g, orc, s, 64, 64, 1)
    got = _compare(g, orc, s, torch, hits, pi, vi, 64, 64, 1, 17)
    assert len(np.unique(got[:, 1])) > 10
    m = _mesh_scene(g)
    col = torch.zeros((64 * 64, 4), dtype=torch.float32, device="cuda")
    for mode in (17, 21, 23):
        p["render_mode"] = mode
        m.SetPreset(p)
        with pytest.raises(g.LrtError) as e:
            m.resolve(hits, pi, vi, 64, 64, col)
        assert e.value.code == 4
