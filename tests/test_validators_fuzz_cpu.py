"""The host checks of the SDF formats (AddGeom_SdfFrameOctree, AddGeom_SdfSVS, AddGeom_SdfSBS,
AddGeom_SdfMultiOctree, AddGeom_COctreeV3, AddGeom_SdfDAG, AddGeom_SCom2) stand between untrusted
arrays and kernels that read without bounds checks: every word a ray can make a walk read must
lie inside the geometry's arrays. Random corruptions of the reference's files and of an encoded
plain SCom2 tree are either refused with LrtError or traced
by the oracle, whose walks read what the kernels read; under the host-sanitizer build (the
oracle and library built with AddressSanitizer) a read outside the arrays of an accepted tree is
reported there. Outside that build the test checks that every corruption is refused or traced
to the end."""
import os

import numpy as np
import pytest

from conftest import host_scene
from test_coctree_cpu import SPHERES, sphere_rays

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _corrupt(rng, words, n_flips):
    w = words.copy()
    for _ in range(n_flips):
        i = int(rng.integers(0, w.size))
        kind = int(rng.integers(0, 3))
        if kind == 0:
            w[i] ^= np.uint32(1) << np.uint32(rng.integers(0, 32))    # one bit
        elif kind == 1:
            w[i] = np.uint32(rng.integers(0, 2 ** 32, dtype=np.uint64))  # a random word
        else:
            w[i] = w[int(rng.integers(0, w.size))]                    # another word of the tree
    return w


def _trace(lrt, orc, build, pos, dir_):
    try:
        sc = host_scene(lrt, build)
    except lrt.LrtError:
        return False
    for mode, render in ((0, 8), (3, 23)):
        p = lrt.default_preset()
        p["sdf_node_intersect"], p["render_mode"] = mode, render
        sc.SetPreset(p)
        orc.trace_closest(orc.Scene.from_scene(sc), pos, dir_)
    return True


def test_corrupted_compact_formats_are_refused_or_traced(lrt, orc):
    import scom2_encode as se
    rng = np.random.default_rng(2024)
    pos, dir_ = sphere_rays(300)
    coct_h, coct_w, _ = lrt.load_coctree_v3(os.path.join(GOLD, "engine_sdf_coctree_v3_0.bin"))
    dag = lrt.load_sdf_dag(os.path.join(GOLD, "engine_sdf_dag_0.bin"))
    sc2 = lrt.load_scom2(os.path.join(GOLD, "engine_scom2_0.bin"))
    nodes, values = lrt.sdf_spheres_global_octree(SPHERES, 4, 2, 0)
    plain = se.encode(nodes, values, rotate=True, shift=True, seed=3)
    accepted = refused = 0
    for trial in range(24):
        flips = 1 + trial % 4
        cases = []
        w = _corrupt(rng, coct_w, flips)
        cases.append(lambda s, w=w: s.AddInstance(s.AddGeom_COctreeV3(coct_h, w, 0)))
        ce = dag[2].copy()
        cw = _corrupt(rng, ce.view(np.uint32).reshape(-1), flips).view(ce.dtype).reshape(ce.shape)
        cases.append(lambda s, cw=cw: s.AddInstance(s.AddGeom_SdfDAG(dag[0], dag[1], cw, dag[3], dag[4])))
        de = dag[3].copy()
        dw = _corrupt(rng, de.view(np.uint32).reshape(-1), flips).view(de.dtype).reshape(de.shape)
        cases.append(lambda s, dw=dw: s.AddInstance(s.AddGeom_SdfDAG(dag[0], dag[1], dag[2], dw, dag[4])))
        nd = dag[1].copy()
        nw = _corrupt(rng, nd.view(np.uint32).reshape(-1), flips).view(nd.dtype).reshape(nd.shape)
        cases.append(lambda s, nw=nw: s.AddInstance(s.AddGeom_SdfDAG(dag[0], nw, dag[2], dag[3], dag[4])))
        sn = _corrupt(rng, sc2[1], flips)
        cases.append(lambda s, sn=sn: s.AddInstance(s.AddGeom_SCom2(sc2[0], sn, sc2[2])))
        pn = _corrupt(rng, plain[1], flips)
        cases.append(lambda s, pn=pn: s.AddInstance(s.AddGeom_SCom2(plain[0], pn, plain[2])))
        # headers: shifts, masks, sizes and offsets the walks take as given
        ch = coct_h.copy()
        chw = _corrupt(rng, ch.reshape(1).view(np.uint32), 1).view(ch.dtype).reshape(ch.shape)
        cases.append(lambda s, chw=chw: s.AddInstance(s.AddGeom_COctreeV3(chw, coct_w, 0)))
        for hsrc, n_, b_ in ((sc2[0], sc2[1], sc2[2]), (plain[0], plain[1], plain[2])):
            hh = np.array(hsrc).copy()
            hw = _corrupt(rng, hh.reshape(1).view(np.uint32), 1).view(hh.dtype).reshape(hh.shape)
            cases.append(lambda s, hw=hw, n_=n_, b_=b_: s.AddInstance(s.AddGeom_SCom2(hw, n_, b_)))
        dh = np.array(dag[0]).copy()
        dhw = _corrupt(rng, dh.reshape(1).view(np.uint32), 1).view(dh.dtype).reshape(dh.shape)
        cases.append(lambda s, dhw=dhw: s.AddInstance(s.AddGeom_SdfDAG(dhw, dag[1], dag[2], dag[3], dag[4])))
        for build in cases:
            if _trace(lrt, orc, build, pos, dir_):
                accepted += 1
            else:
                refused += 1
    assert accepted > 20 and refused > 20, (accepted, refused)


def test_corrupted_hot_path_formats_are_refused_or_traced(lrt, orc):
    rng = np.random.default_rng(7)
    pos, dir_ = sphere_rays(300)
    frame = lrt.load_sdf_frame_octree(os.path.join(GOLD, "engine_sdf_frame_octree_0.bin"))
    svs = lrt.load_sdf_svs(os.path.join(GOLD, "test_10_svs.bin"))
    sbs = lrt.load_sdf_sbs(os.path.join(GOLD, "test_10_sbs.bin"))
    mo_n, mo_v = lrt.load_sdf_multi_octree(os.path.join(GOLD, "engine_sdf_multi_octree_0.bin"))

    def corrupt_rec(a, flips):
        return _corrupt(rng, a.view(np.uint32).reshape(-1), flips).view(a.dtype).reshape(a.shape)

    accepted = refused = 0
    for trial in range(24):
        flips = 1 + trial % 4
        fr = corrupt_rec(frame, flips)
        sv = corrupt_rec(svs, flips)
        sb_n = corrupt_rec(sbs[1], flips)
        sb_v = _corrupt(rng, sbs[2], flips)
        mn = corrupt_rec(mo_n, flips)
        tex = np.zeros(frame.size, lrt.FRAME_TEX_NODE_DTYPE)
        tex["values"], tex["offset"] = frame["values"], frame["offset"]
        tx = corrupt_rec(tex, flips)
        cases = [lambda s, fr=fr: s.AddInstance(s.AddGeom_SdfFrameOctree(fr)),
                 lambda s, tx=tx: s.AddInstance(s.AddGeom_SdfFrameOctreeTex(tx)),
                 lambda s, sv=sv: s.AddInstance(s.AddGeom_SdfSVS(sv)),
                 lambda s, sb_n=sb_n: s.AddInstance(s.AddGeom_SdfSBS(sbs[0], sb_n, sbs[2], *sbs[3:])),
                 lambda s, sb_v=sb_v: s.AddInstance(s.AddGeom_SdfSBS(sbs[0], sbs[1], sb_v, *sbs[3:])),
                 lambda s, sh=_corrupt(rng, np.array(sbs[0]).reshape(1).view(np.uint32), 1).view(
                     np.array(sbs[0]).dtype).reshape(()): s.AddInstance(s.AddGeom_SdfSBS(sh, sbs[1], sbs[2], *sbs[3:])),
                 lambda s, mn=mn: s.AddInstance(s.AddGeom_SdfMultiOctree(mn, mo_v))]
        for build in cases:
            if _trace(lrt, orc, build, pos, dir_):
                accepted += 1
            else:
                refused += 1
    assert accepted > 10 and refused > 10, (accepted, refused)
