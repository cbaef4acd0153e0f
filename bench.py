#!/usr/bin/env python3
"""Ray-query benchmark on one MI355X.

A step is one pass of a workload's hot path over its full ray batch, inputs resident in HBM: a
closest-hit trace call (both passes), or one fused forward + adjoint of the differentiable brick
set. The default run (--workload all) prints ONE JSON line whose headline is the closest-hit
icosphere7_512 workload (512x512 primary rays on a 327,680-triangle mesh, the reference's
closest-hit configuration) and which carries every other workload under "sub_results", each with
its ms/step, Mrays/s, roofline (algorithmic bytes per step from traversal counters / step time,
against 8 TB/s; measured HBM bytes per step from the committed counter summaries; the limiting
stage named by the committed counter ratios) and the CPU oracle timed on the host cores.

Workloads (synthetic, seeded):
  sdf_octree_d9_1024  64-sphere union, frame octree depth 9, 1024x1024 jittered primary rays
  sdf_multi_d9_1024   the same depth-9 octree and rays as sdf_octree_d9_1024, walked as a multi-octree
                      (front-to-back octree walk instead of a BVH over the leaf cells)
  icosphere7_512      icosphere subdivided 7x (327,680 tris), radial value noise 0.05, 512x512 rays
  tlas_4096x32_4m     two-level BVH: 4,096 instances (random rotation, scale U[0.5,2], translation
                      U[0,100]^3, seed 8) of 32 noise-displaced geodesic spheres with 20 n^2
                      triangles, n log-uniform in [5,100] (seed 7); 4,194,304 cosine-hemisphere
                      secondary rays leaving random instance surfaces (seed 9)
  sbs_batch_64x4x128  64 differentiable brick-set scenes (64-sphere unions, seeds 1000+k) with grid
                      resolutions drawn from {32,64,128,256}^3 (seed 10), 8^3 bricks, 4 ring views
                      of 128x128 each; target = each scene rendered with inverted albedo; a step =
                      one batched run (forward + L2 adjoint of all 4.2 M rays, one persistent
                      kernel, HIP-graph replay); also timed: the same work as per-scene launches
                      (forward then L2 backward per scene, and the fused per-scene call)
  sbs_diff_256_8x256  differentiable brick set: 64-sphere union (seed 4) on a 256^3 grid of 8^3
                      fp32 bricks with per-lattice-point RGB albedo; 8 ring views of 256x256; a
                      step = forward render + adjoint of the L2 loss against the image of a second
                      scene (seed 5), one fused traversal per ray; rays/s counts each ray once per
                      step
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s closest-hit & fwd+bwd SDF rays/s on 1 MI355X; % of HBM peak"

# Algorithmic bytes per traversal event (per-ray counters of the instrumented kernel):
# ray in + hit out, fat node, TLAS node, leaf start/count word, triangle (3 indices +
# 3 float4 vertices), instance entered (transformInv + GeomData), frame-octree leaf
# (leaf box + 8 corner values + offset), primId remap word per triangle hit.
B_RAY_IO, B_FAT_NODE, B_TLAS_NODE, B_LEAF, B_TRI = 64, 64, 32, 4, 60
B_INSTANCE, B_FRAME_LEAF, B_REMAP = 128, 32 + 36, 4
B_BRICK, B_VOXEL = 16, 64          # brick header (SdfSBSNode); voxel stepped: 8 indices + 8 values


def camera():
    import graphic_lab_task_amd as g
    return g.camera_inverse([0, 0, 3], [0, 0, 0], [0, 1, 0], 60.0, 1.0, 0.01, 100.0)


def spheres(seed=2, n=64):
    """64-sphere union: centres U[-0.8, 0.8]^3, radii U[0.05, 0.2]."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-0.8, 0.8, size=(n, 3))
    r = rng.uniform(0.05, 0.2, size=(n, 1))
    return np.concatenate([c, r], axis=1).astype(np.float32)


def rotation(q):
    a, b, c, d = q / np.linalg.norm(q)
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def tlas_scene(n_inst=4096, n_mesh=32):
    """The batched two-level scene: (scene, meshes [(verts, tris)], instance matrices (row-major))."""
    import graphic_lab_task_amd as g
    rng = np.random.default_rng(7)
    freqs = np.exp(rng.uniform(np.log(5), np.log(100), n_mesh)).round().astype(int)
    scene = g.SceneRT()
    meshes = []
    for k, nf in enumerate(freqs):
        v, f = g.gen_geodesic(int(nf), 0.05, 4.0, 100 + k)
        meshes.append((v, f))
        scene.AddGeom_Triangles3f(v, f)
    rng = np.random.default_rng(8)
    mats = np.zeros((n_inst, 4, 4), np.float32)
    for i in range(n_inst):
        m = np.eye(4)
        m[:3, :3] = rotation(rng.normal(size=4)) * rng.uniform(0.5, 2.0)
        m[:3, 3] = rng.uniform(0, 100, 3)
        mats[i] = m
        scene.AddInstance(i % n_mesh, mats[i])
    try:
        scene.CommitScene()
    except g.LrtError as e:   # no device (host-side checks): the host arrays are built
        if e.code != 5:
            raise
    return scene, meshes, mats


def hemisphere_rays(meshes, mats, n, seed=9):
    """Rays leaving random points of random instance surfaces, cosine-distributed about the
    geometric normal; origin offset 1e-3 along the normal, near 1e-4, far 1e9."""
    rng = np.random.default_rng(seed)
    n_mesh = len(meshes)
    inst = rng.integers(0, mats.shape[0], n)
    pos = np.zeros((n, 4), np.float32)
    dirs = np.zeros((n, 4), np.float32)
    for k in range(n_mesh):
        sel = np.nonzero(inst % n_mesh == k)[0]
        if sel.size == 0:
            continue
        v, f = meshes[k]
        v = v[:, :3].astype(np.float64)
        tri = f[rng.integers(0, f.shape[0], sel.size)]
        A, B, C = v[tri[:, 0]], v[tri[:, 1]], v[tri[:, 2]]
        r1, r2 = np.sqrt(rng.random(sel.size))[:, None], rng.random(sel.size)[:, None]
        p = (1 - r1) * A + r1 * (1 - r2) * B + r1 * r2 * C
        nrm = np.cross(B - A, C - A)
        M = mats[inst[sel]].astype(np.float64)
        pw = np.einsum("nij,nj->ni", M[:, :3, :3], p) + M[:, :3, 3]
        nw = np.einsum("nij,nj->ni", np.linalg.inv(M[:, :3, :3]).transpose(0, 2, 1), nrm)
        nw /= np.linalg.norm(nw, axis=1, keepdims=True)
        t1 = np.where(np.abs(nw[:, :1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
        t1 = np.cross(nw, t1)
        t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
        t2 = np.cross(nw, t1)
        u1, u2 = rng.random(sel.size)[:, None], rng.random(sel.size)[:, None]
        rr, phi = np.sqrt(u1), 2 * np.pi * u2
        d = rr * np.cos(phi) * t1 + rr * np.sin(phi) * t2 + np.sqrt(1 - u1) * nw
        pos[sel, :3] = pw + 1e-3 * nw
        pos[sel, 3] = 1e-4
        dirs[sel, :3] = d / np.linalg.norm(d, axis=1, keepdims=True)
        dirs[sel, 3] = 1e9
    return pos, dirs


def build_workload(name, threads):
    """Returns (scene, width, height, jitter or None, description dict, leaf kind)."""
    import graphic_lab_task_amd as g
    t0 = time.time()
    if name == "icosphere7_512":
        v, f = g.gen_icosphere(7, 0.05, 4.0, 1)
        scene = g.SceneRT()
        gid = scene.AddGeom_Triangles3f(v, f)
        scene.AddInstance(gid)
        scene.CommitScene()
        desc = {"workload": name, "triangles": int(f.shape[0]), "rays": 512 * 512, "width": 512, "height": 512}
        return scene, 512, 512, None, desc, "tri", time.time() - t0
    if name == "sdf_octree_d9_1024":
        nodes = g.sdf_spheres_frame_octree(spheres(), 9, threads)
        scene = g.SceneRT()
        gid = scene.AddGeom_SdfFrameOctree(nodes)
        scene.AddInstance(gid)
        scene.CommitScene()
        jitter = np.random.default_rng(3).random((1024 * 1024, 2), dtype=np.float32)
        desc = {"workload": name, "octree_nodes": int(nodes.size), "depth": 9, "spheres": 64,
                "surface_leaves": int(scene.get_array("orig_nodes").size), "rays": 1024 * 1024,
                "width": 1024, "height": 1024}
        return scene, 1024, 1024, jitter, desc, "frame", time.time() - t0
    if name == "sdf_multi_d9_1024":
        mn, mv = g.sdf_spheres_multi_octree(spheres(), 9, threads)
        scene = g.SceneRT()
        scene.AddInstance(scene.AddGeom_SdfMultiOctree(mn, mv))
        scene.CommitScene()
        jitter = np.random.default_rng(3).random((1024 * 1024, 2), dtype=np.float32)
        desc = {"workload": name, "octree_nodes": int(mn.size), "depth": 9, "spheres": 64, "rays": 1024 * 1024,
                "width": 1024, "height": 1024}
        return scene, 1024, 1024, jitter, desc, "multi", time.time() - t0
    if name == "tlas_4096x32_4m":
        scene, meshes, mats = tlas_scene()
        pos, dirs = hemisphere_rays(meshes, mats, 4 * 1024 * 1024)
        desc = {"workload": name, "instances": int(mats.shape[0]), "meshes": len(meshes),
                "triangles_unique": int(sum(f.shape[0] for _, f in meshes)),
                "triangles_instanced": int(sum(meshes[i % len(meshes)][1].shape[0] for i in range(mats.shape[0]))),
                "rays": int(pos.shape[0]), "ray_kind": "cosine-hemisphere secondary"}
        return scene, (pos, dirs), None, None, desc, "tri", time.time() - t0
    raise SystemExit(f"unknown workload {name}")


def sbs_batch_scenes(torch, g, n_scenes=64, res=128, views=4, threads=0, seed=10, stream=0):
    """Items (dicts for graphic_lab_task_amd.SbsBatch) of the batched differentiable workload."""
    import graphic_lab_task_amd.diff as gd
    grids = np.random.default_rng(seed).choice([32, 64, 128, 256], n_scenes)
    pos, dirs = gd.ring_rays(torch, res, res, count=views, stream=stream)
    items = []
    for k, gr in enumerate(grids):
        depth = int(np.log2(gr // 8))
        h, nodes, vals, vf, nd = g.sdf_spheres_sbs_indexed(spheres(1000 + k), depth, 8, 1 + k, threads)
        s = g.SceneRT()
        gid = s.AddGeom_SdfSBS(h, nodes, vals, vf)
        s.AddInstance(gid)
        s.CommitScene()
        params = torch.from_numpy(vf).cuda()
        tparams = params.clone()
        tparams[nd:] = 1.0 - tparams[nd:]
        target = torch.zeros((pos.shape[0], 4), dtype=torch.float32, device="cuda")
        s.sbs_forward(gid, tparams, pos, dirs, target, stream=stream)
        items.append({"scene": s, "geom_id": gid, "params": params, "grad": torch.zeros_like(params),
                      "loss": torch.zeros(1, dtype=torch.float32, device="cuda"), "pos": pos, "dir": dirs,
                      "target": target, "colour": torch.zeros_like(target), "grid": int(gr), "n_dist": int(nd),
                      "bricks": int(nodes.size)})
    return items


def timed(torch, stream, fn, steps, warmup):
    """Mean ms per call of fn over `steps` calls after `warmup`, HIP events on `stream`."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def sbs_traversal_bytes(torch, g, scene, pos, dirs, sh):
    """Traversal bytes (brick headers, voxels stepped, nodes) and hit count of one brick-set trace
    of the given rays, from the instrumented launch's per-ray counters (untimed)."""
    n = pos.shape[0]
    hits = torch.zeros((n, 8), dtype=torch.int32, device="cuda")
    counts = torch.zeros((n, 8), dtype=torch.int32, device="cuda")
    scene.trace_closest_counted(pos, dirs, hits, counts, stream=sh)
    torch.cuda.synchronize()
    c = counts.cpu().numpy().view(np.uint32).astype(np.float64)
    hit_n = float((hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)["primId"] != g.MISS).sum())
    trav = B_FAT_NODE * c[:, 0].sum() + B_TLAS_NODE * c[:, 3].sum() + B_LEAF * c[:, 1].sum() + \
        B_BRICK * c[:, 1].sum() + B_VOXEL * c[:, 5].sum() + B_INSTANCE * c[:, 4].sum()
    return trav, hit_n


# Per hit of the fused forward + L2 adjoint: 32 parameters read (8 distances + 24 albedo floats)
# and 32 float atomics (read-modify-write, 8 B each).
B_DIFF_HIT = 32 * 4 + 32 * 8
# Per ray of the fused pass: ray in (32 B), target in (16 B), colour out (16 B).
B_DIFF_RAY = 32 + 16 + 16


def run_batch(args, torch, g, cpu_target_s):
    """Batched differentiable workload: one persistent launch per step vs. per-scene launches."""
    t0 = time.time()
    threads = min(16, len(os.sched_getaffinity(0)))
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    items = sbs_batch_scenes(torch, g, threads=threads, stream=sh)
    build_s = time.time() - t0
    n = sum(it["pos"].shape[0] for it in items)
    batch = g.SbsBatch(items)
    tune = None
    if args.batch_config:
        lds, rec, grid = (int(x) for x in args.batch_config.split(","))
        batch.configure(lds, rec, grid)
        tune = {"lds_stack": lds, "record": rec, "grid": grid, "given": True}
    elif not args.no_autotune:
        tune = batch.autotune(sh)
    if args.copy_calibration:
        copy_calibration(torch)

    # algorithmic bytes of one batched step: per item, the traversal of its rays (instrumented
    # launch), ray I/O and the per-hit gradient scatter, plus the set-up kernels: gradient zeroing
    # (4 B per parameter) and the occupancy masks (one pass over each brick's (bs+1)^3 value
    # indices and values, 8 B each). The hit brick's second march is not counted.
    alg_bytes, hit_n = 0.0, 0.0
    for it in items:
        trav, hn = sbs_traversal_bytes(torch, g, it["scene"], it["pos"], it["dir"], sh)
        alg_bytes += it["pos"].shape[0] * B_DIFF_RAY + trav + hn * B_DIFF_HIT
        alg_bytes += it["params"].numel() * 4 + it["bricks"] * 729 * 8
        hit_n += hn

    def plain():
        for it in items:
            it["grad"].zero_()
            it["loss"].zero_()
            it["scene"].sbs_forward(it["geom_id"], it["params"], it["pos"], it["dir"], it["colour"], stream=sh)
            it["scene"].sbs_l2_backward(it["geom_id"], it["params"], it["pos"], it["dir"], it["target"], it["grad"],
                                        it["loss"], stream=sh)

    def plain_fused():
        for it in items:
            it["grad"].zero_()
            it["loss"].zero_()
            it["scene"].sbs_render_l2_backward(it["geom_id"], it["params"], it["pos"], it["dir"], it["target"],
                                               it["colour"], it["grad"], it["loss"], stream=sh)

    ms_plain = timed(torch, stream, plain, max(3, args.steps // 5), args.warmup)
    ms_plain_fused = timed(torch, stream, plain_fused, max(3, args.steps // 5), args.warmup)
    ms_eager = timed(torch, stream, lambda: batch.run(sh), args.steps, args.warmup)
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(stream)
    with torch.cuda.stream(side):
        batch.run(side.cuda_stream)            # warm-up on the capture stream
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=side):
            batch.run(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ms = timed(torch, stream, graph.replay, args.steps, args.warmup)
    grids = [it["grid"] for it in items]
    desc = {"workload": "sbs_batch_64x4x128", "scenes": len(items), "views_per_scene": 4, "width": 128, "height": 128,
            "rays": int(n), "grids": {str(r): grids.count(r) for r in (32, 64, 128, 256)}, "brick": 8,
            "hit_fraction": round(hit_n / n, 4), "scene_build_s": round(build_s, 1),
            "ms_batched_eager": round(ms_eager, 4), "ms_per_scene_launches": round(ms_plain, 4),
            "batched_vs_per_scene": round(ms_plain / ms, 2),
            "ms_per_scene_fused_launches": round(ms_plain_fused, 4),
            "batched_vs_per_scene_fused": round(ms_plain_fused / ms, 2),
            "recompute_misses": int(sum(it["scene"].sbs_recompute_misses() for it in items)),
            "launch": tune, "autotuned": tune is not None,
            "algorithmic_bytes_per_ray": round(alg_bytes / n, 1)}
    cpu = None
    if cpu_target_s > 0:
        # the oracle's forward walk (closest hit of every ray) over an evenly strided sample of
        # each scene's rays: the CPU side of the forward half of the step
        pos_np, dir_np = items[0]["pos"].cpu().numpy(), items[0]["dir"].cpu().numpy()
        cpu = cpu_baseline([it["scene"] for it in items], pos_np, dir_np, threads, cpu_target_s,
                           what="forward closest hit of the same rays (the oracle has no adjoint)")
    return result("sbs_batch_64x4x128", n, ms, args, desc, alg_bytes, cpu, unit="Mrays/s (fwd+bwd, batched)")


def copy_calibration(torch):
    """A 256 MiB elementwise copy: the reference kernel that scales the HBM byte counters."""
    src = torch.empty(64 << 20, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.neg(src, out=dst)   # one elementwise kernel: reads 256 MiB, writes 256 MiB
    torch.cuda.synchronize()
    del src, dst


def run_diff(args, torch, g, cpu_target_s):
    """Forward + adjoint steps of the differentiable brick-set workload."""
    import graphic_lab_task_amd.diff as gd
    t0 = time.time()
    threads = min(16, len(os.sched_getaffinity(0)))

    def build(seed, albedo_seed):
        h, nodes, vals, vf, nd = g.sdf_spheres_sbs_indexed(spheres(seed), 5, 8, albedo_seed, threads)
        s = g.SceneRT()
        gid = s.AddGeom_SdfSBS(h, nodes, vals, vf)
        s.AddInstance(gid)
        s.CommitScene()
        return s, gid, vf, nd, nodes.size

    scene, gid, vf, nd, n_bricks = build(4, 1)
    tscene, tgid, tvf, _, _ = build(5, 2)
    build_s = time.time() - t0
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    pos, dirs = gd.ring_rays(torch, 256, 256, count=8, stream=sh)
    n = pos.shape[0]
    params = torch.from_numpy(vf).cuda()
    target = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    tscene.sbs_forward(tgid, torch.from_numpy(tvf).cuda(), pos, dirs, target, stream=sh)
    colour = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    grad = torch.zeros_like(params)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")

    # one fused pass: ray I/O, one traversal (the hit voxel is recorded by the walk's own brick
    # march, not marched again), 32 parameters read and 32 atomics per hit
    trav, hit_n = sbs_traversal_bytes(torch, g, scene, pos, dirs, sh)
    alg_bytes = n * B_DIFF_RAY + trav + hit_n * B_DIFF_HIT

    if args.copy_calibration:
        copy_calibration(torch)

    def step():
        grad.zero_()
        loss.zero_()
        # forward colour and L2 adjoint in one traversal per ray
        scene.sbs_render_l2_backward(gid, params, pos, dirs, target, colour, grad, loss, stream=sh)

    ms = timed(torch, stream, step, args.steps, args.warmup)
    desc = {"workload": "sbs_diff_256_8x256", "virtual_grid": 256, "brick": 8, "bricks": int(n_bricks),
            "distance_values": int(nd), "views": 8, "width": 256, "height": 256, "rays": int(n),
            "hit_fraction": round(hit_n / n, 4), "scene_build_s": round(build_s, 2),
            "recompute_misses": scene.sbs_recompute_misses(), "loss": float(loss.item()),
            "algorithmic_bytes_per_ray": round(alg_bytes / n, 1)}
    cpu = None
    if cpu_target_s > 0:
        cpu = cpu_baseline([scene], pos.cpu().numpy(), dirs.cpu().numpy(), threads, cpu_target_s,
                           what="forward closest hit of the same rays (the oracle has no adjoint)")
    return result("sbs_diff_256_8x256", n, ms, args, desc, alg_bytes, cpu, unit="Mrays/s (fwd+bwd)")


def algorithmic_bytes(cnt, hits, kind):
    import graphic_lab_task_amd as g
    c = cnt.astype(np.float64)
    b = B_RAY_IO * c.shape[0] + B_FAT_NODE * c[:, 0].sum() + B_TLAS_NODE * c[:, 3].sum() + B_LEAF * c[:, 1].sum()
    b += B_TRI * c[:, 2].sum() + B_INSTANCE * c[:, 4].sum()
    if kind == "frame":
        b += B_FRAME_LEAF * c[:, 1].sum()
    if kind == "multi":   # the walk reports its octree reads in 16-byte units (node records, voxel values)
        b += 16 * c[:, 5].sum()
    if kind == "tri":
        b += B_REMAP * float((hits["primId"] != g.MISS).sum())
    return b


def strided(a, k):
    """Every n/k-th ray: a sample spread over the whole image (rays are in screen-tile order)."""
    step = max(1, a.shape[0] // k)
    return np.ascontiguousarray(a[::step][:k])


def cpu_baseline(scenes, pos, dir, threads, target_s=20.0, what="closest hit of the same rays"):
    """Oracle (reference flags) on the host cores over an evenly strided sample of the same rays
    (of every scene, when several share one ray set); median of 5 after 1 warm-up; plus a 1-thread
    run (the reference loop is serial). The sample is sized for about target_s seconds in total."""
    import oracle
    if not oracle.available("fma"):
        return None
    scs = [oracle.Scene.from_scene(s) for s in scenes]
    n = pos.shape[0] * len(scs)

    def run(p, d, th):
        t = time.perf_counter()
        for sc in scs:
            oracle.trace_closest(sc, p, d, threads=th, flavor="fma")
        return time.perf_counter() - t

    probe = min(pos.shape[0], max(1024, 65536 // len(scs)))
    first = run(strided(pos, probe), strided(dir, probe), threads)
    est_full = first * pos.shape[0] / probe
    per_scene = pos.shape[0] if est_full * 6 < target_s else max(probe, int(pos.shape[0] * target_s / (6 * est_full)))
    per_scene = min(per_scene, pos.shape[0])
    p, d = strided(pos, per_scene), strided(dir, per_scene)
    run(p, d, threads)
    times = sorted(run(p, d, threads) for _ in range(5))
    rays = p.shape[0] * len(scs)
    mrays = rays / times[2] / 1e6
    s1 = min(p.shape[0], max(256, 32768 // len(scs)))
    p1, d1 = strided(p, s1), strided(d, s1)
    run(p1, d1, 1)
    t1 = sorted(run(p1, d1, 1) for _ in range(3))[1]
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    scenes_txt = f" in each of {len(scs)} scenes" if len(scs) > 1 else ""
    return {"value": round(mrays, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{p.shape[0]} of {pos.shape[0]} rays{scenes_txt} (every {max(1, pos.shape[0] // p.shape[0])}th), "
                      f"{what}, median of 5 after 1 warm-up, OpenMP dynamic(256); oracle built with the reference "
                      f"flags (-O3, FMA contraction); 1-thread value on {p1.shape[0] * len(scs)} rays",
            "value_1thread": round(p1.shape[0] * len(scs) / t1 / 1e6, 3), "cpu_model": cpu_model}


def load_traffic(workload):
    """HBM bytes per step from the committed counter summary (profiles/traffic_<workload>.txt)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith("hbm_bytes_per_step:"):
                return float(line.split(":", 1)[1])
    return None


def load_counters(workload):
    """Derived ratios of the committed counter summary (profiles/counters_<workload>.txt)."""
    path = os.path.join(ROOT, "profiles", f"counters_{workload}.txt")
    out = {}
    if not os.path.exists(path):
        return out
    with open(path) as fh:
        derived = False
        for line in fh:
            if line.startswith("derived:"):
                derived = True
                continue
            if derived and line.strip():
                key, _, val = line.rstrip().rpartition(" ")
                try:
                    out[key.strip()] = float(val)
                except ValueError:
                    pass
    return out


def limiter(workload, ms, traffic):
    """What the counters say bounds the kernel. The only roofline that applies is HBM (no MFMA:
    traversal is not a contraction); `limiter` names the stage closest to saturation."""
    hbm = None if traffic is None else traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS
    cnt = load_counters(workload)
    ta = cnt.get("TA busy / active cycles (8 XCDs)")
    if hbm is not None and hbm > 0.6:
        name = "hbm bandwidth"
    elif ta is not None and ta > 0.5:
        name = "texture addresser (address expansion of divergent gathers)"
    else:
        name = "latency (dependent node loads on the critical path of the slowest waves)"
    out = {"limiter": name}
    if hbm is not None:
        out["hbm_measured_frac"] = round(hbm, 4)
    if cnt.get("L2 hit rate") is not None:
        out["l2_hit_rate"] = cnt["L2 hit rate"]
    if ta is not None:
        out["ta_busy"] = ta
    return out


def result(workload, n, ms, args, desc, alg_bytes, cpu, unit="Mrays/s"):
    mrays = n / (ms * 1e-3) / 1e6
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    traffic = load_traffic(workload)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic}
    roof.update(limiter(workload, ms, traffic))
    out = {"value": round(mrays, 2), "unit": unit, "ms_per_step": round(ms, 4), "steps": args.steps,
           "warmup": args.warmup, "dtype": "f32", "config": desc, "roofline": roof, "cpu_baseline": cpu}
    if cpu:
        out["speedup_vs_cpu"] = round(mrays / cpu["value"], 1)
    return out


def run_trace(name, args, torch, g, cpu_target_s):
    """A closest-hit workload: scene, rays, counted launch, autotune, timed steps, CPU oracle."""
    threads = min(16, len(os.sched_getaffinity(0)))
    scene, w, h, jitter, desc, kind, build_s = build_workload(name, threads)
    scene.set_launch_config(args.block, args.lds_stack, bool(args.xcd_remap))
    scene.set_tail_config(args.budget, args.tail_waves)
    scene.set_leaf_postponement(args.leaf_postpone)
    scene.set_ray_interleave(args.ray_interleave)
    scene.set_persistent(bool(args.persistent))
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    if isinstance(w, tuple):   # explicit ray buffers
        pos, dirs = (torch.from_numpy(a).cuda() for a in w)
        n = pos.shape[0]
    else:
        n = w * h
        pos = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
        dirs = torch.zeros_like(pos)
        jt = None if jitter is None else torch.from_numpy(jitter).cuda()
        pi, vi = camera()
        g.eye_rays(pi, vi, w, h, pos, dirs, jitter=jt, order=1, stream=sh)
    hits = torch.zeros((n, 8), dtype=torch.int32, device="cuda")

    # instrumented launch (untimed): traversal counters for the algorithmic-bytes model
    counts = torch.zeros((n, 8), dtype=torch.int32, device="cuda")
    scene.trace_closest_counted(pos, dirs, hits, counts, stream=sh)
    torch.cuda.synchronize()
    cnt = counts.cpu().numpy().view(np.uint32)
    del counts
    hits_np = hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)
    alg_bytes = algorithmic_bytes(cnt, hits_np, kind)

    tune = None
    if not args.no_autotune:
        tune = scene.autotune(pos, dirs, hits, stream=sh)
    if args.copy_calibration:
        copy_calibration(torch)

    ms_eager = timed(torch, stream, lambda: scene.trace_closest(pos, dirs, hits, stream=sh), args.steps, args.warmup)
    # the step as a captured HIP graph (both passes; the library's work buffer for the capture
    # stream is made by an eager call first, so the captured call allocates nothing)
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(stream)
    with torch.cuda.stream(side):
        scene.trace_closest(pos, dirs, hits, stream=side.cuda_stream)
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=side):
            scene.trace_closest(pos, dirs, hits, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ms = timed(torch, stream, graph.replay, args.steps, args.warmup)
    # records of the timed (tuned, graph-replayed) launch against the instrumented single-pass one
    replay = hits.cpu().numpy().view(np.uint8).reshape(n, 32)
    differ = int((replay != hits_np.view(np.uint8).reshape(n, 32)).any(axis=1).sum())
    overflow = scene.stack_overflow_count()
    cpu = None
    if cpu_target_s > 0:
        cpu = cpu_baseline([scene], pos.cpu().numpy(), dirs.cpu().numpy(), threads, cpu_target_s)
    launch = ({k: tune[k] for k in ("block", "lds_stack", "xcd_remap", "ray_interleave", "budget", "tail_waves",
                                    "leaf_postpone", "persistent")}
              if tune else
              {"block": args.block, "lds_stack": args.lds_stack, "xcd_remap": args.xcd_remap,
               "ray_interleave": args.ray_interleave, "budget": args.budget, "tail_waves": args.tail_waves,
               "leaf_postpone": args.leaf_postpone, "persistent": args.persistent})
    desc.update({"hit_fraction": round(float((hits_np["primId"] != g.MISS).mean()), 4),
                 "scene_build_s": round(build_s, 2), "launch": launch, "autotuned": tune is not None,
                 "stack_overflow_rays": overflow, "ms_eager": round(ms_eager, 4), "timed": "hip graph replay",
                 "records_differing_from_single_pass": differ,
                 "mean_fat_nodes_per_ray": round(float(cnt[:, 0].mean()), 2),
                 "algorithmic_bytes_per_ray": round(alg_bytes / n, 1)})
    return result(name, n, ms, args, desc, alg_bytes, cpu)


def run_build(args, torch, g):
    """Acceleration-structure construction: the 327,680-triangle BLAS built on the GPU (linear
    BVH, build name gpu_lbvh2) against the host's binned SAH, the primary-ray trace on each tree,
    and the 4096-instance TLAS after moving every instance: GPU refit against host rebuild.
    Wall-clock times of the ABI calls (uploads and downloads included); step = one GPU build."""
    v, f = g.gen_icosphere(7, 0.05, 4.0, 1)
    n_tris = int(f.shape[0])
    t = time.perf_counter()
    sh = g.SceneRT()
    sh.AddInstance(sh.AddGeom_Triangles3f(v, f))
    host_s = time.perf_counter() - t
    host_depth = sh.commit_info()["blas_max_depth"]
    gpu_times = []
    for _ in range(max(args.warmup, 1) + 3):
        t = time.perf_counter()
        sg = g.SceneRT(build_name="gpu_lbvh2")
        gid = sg.AddGeom_Triangles3f(v, f)
        gpu_times.append(time.perf_counter() - t)
    gpu_s = float(np.median(gpu_times[-3:]))
    info = sg.commit_info()
    sg.AddInstance(gid)
    sg.CommitScene()
    sh.CommitScene()
    stream = torch.cuda.current_stream()
    h = stream.cuda_stream
    pos = torch.zeros((512 * 512, 4), dtype=torch.float32, device="cuda")
    dirs = torch.zeros_like(pos)
    pi, vi = camera()
    g.eye_rays(pi, vi, 512, 512, pos, dirs, order=1, stream=h)
    hits = torch.zeros((512 * 512, 8), dtype=torch.int32, device="cuda")
    trace_ms = {}
    for name, sc in (("sah_host", sh), ("lbvh_gpu", sg)):
        if not args.no_autotune:
            sc.autotune(pos, dirs, hits, stream=h)
        trace_ms[name] = round(timed(torch, stream, lambda: sc.trace_closest(pos, dirs, hits, stream=h),
                                     args.steps, args.warmup), 4)
    # TLAS: every instance moved by up to 0.5 units, then one commit
    scene, meshes, mats = tlas_scene()
    rng = np.random.default_rng(44)
    commit_ms = {}
    for mode, label in ((1, "refit_gpu"), (0, "rebuild_host")):
        scene.set_tlas_update(mode)
        ts = []
        for _ in range(3):
            for i in range(mats.shape[0]):
                m = mats[i].copy()
                m[:3, 3] += rng.uniform(-0.5, 0.5, 3).astype(np.float32)
                scene.UpdateInstance(i, m)
            t = time.perf_counter()
            scene.CommitScene()
            ts.append(time.perf_counter() - t)
        commit_ms[label] = round(1e3 * float(np.median(ts)), 3)
        commit_ms[label + "_kind"] = scene.commit_info()["kind"]
    desc = {"workload": "bvh_build_icosphere7", "triangles": n_tris, "builder": "linear BVH, 63-bit Morton",
            "gpu_build_ms": round(1e3 * gpu_s, 2), "host_sah_build_ms": round(1e3 * host_s, 1),
            "gpu_vs_host_build": round(host_s / gpu_s, 1), "blas_leaves": info["blas_leaves"],
            "blas_max_depth": {"lbvh_gpu": info["blas_max_depth"], "sah_host": host_depth},
            "trace_ms_512x512": trace_ms, "tlas_instances": int(mats.shape[0]), "tlas_commit_ms": commit_ms,
            "timed": "wall clock of the ABI call, median of 3 after warm-up"}
    mtris = n_tris / gpu_s / 1e6
    return {"value": round(mtris, 2), "unit": "Mtris/s", "ms_per_step": round(1e3 * gpu_s, 3), "steps": 3,
            "warmup": max(args.warmup, 1), "dtype": "f32", "config": desc, "roofline": None,
            "cpu_baseline": {"value": round(n_tris / host_s / 1e6, 3), "unit": "Mtris/s", "cores": 1,
                             "kind": "port", "sample": "the host binned-SAH build of the same mesh"},
            "speedup_vs_cpu": round(host_s / gpu_s, 1)}


TRACE_WORKLOADS = ["icosphere7_512", "sdf_octree_d9_1024", "sdf_multi_d9_1024", "tlas_4096x32_4m"]
ALL_WORKLOADS = TRACE_WORKLOADS + ["sbs_diff_256_8x256", "sbs_batch_64x4x128", "bvh_build_icosphere7"]


def run_one(name, args, torch, g, cpu_target_s):
    if name == "sbs_diff_256_8x256":
        return run_diff(args, torch, g, cpu_target_s)
    if name == "sbs_batch_64x4x128":
        return run_batch(args, torch, g, cpu_target_s)
    if name == "bvh_build_icosphere7":
        return run_build(args, torch, g)
    return run_trace(name, args, torch, g, cpu_target_s)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="all", choices=["all"] + ALL_WORKLOADS,
                    help="all: the closest-hit headline (icosphere7_512) with every other workload as a "
                         "sub-result of the same line")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--lds-stack", type=int, default=16)
    ap.add_argument("--xcd-remap", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="approximate CPU time of each workload's oracle baseline (all cores)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the given launch configuration")
    ap.add_argument("--budget", type=int, default=64, help="tail step budget when not autotuning")
    ap.add_argument("--tail-waves", type=int, default=1024, help="tail waves when not autotuning")
    ap.add_argument("--ray-interleave", type=int, default=32,
                    help="pass-1 lanes per interleaved group when not autotuning (0: thread order)")
    ap.add_argument("--leaf-postpone", type=int, default=-1, choices=[-1, 0, 1],
                    help="triangle leaf postponement when not autotuning (-1: on for one instance)")
    ap.add_argument("--persistent", type=int, default=0, choices=[0, 1],
                    help="persistent single-pass kernel when not autotuning")
    ap.add_argument("--batch-config", default="",
                    help="lds_stack,record,grid of the batched adjoint kernel instead of its autotuner")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB device copy (reference kernel for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); sharding the rays or scenes over "
              f"several GPUs is designed but not built", file=sys.stderr)
        sys.exit(2)

    import torch
    import graphic_lab_task_amd as g

    torch.cuda.set_device(0)
    cpu_s = 0.0 if args.no_cpu_baseline else args.cpu_seconds
    names = ALL_WORKLOADS if args.workload == "all" else [args.workload]
    head_name = names[0]
    t0 = time.time()
    head = run_one(head_name, args, torch, g, cpu_s)
    print(f"# {head_name}: {head['value']} {head['unit']} ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)
    subs = {}
    for name in names[1:]:
        torch.cuda.empty_cache()
        t1 = time.time()
        subs[name] = run_one(name, args, torch, g, cpu_s)
        print(f"# {name}: {subs[name]['value']} {subs[name]['unit']} ({time.time() - t1:.0f} s)", file=sys.stderr,
              flush=True)
    out = {
        "metric": METRIC, "value": head["value"], "unit": head["unit"], "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": head["config"],
        "roofline": head["roofline"], "cpu_baseline": head["cpu_baseline"],
    }
    if "speedup_vs_cpu" in head:
        out["speedup_vs_cpu"] = head["speedup_vs_cpu"]
    if subs:
        out["sub_results"] = subs
    print(json.dumps(out))


if __name__ == "__main__":
    main()
