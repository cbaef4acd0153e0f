#!/usr/bin/env python3
"""Ray-query benchmark on one MI355X.

A step is one closest-hit launch over the workload's full ray batch, inputs resident in HBM.
Prints one JSON line (see DESIGN.md for the metric, the roofline model and the CPU baseline).

Workloads (synthetic, seeded):
  sdf_octree_d9_1024  64-sphere union, frame octree depth 9, 1024x1024 jittered primary rays
  icosphere7_512      icosphere subdivided 7x (327,680 tris), radial value noise 0.05, 512x512 rays
  sbs_diff_256_8x256  differentiable brick set: 64-sphere union (seed 4) on a 256^3 grid of 8^3
                      fp32 bricks with per-lattice-point RGB albedo; 8 ring views of 256x256; a
                      step = forward render + adjoint of the L2 loss against the image of a second
                      scene (seed 5); rays/s counts each ray once per step
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
    raise SystemExit(f"unknown workload {name}")


def run_diff(args, torch, g):
    """Forward + adjoint steps of the differentiable brick-set workload (one JSON line)."""
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

    # traversal counters of the same rays (untimed) for the algorithmic-bytes model
    hits = torch.zeros((n, 8), dtype=torch.int32, device="cuda")
    counts = torch.zeros((n, 8), dtype=torch.int32, device="cuda")
    scene.trace_closest_counted(pos, dirs, hits, counts, stream=sh)
    torch.cuda.synchronize()
    c = counts.cpu().numpy().view(np.uint32).astype(np.float64)
    hit_n = float((hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)["primId"] != g.MISS).sum())
    trav = B_FAT_NODE * c[:, 0].sum() + B_TLAS_NODE * c[:, 3].sum() + B_LEAF * c[:, 1].sum() + \
        B_BRICK * c[:, 1].sum() + B_VOXEL * c[:, 5].sum() + B_INSTANCE * c[:, 4].sum()
    # forward: rays in, colour out, traversal; backward: rays + target in, traversal again, the hit
    # brick re-marched (bounded by one brick's voxels), 32 parameters read and 32 atomics per hit
    alg_bytes = n * (32 + 16) + trav + n * (32 + 16) + trav + hit_n * (B_VOXEL * 8 + 32 * 4 + 32 * 8)

    def step():
        grad.zero_()
        loss.zero_()
        scene.sbs_forward(gid, params, pos, dirs, colour, stream=sh)
        scene.sbs_l2_backward(gid, params, pos, dirs, target, grad, loss, stream=sh)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        step()
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.steps
    mrays = n / (ms * 1e-3) / 1e6
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    desc = {"workload": "sbs_diff_256_8x256", "virtual_grid": 256, "brick": 8, "bricks": int(n_bricks),
            "distance_values": int(nd), "views": 8, "width": 256, "height": 256, "rays": int(n),
            "hit_fraction": round(hit_n / n, 4), "scene_build_s": round(build_s, 2),
            "recompute_misses": scene.sbs_recompute_misses(), "loss": float(loss.item()),
            "algorithmic_bytes_per_ray": round(alg_bytes / n, 1)}
    return {
        "metric": METRIC, "value": round(mrays, 2), "unit": "Mrays/s (fwd+bwd)", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": desc,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic("sbs_diff_256_8x256")},
        "cpu_baseline": None,
    }


def algorithmic_bytes(cnt, hits, kind):
    import graphic_lab_task_amd as g
    c = cnt.astype(np.float64)
    b = B_RAY_IO * c.shape[0] + B_FAT_NODE * c[:, 0].sum() + B_TLAS_NODE * c[:, 3].sum() + B_LEAF * c[:, 1].sum()
    b += B_TRI * c[:, 2].sum() + B_INSTANCE * c[:, 4].sum()
    if kind == "frame":
        b += B_FRAME_LEAF * c[:, 1].sum()
    if kind == "tri":
        b += B_REMAP * float((hits["primId"] != g.MISS).sum())
    return b


def strided(a, k):
    """Every n/k-th ray: a sample spread over the whole image (rays are in screen-tile order)."""
    step = max(1, a.shape[0] // k)
    return np.ascontiguousarray(a[::step][:k])


def cpu_baseline(scene, pos, dir, threads):
    """Oracle (reference flags) on the host cores over an evenly strided sample of the same rays;
    median of 5 after 1 warm-up; plus a 1-thread run (the reference loop is serial)."""
    import oracle
    if not oracle.available("fma"):
        return None
    sc = oracle.Scene.from_scene(scene)
    n = pos.shape[0]

    def run(p, d, th):
        t = time.perf_counter()
        oracle.trace_closest(sc, p, d, threads=th, flavor="fma")
        return time.perf_counter() - t

    probe = min(n, 65536)
    first = run(strided(pos, probe), strided(dir, probe), threads)
    est_full = first * n / probe
    sample = n if est_full * 6 < 20.0 else max(probe, int(n * 20.0 / (6 * est_full)))
    sample = min(sample, n)
    p, d = strided(pos, sample), strided(dir, sample)
    run(p, d, threads)
    times = sorted(run(p, d, threads) for _ in range(5))
    mrays = p.shape[0] / times[2] / 1e6
    s1 = min(p.shape[0], 32768)
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
    return {"value": round(mrays, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{p.shape[0]} of {n} rays of the same workload (every {max(1, n // p.shape[0])}th), median "
                      f"of 5 after 1 warm-up, OpenMP dynamic(256); oracle built with the reference flags (-O3, FMA "
                      f"contraction); 1-thread value on {p1.shape[0]} of them",
            "value_1thread": round(p1.shape[0] / t1 / 1e6, 3), "cpu_model": cpu_model}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="sdf_octree_d9_1024",
                    choices=["sdf_octree_d9_1024", "icosphere7_512", "sbs_diff_256_8x256"])
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--lds-stack", type=int, default=16)
    ap.add_argument("--xcd-remap", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the given launch configuration")
    ap.add_argument("--budget", type=int, default=64, help="tail step budget when not autotuning")
    ap.add_argument("--tail-waves", type=int, default=1024, help="tail waves when not autotuning")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB device copy (reference kernel for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the sharded multi-GPU path is "
              f"designed in DESIGN.md but not built", file=sys.stderr)
        sys.exit(2)

    import torch
    import graphic_lab_task_amd as g

    threads = min(16, len(os.sched_getaffinity(0)))
    torch.cuda.set_device(0)
    if args.workload == "sbs_diff_256_8x256":
        print(json.dumps(run_diff(args, torch, g)))
        return
    scene, w, h, jitter, desc, kind, build_s = build_workload(args.workload, threads)
    scene.set_launch_config(args.block, args.lds_stack, bool(args.xcd_remap))
    scene.set_tail_config(args.budget, args.tail_waves)
    n = w * h
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    pos = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    dirs = torch.zeros_like(pos)
    hits = torch.zeros((n, 8), dtype=torch.int32, device="cuda")
    jt = None if jitter is None else torch.from_numpy(jitter).cuda()
    pi, vi = camera()
    g.eye_rays(pi, vi, w, h, pos, dirs, jitter=jt, order=1, stream=sh)

    # instrumented launch (untimed): traversal counters for the algorithmic-bytes model
    counts = torch.zeros((n, 8), dtype=torch.int32, device="cuda")
    scene.trace_closest_counted(pos, dirs, hits, counts, stream=sh)
    torch.cuda.synchronize()
    cnt = counts.cpu().numpy().view(np.uint32)
    hits_np = hits.cpu().numpy().view(g.HIT_DTYPE).reshape(-1)
    alg_bytes = algorithmic_bytes(cnt, hits_np, kind)

    tune = None
    if not args.no_autotune:
        tune = scene.autotune(pos, dirs, hits, stream=sh)
    if args.copy_calibration:
        src = torch.empty(64 << 20, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)   # one elementwise kernel: reads 256 MiB, writes 256 MiB
        torch.cuda.synchronize()
        del src, dst

    for _ in range(args.warmup):
        scene.trace_closest(pos, dirs, hits, stream=sh)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        scene.trace_closest(pos, dirs, hits, stream=sh)
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.steps
    mrays = n / (ms * 1e-3) / 1e6
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    overflow = scene.stack_overflow_count()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene, pos.cpu().numpy(), dirs.cpu().numpy(), threads)

    launch = ({k: tune[k] for k in ("block", "lds_stack", "xcd_remap", "budget", "tail_waves")} if tune else
              {"block": args.block, "lds_stack": args.lds_stack, "xcd_remap": args.xcd_remap, "budget": args.budget,
               "tail_waves": args.tail_waves})
    desc.update({"hit_fraction": round(float((hits_np["primId"] != g.MISS).mean()), 4),
                 "scene_build_s": round(build_s, 2), "launch": launch, "autotuned": tune is not None,
                 "stack_overflow_rays": overflow,
                 "mean_fat_nodes_per_ray": round(float(cnt[:, 0].mean()), 2),
                 "algorithmic_bytes_per_ray": round(alg_bytes / n, 1)})
    out = {
        "metric": METRIC, "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": desc,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(args.workload)},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(mrays / cpu["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
