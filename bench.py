#!/usr/bin/env python3
"""Headline benchmark: incoherent closest-hit throughput on the 1M-triangle
scene (Mrays/s), one MI355X.

Workload: 2^20 triangles (centroids uniform in the unit cube, vertex offsets
Gaussian with lognormal sigma, median 0.005), 2^24 rays from a radius-2 sphere
towards uniform points of the cube, SAH BVH with <= 4 primitives per leaf,
FP32. A step = one closest-hit launch over all 2^24 rays with rays and BVH
already resident in HBM. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
N_TRIS = 1 << 20
N_RAYS = 1 << 24


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(v, idx, rays, n_nodes_budget_s=15.0):
    """The oracle's pbrt-v3 BVH timed on host cores over a bounded ray sample
    of the same workload (BVH build excluded)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.time()
    sc = pyoracle.OracleScene(v, idx, None, 4)
    build_s = time.time() - t0
    n = 1 << 16
    sample = rays[:: len(rays) // n][:n]
    t0 = time.time()
    sc.intersect(sample, n_threads=threads)
    dt = time.time() - t0
    # scale the sample so that the timed CPU work is ~10-20 s
    target = min(len(rays), max(n, int(n * n_nodes_budget_s / max(dt, 1e-3))))
    target = 1 << int(np.log2(target))
    if target > n:
        sample = rays[:: len(rays) // target][:target]
        t0 = time.time()
        sc.intersect(sample, n_threads=threads)
        dt = time.time() - t0
    return {"value": len(sample) / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": "%d of the %d rays (strided), closest hit, oracle pbrt-v3 SAH BVH (build %.1f s excluded)"
                      % (len(sample), len(rays), build_s)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--rays", type=int, default=N_RAYS)
    ap.add_argument("--width", type=int, default=2, choices=(2, 4), help="BVH node width")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (rays shard with no exchange; "
              "see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    from rs_illume_amd import workloads as W
    from rs_illume_amd.accel import BVHAccel, rays_to_device

    log("generating workload")
    v, idx = W.triangle_soup(N_TRIS, seed=1)
    rays = W.sphere_to_cube_rays(args.rays, seed=2)
    log("building BVH")
    bvh = BVHAccel(v, idx, None, 4, width=args.width)
    info = bvh.info()
    log("bvh", info)
    rd = rays_to_device(rays)
    n = rd.shape[0]
    hits = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()

    # algorithmic work per ray from the instrumented traversal (outside the timed region)
    counts = bvh.intersect_counts(rd).double().mean(0).cpu().numpy()
    box_tests, prim_tests = float(counts[0]), float(counts[1])
    bytes_per_ray = 32 + 8 + 32 * box_tests + 36 * prim_tests

    for _ in range(args.warmup):
        bvh.intersect(rd, hits)
    torch.cuda.synchronize()

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        bvh.intersect(rd, hits)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = gpu_ms / args.steps
    mrays = n / (ms_per_step * 1e-3) / 1e6
    achieved_gbs = bytes_per_ray * n / (ms_per_step * 1e-3) / 1e9

    hit_frac = float(((hits[:, 0].to(torch.int64) & 0xFFFFFFFF) != 0xFFFFFFFF).float().mean())
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(v, idx, rays)

    out = {
        "metric": "Mrays/s closest-hit; time-to-relMSE 0.01, guided vs plain GPU path tracer",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "incoherent closest-hit, 1M-triangle scene (2^20 tris, lognormal sigma median "
                               "0.005), 2^24 rays sphere->cube, SAH BVH <=4 prims/leaf",
                   "n_tris": N_TRIS, "n_rays": n, "max_prims_in_node": 4, "bvh_width": args.width,
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_ray": round(bytes_per_ray, 1), "box_tests_per_ray": round(box_tests, 2),
                     "prim_tests_per_ray": round(prim_tests, 2)},
        "cpu_baseline": cpu,
        "hit_fraction": round(hit_frac, 4),
        "bvh": {"inner_nodes": info["n_inner_nodes"], "max_depth": info["max_depth"],
                "device_mb": round(info["device_bytes"] / 2 ** 20, 1), "build_s": round(info["build_seconds"], 2)},
        "wall_s_timed": round(wall, 4),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
