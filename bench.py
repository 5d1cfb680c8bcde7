#!/usr/bin/env python3
"""Headline benchmark: incoherent closest-hit throughput on the 1M-triangle
scene (Mrays/s), one MI355X, plus the path-tracing measurements that share the
metric line.

Headline workload: 2^20 triangles (centroids uniform in the unit cube, vertex
offsets Gaussian with lognormal sigma, median 0.005), 2^24 rays from a
radius-2 sphere towards uniform points of the cube, SAH BVH with <= 4
primitives per leaf, FP32. A step = one closest-hit launch over all 2^24 rays
with rays and BVH already resident in HBM.

Also measured (keys of the same JSON line):
  cornell  - Cornell box 256x256, 16 spp, depth 5: GPU samples/s and rays/s
             (megakernel), and the oracle's CPU path tracer on host cores;
  guided   - glossy small-light Cornell box 512x512, depth 8: wall time until
             relMSE <= 0.01 against a 16384-spp plain render, plain GPU path
             tracer vs the guided one (online training included in its time);
  guided_indirect - the same comparison in a box lit only indirectly (its
             light facing the ceiling), where the proposal pays;
  mlp      - the proposal network's MFMA kernels: inference and training
             throughput as a fraction of the bf16 dense peak;
  batched  - 64 procedural scenes of uneven size in one batched call: the
             wavefront pool with the autotuned launch shape and the
             megakernel batch, the faster one reported on top;
  combined - terrain + spheres under the analytic sky, 16 views batched,
             plain vs guided time-to-relMSE per view and samples/s.

Roofline of the headline kernel: the HBM-side bytes per closest-hit launch
(FETCH_SIZE + WRITE_SIZE of the traversal kernel, measured in this run by
child rocprofv3 --pmc passes before the GPU is touched, with the L2 hit rate)
over the launch time from HIP events; the algorithmic bytes per ray (the
kernel's own 64-byte nodes and 48-byte primitive records, and the reference's
pbrt layout) are reported beside it under their own names.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-render] [--no-counters]
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
BF16_PEAK_TFLOPS = 2500.0
N_TRIS = 1 << 20
N_RAYS = 1 << 24
MLP_INFER_FLOP = 2 * (144 * 64 + 64 * 64 + 64 * 64 + 64 * 40)  # 39,936 per sample
MLP_TRAIN_FLOP = 3 * MLP_INFER_FLOP                              # forward + two backward products


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _oracle():
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle
    return pyoracle


def host_cpu():
    """(threads, model): every host thread this process may run on, and the
    CPU model, for the CPU-baseline legs."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return max(1, n), model


def cpu_threads():
    return host_cpu()[0]


def cpu_progressive(osc, cam, ref, max_depth, seed, spp_pass, budget_s, target=0.01):
    """The oracle's path tracer on host cores in progressive passes of
    spp_pass samples per pixel until the wall budget is spent: the time and spp
    at which relMSE against `ref` (host (H, W, 3)) first reaches `target`, and
    the relMSE at the budget."""
    from rs_illume_amd.scenes import rel_mse
    threads = cpu_threads()
    acc = None
    spp, elapsed, reached = 0, 0.0, None
    while elapsed < budget_s:
        t0 = time.time()
        s, _ = osc.render_sum(cam, spp_pass, max_depth, seed=seed, sample_offset=spp, n_threads=threads)
        elapsed += time.time() - t0
        acc = s.astype(np.float64) if acc is None else acc + s
        spp += spp_pass
        err = rel_mse(acc / spp, ref)
        if reached is None and err <= target:
            reached = (elapsed, spp)
    w, h = int(cam["width"][0]), int(cam["height"][0])
    return {"seconds": round(elapsed, 2), "spp": spp, "rel_mse": round(float(err), 6),
            "samples_per_s": round(w * h * spp / elapsed, 1),
            "time_to_target_s": round(reached[0], 3) if reached else None,
            "spp_to_target": reached[1] if reached else None, "cores": threads, "cpu_model": host_cpu()[1],
            "kind": "port"}


def cpu_baseline(v, idx, rays, budget_s=15.0):
    """The oracle's pbrt-v3 BVH timed on host cores over a bounded ray sample
    of the same workload (BVH build excluded). Also returns the oracle's mean
    node and primitive tests per ray on that sample: the per-ray work figure
    of the algorithmic-bytes roofline."""
    pyoracle = _oracle()
    threads = cpu_threads()
    t0 = time.time()
    sc = pyoracle.OracleScene(v, idx, None, 4)
    build_s = time.time() - t0
    n = 1 << 16
    sample = rays[:: len(rays) // n][:n]
    t0 = time.time()
    _, st = sc.intersect(sample, n_threads=threads, with_stats=True)
    dt = time.time() - t0
    # scale the sample so that the timed CPU work is ~10-20 s
    target = min(len(rays), max(n, int(n * budget_s / max(dt, 1e-3))))
    target = 1 << int(np.log2(target))
    if target > n:
        sample = rays[:: len(rays) // target][:target]
        t0 = time.time()
        _, st = sc.intersect(sample, n_threads=threads, with_stats=True)
        dt = time.time() - t0
    counts = (float(st[0]) / len(sample), float(st[1]) / len(sample))
    return {"value": round(len(sample) / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads,
            "cpu_model": host_cpu()[1], "kind": "port",
            "sample": "%d of the %d rays (strided), closest hit, oracle pbrt-v3 SAH BVH (build %.1f s excluded)"
                      % (len(sample), len(rays), build_s)}, counts


def bench_cornell(no_cpu):
    """The Cornell box (256x256, 16 spp, depth 5) on the megakernel path tracer; the oracle renders
    a bounded number of samples per pixel on host cores."""
    import torch
    from rs_illume_amd import scenes as S
    from rs_illume_amd.render import PathIntegrator, Scene
    desc = S.cornell_box(0)
    sc = Scene(desc)
    cam = S.cornell_camera(256, 256)
    spp = 16
    pi = PathIntegrator(cam, spp=spp, max_depth=5, seed=0)
    pi.render_sum(sc)
    counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    reps = 10
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        pi.render_sum(sc, ray_counts=counts)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    rays = counts.cpu().numpy().astype(np.float64) / reps
    out = {"workload": "Cornell box 256x256, 16 spp, max depth 5 (10 triangles, 2 spheres, 0.25^2 light)",
           "samples_per_s": round(256 * 256 * spp / dt, 1), "mrays_per_s": round(float(rays.sum()) / dt / 1e6, 1),
           "ms_per_frame": round(dt * 1e3, 3), "closest_rays": int(rays[0]), "any_rays": int(rays[1])}
    # the same frame through pbrt-v3's default Gaussian filter (radius 2): samples over the
    # 260x260 sample bounds, each splatted into up to 16 pixels
    from rs_illume_amd.film import Filter, render_filtered
    filt = Filter("gaussian")
    rgbw = render_filtered(pi, sc, filt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        render_filtered(pi, sc, filt, rgbw=rgbw)
    torch.cuda.synchronize()
    dtf = (time.perf_counter() - t0) / reps
    out["gaussian_filter"] = {"samples_per_s": round(260 * 260 * spp / dtf, 1), "ms_per_frame": round(dtf * 1e3, 3)}
    # the same frame with pbrt-v3's Halton sampler (scrambled radical inverses per dimension)
    ph = PathIntegrator(cam, spp=spp, max_depth=5, seed=0, sampler="halton")
    acc = ph.render_sum(sc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ph.render_sum(sc, rgb_sum=acc)
    torch.cuda.synchronize()
    dth = (time.perf_counter() - t0) / reps
    out["halton_sampler"] = {"samples_per_s": round(256 * 256 * spp / dth, 1), "ms_per_frame": round(dth * 1e3, 3)}
    # and with pbrt-v3's Sobol' sampler
    ps = PathIntegrator(cam, spp=spp, max_depth=5, seed=0, sampler="sobol")
    acc = ps.render_sum(sc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ps.render_sum(sc, rgb_sum=acc)
    torch.cuda.synchronize()
    dts = (time.perf_counter() - t0) / reps
    out["sobol_sampler"] = {"samples_per_s": round(256 * 256 * spp / dts, 1), "ms_per_frame": round(dts * 1e3, 3)}
    # the reference's AOIntegrator on the same frame: 16 cosine-distributed
    # directions per camera sample
    from rs_illume_amd.render import AOIntegrator
    ao = AOIntegrator(cam, spp=spp, n_samples=16, cos_sample=True, seed=0)
    acc = ao.render_sum(sc)
    ao_counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ao.render_sum(sc, rgb_sum=acc, ray_counts=ao_counts)
    torch.cuda.synchronize()
    dta = (time.perf_counter() - t0) / reps
    ao_rays = float(ao_counts.sum().item()) / reps
    out["ao_integrator"] = {"samples_per_s": round(256 * 256 * spp / dta, 1), "ms_per_frame": round(dta * 1e3, 3),
                            "mrays_per_s": round(ao_rays / dta / 1e6, 1), "ao_samples": 16}
    if not no_cpu:
        pyoracle = _oracle()
        osc = pyoracle.OracleRenderScene(desc.arrays())
        threads = cpu_threads()
        t0 = time.time()
        _, st = osc.render_sum(cam, 1, 5, seed=0, n_threads=threads)
        dt1 = time.time() - t0
        cspp = int(max(1, min(spp, 10.0 / max(dt1, 1e-3))))
        t0 = time.time()
        _, st = osc.render_sum(cam, cspp, 5, seed=0, n_threads=threads)
        dtc = time.time() - t0
        out["cpu_baseline"] = {"samples_per_s": round(256 * 256 * cspp / dtc, 1),
                               "mrays_per_s": round(float(st.sum()) / dtc / 1e6, 3), "cores": threads,
                               "cpu_model": host_cpu()[1],
                               "kind": "port", "sample": "%d spp of the 256x256 frame" % cspp}
    sc.close()
    return out


def bench_guided(ref_spp=16384, res=512, log=log, train_steps=4, warmup_steps=32, plain="megakernel", no_cpu=False,
                 cpu_budget_s=20.0, guide_frac=0.25, product=True):
    """Glossy small-light Cornell box (512x512, depth 8): time to relMSE 0.01,
    guided against plain (rs_illume_amd.converge); the CPU
    leg runs the oracle's path tracer on host cores in 2-spp passes for a fixed
    wall budget against the same reference (time to relMSE 0.01 and the relMSE
    at the budget)."""
    import torch
    from rs_illume_amd import converge as C
    from rs_illume_amd import scenes as S
    from rs_illume_amd.render import PathIntegrator, Scene
    desc = S.cornell_box(0, glossy=True, light_size=0.2, light_L=26.5625)
    sc = Scene(desc)
    cam = S.cornell_camera(res, res)
    pi = PathIntegrator(cam, spp=256, max_depth=8, seed=987654321)
    ref = torch.zeros((res, res, 3), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(ref_spp // 256):
        pi.render_sum(sc, rgb_sum=ref, spp=256, sample_offset=256 * k)
    torch.cuda.synchronize()
    ref_s = time.perf_counter() - t0
    ref /= ref_spp
    log("reference %d spp in %.1f s" % (ref_spp, ref_s))
    r = C.guided_vs_plain(sc, cam, ref, target=0.01, spp_per_pass=4, max_spp=4096, max_depth=8, seed=1,
                          train_steps=train_steps, warmup_steps=warmup_steps, plain=plain, guide_frac=guide_frac,
                          product=product)
    out = {"workload": "glossy small-light Cornell box %dx%d, max depth 8, TR copper walls alpha~U[0.05,0.5], "
                       "light 1%% of the ceiling; relMSE vs %d-spp plain render" % (res, res, ref_spp),
           "target_rel_mse": 0.01, "reference_spp": ref_spp, "reference_s": round(ref_s, 2),
           "spp_per_pass": r["spp_per_pass"], "plain_renderer": r["plain_renderer"], "train_steps": train_steps,
           "warmup_steps": warmup_steps, "guide_frac": guide_frac, "product_with_bsdf_lobe": product}
    for k in ("plain", "guided"):
        rr = r[k]["reached"]
        out[k] = {"seconds": round(rr["seconds"], 5), "spp": rr["spp"], "rel_mse": round(rr["rel_mse"], 5)} \
            if rr else {"seconds": None, "spp": r[k]["spp"], "rel_mse": round(r[k]["curve"][-1][2], 5)}
    out["time_ratio_guided_over_plain"] = round(r["time_ratio"], 3) if "time_ratio" in r else None
    # the measured bound: a proposal trained long, then frozen, at equal spp
    log("guided: frozen-proposal ceiling")
    out["frozen_ceiling"] = C.frozen_ceiling(sc, cam, ref, guide_frac=guide_frac, product=product)
    if not no_cpu:
        log("guided: CPU leg (oracle, %.0f s budget)" % cpu_budget_s)
        osc = _oracle().OracleRenderScene(desc.arrays())
        cpu = cpu_progressive(osc, cam, ref.cpu().numpy(), 8, 1, 2, cpu_budget_s)
        cpu["sample"] = "the whole %dx%d frame in 2-spp passes for a %.0f s wall budget" % (res, res, cpu_budget_s)
        out["cpu_baseline"] = cpu
    sc.close()
    return out


def bench_guided_indirect(ref_spp=16384, res=512, log=log, spp_per_pass=16, warmup_steps=16, train_steps=8,
                          train_budget=256, guide_frac=0.5):
    """The proposal where guiding pays: the Cornell box with its light turned
    over (facing the ceiling, L = 60), so every surface but the ceiling is lit
    only indirectly, 512x512, max depth 8. Time to relMSE 0.01 against a
    16384-spp plain render, plain megakernel vs the guided wavefront tracer
    (trained online from scratch: BSDF-only passes until 16 Adam steps, then
    one-sample MIS at fraction 0.5; training stops after train_budget steps and
    its time is included)."""
    import torch
    from rs_illume_amd import converge as C
    from rs_illume_amd import scenes as S
    from rs_illume_amd.render import PathIntegrator, Scene
    sc = Scene(S.cornell_box(0, light_up=True, light_L=60.0))
    cam = S.cornell_camera(res, res)
    pi = PathIntegrator(cam, spp=256, max_depth=8, seed=987654321)
    ref = torch.zeros((res, res, 3), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(ref_spp // 256):
        pi.render_sum(sc, rgb_sum=ref, spp=256, sample_offset=256 * k)
    torch.cuda.synchronize()
    ref_s = time.perf_counter() - t0
    ref /= ref_spp
    log("indirect room: reference %d spp in %.1f s" % (ref_spp, ref_s))
    r = C.guided_vs_plain(sc, cam, ref, target=0.01, spp_per_pass=spp_per_pass, max_spp=16384, max_depth=8, seed=1,
                          train_steps=train_steps, warmup_steps=warmup_steps, plain="megakernel",
                          guide_frac=guide_frac, product=False, train_budget=train_budget)
    out = {"workload": "Cornell box %dx%d lit only indirectly (the 0.25^2 light at y = 0.9 facing the ceiling, "
                       "L = 60), Lambertian, max depth 8; relMSE vs %d-spp plain render" % (res, res, ref_spp),
           "target_rel_mse": 0.01, "reference_spp": ref_spp, "reference_s": round(ref_s, 2),
           "spp_per_pass": spp_per_pass, "plain_renderer": "megakernel", "warmup_steps": warmup_steps,
           "train_steps_per_replay": train_steps, "train_budget": train_budget, "adam_steps": r["adam_steps"],
           "guide_frac": guide_frac}
    for k in ("plain", "guided"):
        rr = r[k]["reached"]
        out[k] = {"seconds": round(rr["seconds"], 4), "spp": rr["spp"], "rel_mse": round(rr["rel_mse"], 5)} \
            if rr else {"seconds": None, "spp": r[k]["spp"], "rel_mse": round(r[k]["curve"][-1][2], 5)}
        out[k]["seconds_per_spp_ms"] = round(1e3 * r[k]["seconds"] / max(r[k]["spp"], 1), 4)
    out["time_ratio_guided_over_plain"] = round(r["time_ratio"], 3) if "time_ratio" in r else None
    out["spp_ratio_guided_over_plain"] = round(out["guided"]["spp"] / out["plain"]["spp"], 3) \
        if r["plain"]["reached"] and r["guided"]["reached"] else None
    sc.close()
    return out


def bench_batched(no_cpu, log=log, builder="SAH_GPU", pmc=None):
    """Batched workload: 64 procedural scenes (seeds 1-64) in one batched call, twice:
    through one wavefront pool (launch shape chosen by the autotuner on a 2-spp
    probe of the same batch) and through the megakernel batch
    (illume_render_batch); the faster is the config's figure. The oracle
    renders scene 1 at 1 spp on host cores."""
    import torch
    from rs_illume_amd import autotune
    from rs_illume_amd import scenes as S
    from rs_illume_amd.render import Scene
    t0 = time.time()
    descs = [S.procedural_scene(seed) for seed in range(1, 65)]
    gen_s = time.time() - t0
    # scene set-up from 8 host threads (each scene's tree is built on a stream
    # of its own, so the device builds overlap; the reference builds its BVHs
    # with rayon)
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.time()
    with ThreadPoolExecutor(8) as ex:
        scenes = list(ex.map(lambda d: Scene(d[0], builder=builder), descs))
    setup_s = time.time() - t0
    bvh_s = sum(sc.bvh_info()["build_seconds"] for sc in scenes)
    jobs = [(sc, cam, spp, 8, seed) for seed, (sc, (_, cam, spp)) in enumerate(zip(scenes, descs), 1)]
    n_tris = sum(len(d.idx) for d, _, _ in descs)
    n_px = sum(int(c["width"][0]) * int(c["height"][0]) for _, c, _ in descs)
    n_paths = sum(int(c["width"][0]) * int(c["height"][0]) * spp for _, c, spp in descs)
    log("batched: %d scenes, %.1fM tris, %.1fM px, %.0fM paths, generation %.1f s, scene set-up (%s) %.1f s"
        % (len(jobs), n_tris / 1e6, n_px / 1e6, n_paths / 1e6, gen_s, builder, setup_s))
    t0 = time.time()
    best, table = autotune.tune(jobs, max_depth=8, sample_spp=2)
    tune_s = time.time() - t0
    log("autotune %.1f s: %s" % (tune_s, best))
    wf = autotune.make_wavefront(best, max_depth=8)
    counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    wf.render_batch(jobs[:1])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    film, _ = wf.render_batch(jobs, ray_counts=counts)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    rays = counts.cpu().numpy().astype(np.float64)
    wave = {"seconds": round(dt, 4), "samples_per_s": round(n_paths / dt, 1),
            "mrays_per_s": round(float(rays.sum()) / dt / 1e6, 1), "iterations": wf.last_iterations(),
            "autotune": {"best": {k: best[k] for k in ("max_paths", "block", "regen_frac", "graph_iters")},
                         "candidates": len(table), "seconds": round(tune_s, 1)},
            "film_finite": bool(torch.isfinite(film).all())}
    wf.close()
    del film
    # the same jobs on the megakernel in one launch (illume_render_batch)
    from rs_illume_amd.render import render_batch
    render_batch(jobs[:1])
    mk_counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    film, _ = render_batch(jobs, ray_counts=mk_counts)
    torch.cuda.synchronize()
    dt_mk = time.perf_counter() - t0
    mk_rays = mk_counts.cpu().numpy().astype(np.float64)
    mega = {"seconds": round(dt_mk, 4), "samples_per_s": round(n_paths / dt_mk, 1),
            "mrays_per_s": round(float(mk_rays.sum()) / dt_mk / 1e6, 1),
            "film_finite": bool(torch.isfinite(film).all())}
    del film
    if pmc:
        # HBM-side bytes of the batched render kernel from the counter passes
        # (same scenes, trees and launch), over this run's launch time
        gbs = pmc["traffic_bytes"] / dt_mk / 1e9
        mega["roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": round(pmc["traffic_bytes"]),
                            "fetch_bytes": round(pmc["fetch_bytes"]), "write_bytes": round(pmc["write_bytes"]),
                            "l2_hit_rate": round(pmc["l2_hit_rate"], 4), "kernel": pmc["kernel"],
                            "bytes_per_ray": round(pmc["traffic_bytes"] / max(float(mk_rays.sum()), 1.0), 1),
                            "basis": "FETCH_SIZE (raw) + WRITE_SIZE of one render_batch_kernel launch from "
                                     "rocprofv3 --pmc passes in this run, over the timed launch"}
    fastest = "megakernel_batch" if dt_mk < dt else "wavefront"
    top = mega if dt_mk < dt else wave
    out = {"workload": "64 procedural scenes (seeds 1-64): displaced icospheres + ground quad, Lambertian/dielectric, "
                       "1-4 quad lights, 64^2..1024^2, 8-64 spp, max depth 8, one batched render call",
           "scenes": len(jobs), "triangles": n_tris, "pixels": n_px, "paths": n_paths, "bvh_builder": builder,
           "scene_setup_s": round(setup_s, 2), "bvh_build_sum_s": round(bvh_s, 3),
           "bvh_build_sum_basis": "the 64 builds' own durations added up (they overlap on 8 host threads)",
           "end_to_end_s": round(setup_s + top["seconds"], 3),
           "end_to_end_basis": "scene set-up from 8 host threads (scene tables, device BVH builds on a stream per "
                               "scene) + the batched render; the numpy scene generation and the autotuner's probe "
                               "excluded", "renderer": fastest,
           "seconds": top["seconds"], "samples_per_s": top["samples_per_s"], "mrays_per_s": top["mrays_per_s"],
           "wavefront": wave, "megakernel_batch": mega}
    if not no_cpu:
        # the oracle renders whole scenes (their resolution and spp) on host
        # cores in seed order, scenes 1-4 at least, until a 15 s render budget
        # is spent; the rest of the batch's CPU time is extrapolated at that rate
        pyoracle = _oracle()
        threads = cpu_threads()
        paths, secs, rays, build, done = 0, 0.0, 0.0, 0.0, 0
        for seed in range(1, len(descs) + 1):
            if seed > 4 and secs >= 15.0:
                break
            d, cam, spp = descs[seed - 1]
            t0 = time.time()
            osc = pyoracle.OracleRenderScene(d.arrays())
            build += time.time() - t0
            t0 = time.time()
            _, st = osc.render_sum(cam, spp, 8, seed=seed, n_threads=threads)
            secs += time.time() - t0
            paths += int(cam["width"][0]) * int(cam["height"][0]) * spp
            rays += float(st.sum())
            done = seed
            del osc
        sps = paths / secs
        out["cpu_baseline"] = {"samples_per_s": round(sps, 1), "mrays_per_s": round(rays / secs / 1e6, 3),
                               "cores": threads, "cpu_model": host_cpu()[1], "kind": "port",
                               "sample": "scenes 1-%d in full (%d of the batch's %d paths, %.1f s; oracle BVH "
                                         "builds %.1f s excluded)" % (done, paths, n_paths, secs, build),
                               "batch_seconds": round(secs + (n_paths - paths) / sps, 1),
                               "batch_seconds_basis": "measured" if done == len(descs) else
                               "scenes 1-%d measured, the rest extrapolated at their rate" % done}
    for sc in scenes:
        sc.close()
    return out


def bench_combined(ref_spp=4096, spp_pass=2, max_spp=256, target=0.01, log=log, train_steps=4, warmup_steps=32,
                   builder="SAH_GPU", no_cpu=False, cpu_budget_s=20.0, guide_frac=0.25):
    """Combined workload: fBm terrain (2M triangles) + 10k spheres under the analytic
    sky, 16 views of uneven size in one batched call per pass, max depth 10.
    Plain (on the megakernel batch and on the wavefront pool; the faster is
    the baseline) and guided (wavefront pool; a proposal network per view,
    trained online on its view's samples, and one network shared by all views
    reported beside it) run progressive passes
    of all views; a view's time-to-relMSE is the batch time at which its image
    first reaches the target against a ref_spp plain reference. The CPU leg
    renders view 0 at its native resolution with the oracle on host cores in
    1-spp passes for a fixed wall budget (samples/s, relMSE reached)."""
    import torch
    from rs_illume_amd import scenes as S
    from rs_illume_amd.guide import Guide
    from rs_illume_amd.render import GUIDE_PRODUCT, GuideCfg, Scene, Wavefront, render_batch
    t0 = time.time()
    desc = S.terrain_scene(0)
    t1 = time.time()
    sc = Scene(desc, builder=builder)
    setup_s = time.time() - t1
    cams = S.terrain_views(0)
    depth = 10
    sizes = [int(c["width"][0]) * int(c["height"][0]) for c in cams]
    n_px = sum(sizes)
    wf = Wavefront(max_paths=1 << 22, max_depth=depth, block=256, graph_iters=4)

    def jobs(spp, off, seed):
        return [(sc, c, spp, depth, seed + v, off) for v, c in enumerate(cams)]

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref = torch.zeros(3 * n_px, dtype=torch.float32, device="cuda")
    for k in range(ref_spp // 32):
        render_batch(jobs(32, 32 * k, 5000), film=ref)
    torch.cuda.synchronize()
    ref_s = time.perf_counter() - t0
    ref /= ref_spp
    offs = np.concatenate([[0], np.cumsum(sizes)]) * 3
    log("combined: %.1fM tris, %d spheres, %d views, %.2f Mpx; reference %d spp in %.1f s"
        % (len(desc.idx) / 1e6, len(desc.spheres), len(cams), n_px / 1e6, ref_spp, ref_s))

    def run(guide=None, guides=None, plain_wavefront=False):
        # guide: one network for every view; guides: one network per view;
        # neither: plain, on the megakernel batch or (plain_wavefront) the pool
        guided_run = guide is not None or guides is not None
        on_pool = guided_run or plain_wavefront
        film = torch.zeros_like(ref)
        gc = GuideCfg(guide_frac, train_steps, 65536, 0.0, GUIDE_PRODUCT) if guided_run else None
        warm = guided_run and warmup_steps > 0
        reached = [None] * len(cams)
        elapsed, spp = 0.0, 0
        if not on_pool:
            render_batch(jobs(1, 1 << 23, 9000))
        elif not guided_run:
            wf.render_batch(jobs(1, 1 << 23, 9000))
        else:
            wf.render_batch(jobs(1, 1 << 23, 9000), guide=Guide(seed=99), gcfg=gc)

        def steps():
            # Adam steps of the network (the mean over the per-view networks)
            return guide.stats()[0] if guide is not None else float(np.mean([g.stats()[0] for g in guides]))

        while spp < max_spp and any(r is None for r in reached):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            if warm:
                gc.guide_frac = 0.0
            if not on_pool:
                render_batch(jobs(spp_pass, spp, 100), film=film)
            elif not guided_run:
                wf.render_batch(jobs(spp_pass, spp, 100), film=film)
            elif guide is not None:
                wf.render_batch(jobs(spp_pass, spp, 200), film=film, guide=guide, gcfg=gc)
            else:
                wf.render_batch(jobs(spp_pass, spp, 200), film=film, gcfg=gc, guides=guides)
            if warm and steps() >= warmup_steps:
                warm = False
            if guided_run and not warm:
                gc.guide_frac = guide_frac
            torch.cuda.synchronize()
            elapsed += time.perf_counter() - t1
            spp += spp_pass
            for v in range(len(cams)):
                if reached[v] is None:
                    a, b = int(offs[v]), int(offs[v + 1])
                    from rs_illume_amd.converge import rel_mse_dev
                    if rel_mse_dev(film[a:b] / spp, ref[a:b]) <= target:
                        reached[v] = (round(elapsed, 4), spp)
        return {"seconds": round(elapsed, 4), "spp": spp, "samples_per_s": round(n_px * spp / elapsed, 1),
                "views_reached": sum(r is not None for r in reached),
                "time_to_target_per_view": [r[0] if r else None for r in reached],
                "spp_to_target_per_view": [r[1] if r else None for r in reached]}

    # plain on both renderers; the faster is the baseline
    plain_mk = run()
    plain_wf = run(plain_wavefront=True)
    plain_on = "megakernel_batch" if plain_mk["seconds"] <= plain_wf["seconds"] else "wavefront"
    plain = plain_mk if plain_on == "megakernel_batch" else plain_wf
    # a proposal network per view, as configured (each trained on its view's
    # samples), and one network shared by all views beside it
    guided = run(guides=[Guide(seed=70 + v) for v in range(len(cams))])
    guided_one = run(guide=Guide(seed=7))
    wf.close()
    sc.close()
    cpu = None
    if not no_cpu:
        log("combined: CPU leg (oracle, view 0, %.0f s budget)" % cpu_budget_s)
        t1 = time.time()
        osc = _oracle().OracleRenderScene(desc.arrays())
        build_s = time.time() - t1
        w0, h0 = int(cams[0]["width"][0]), int(cams[0]["height"][0])
        ref0 = ref[int(offs[0]):int(offs[1])].reshape(h0, w0, 3).cpu().numpy()
        cpu = cpu_progressive(osc, cams[0], ref0, depth, 11, 1, cpu_budget_s, target)
        cpu["sample"] = ("view 0 (%dx%d) in 1-spp passes for a %.0f s wall budget (oracle BVH build %.1f s excluded)"
                         % (w0, h0, cpu_budget_s, build_s))
        del osc
    def median_ratio(g_run):
        both = [(p, g) for p, g in zip(plain["time_to_target_per_view"], g_run["time_to_target_per_view"])
                if p is not None and g is not None]
        return round(float(np.median([g / p for p, g in both])), 3) if both else None

    return {"workload": "fBm terrain 1024^2 (2.09M tris, 6 octaves, lacunarity 2, gain 0.5) + 10k spheres "
                        "(radius lognormal, median 0.02), Lambertian/microfacet, analytic sky; 16 views "
                        "256^2..1920x1080 (%.2f Mpx), max depth 10, %d spp passes of all views per render_batch"
                        % (n_px / 1e6, spp_pass),
            "bvh_builder": builder, "scene_setup_s": round(setup_s, 2), "reference_spp": ref_spp,
            "plain_renderer": plain_on, "guided_renderer": "wavefront",
            "plain_megakernel_batch": plain_mk, "plain_wavefront": plain_wf,
            "reference_s": round(ref_s, 2),
            "target_rel_mse": target, "train_steps": train_steps, "warmup_steps": warmup_steps,
            "guide_frac": guide_frac, "product_with_bsdf_lobe": True,
            "plain": plain, "guided": guided, "guided_networks": "one per view (16), trained on its view's samples",
            "guided_one_network": guided_one,
            "median_time_ratio_guided_over_plain": median_ratio(guided),
            "median_time_ratio_one_network_over_plain": median_ratio(guided_one), "cpu_baseline": cpu}


def bench_mlp():
    """MFMA kernels of the proposal: inference over 2^20 samples and one Adam
    step over 65,536 records, timed with HIP events on the launch stream."""
    import torch
    from rs_illume_amd.guide import Guide
    g = Guide(seed=0)
    n = 1 << 20
    gen = torch.Generator(device="cuda").manual_seed(0)
    x = torch.rand((n, 6), device="cuda", generator=gen)
    x[:, 3:] = torch.nn.functional.normalize(x[:, 3:] * 2 - 1, dim=1)
    out = torch.empty((n, 40), device="cuda")
    recs = torch.rand((65536, 10), device="cuda", generator=gen)
    recs[:, 3:6] = torch.nn.functional.normalize(recs[:, 3:6] * 2 - 1, dim=1)
    recs[:, 6:9] = torch.nn.functional.normalize(recs[:, 6:9] * 2 - 1, dim=1)
    for _ in range(3):
        g.infer(x, out)
        g.train_step(recs)
    st = torch.cuda.current_stream()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    reps = 20
    e0.record(st)
    for _ in range(reps):
        g.infer(x, out)
    e1.record(st)
    for _ in range(reps):
        g.train_step(recs)
    e2.record(st)
    torch.cuda.synchronize()
    inf_ms = e0.elapsed_time(e1) / reps
    trn_ms = e1.elapsed_time(e2) / reps
    inf_tf = MLP_INFER_FLOP * n / (inf_ms * 1e-3) / 1e12
    trn_tf = MLP_TRAIN_FLOP * 65536 / (trn_ms * 1e-3) / 1e12
    return {"infer": {"samples": n, "ms": round(inf_ms, 4), "tflops": round(inf_tf, 2),
                      "frac_bf16_peak": round(inf_tf / BF16_PEAK_TFLOPS, 5)},
            "train_step": {"samples": 65536, "ms": round(trn_ms, 4), "tflops": round(trn_tf, 2),
                           "frac_bf16_peak": round(trn_tf / BF16_PEAK_TFLOPS, 5)},
            "flop_per_sample": {"infer": MLP_INFER_FLOP, "train": MLP_TRAIN_FLOP}, "peak_tflops": BF16_PEAK_TFLOPS}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline legs")
    ap.add_argument("--no-render", action="store_true", help="skip the path-tracing measurements")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter passes (roofline from algorithmic bytes only)")
    ap.add_argument("--only", choices=("cornell", "mlp", "guided", "guided_indirect", "batched", "combined"),
                    help="run one path-tracing measurement alone and print its JSON")
    ap.add_argument("--rays", type=int, default=N_RAYS)
    ap.add_argument("--width", type=int, default=4, choices=(2, 4), help="BVH node width")
    ap.add_argument("--builder", default="SAH_GPU", choices=("SAH", "SAH_GPU", "LBVH"),
                    help="BVH builder: binned SAH on the device (default) or the host (the same tree), or the "
                         "device LBVH")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (rays would shard across GPUs with no "
              "exchange, each GPU holding the whole tree). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # hardware counters of the headline kernel, from child rocprofv3 passes
    # over the same workload, before this process touches the GPU
    pmc, pmc_error = None, None
    if not args.no_counters and not args.only and args.rays == N_RAYS:
        from rs_illume_amd import counters
        try:
            t0 = time.time()
            pmc = counters.collect(args.width, log=log, builder=args.builder)
            log("counter passes %.1f s" % (time.time() - t0))
        except Exception as e:  # reported in the JSON line; the timing still runs
            pmc_error = str(e)[-300:]
            log("counter passes failed:", pmc_error)
    pmc_batched = None
    if not args.no_counters and not args.only and not args.no_render:
        from rs_illume_amd import counters
        try:
            t0 = time.time()
            pmc_batched = counters.collect_batched(log=log, builder=args.builder)
            log("batched counter passes %.1f s" % (time.time() - t0))
        except Exception as e:
            log("batched counter passes failed:", str(e)[-300:])

    import torch
    from rs_illume_amd import workloads as W
    from rs_illume_amd.accel import BVHAccel, rays_to_device

    if args.only:
        fn = {"cornell": lambda: bench_cornell(args.no_cpu), "mlp": bench_mlp,
              "guided": lambda: bench_guided(no_cpu=args.no_cpu), "guided_indirect": bench_guided_indirect,
              "batched": lambda: bench_batched(args.no_cpu),
              "combined": lambda: bench_combined(no_cpu=args.no_cpu)}[args.only]
        print(json.dumps({args.only: fn()}), flush=True)
        return

    log("generating workload")
    v, idx = W.triangle_soup(N_TRIS, seed=1)
    rays = W.sphere_to_cube_rays(args.rays, seed=2)
    log("building BVH")
    bvh = BVHAccel(v, idx, None, 4, split_method=args.builder, width=args.width)
    info = bvh.info()
    log("bvh", info)
    rd = rays_to_device(rays)
    n = rd.shape[0]
    hits = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()

    # algorithmic work per ray from the instrumented traversal (outside the timed region)
    counts = bvh.intersect_counts(rd).double().mean(0).cpu().numpy()
    box_tests, prim_tests = float(counts[0]), float(counts[1])

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

    hit_frac = float(((hits[:, 0].to(torch.int64) & 0xFFFFFFFF) != 0xFFFFFFFF).float().mean())

    # any hit (intersect_p) on the same tree and rays, outside the headline's timed region
    occluded = torch.empty(n, dtype=torch.uint8, device="cuda")
    for _ in range(max(args.warmup, 1)):
        bvh.intersect_p(rd, occluded)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        bvh.intersect_p(rd, occluded)
    ev1.record(stream)
    torch.cuda.synchronize()
    any_ms = ev0.elapsed_time(ev1) / args.steps
    any_hit = {"ms_per_launch": round(any_ms, 4), "mrays_per_s": round(n / (any_ms * 1e-3) / 1e6, 2),
               "occluded_fraction": round(float(occluded.float().mean()), 4)}
    cpu, oracle_counts = None, None
    if not args.no_cpu:
        log("cpu baseline")
        cpu, oracle_counts = cpu_baseline(v, idx, rays)
    # algorithmic bytes per ray. Own layout: what this kernel's data
    # structures hold for the work it does (ray 32 B in, hit 16 B out, one
    # 64-byte node per visit, one 48-byte primitive record per test; counts
    # from the counting traversal). pbrt layout: the reference's 32-byte
    # binary nodes and 36-byte triangles at the oracle's counts (or, without
    # the CPU leg, the GPU's counts).
    visits = box_tests / float(args.width)
    own_bytes = 32 + 16 + 64 * visits + 48 * prim_tests
    if oracle_counts:
        n_node, n_tri = oracle_counts
        src = "oracle pbrt BVH (node tests, triangle tests) on the cpu_baseline sample"
    else:
        n_node, n_tri = box_tests / 2.0, prim_tests
        src = "GPU counting kernel (box tests / 2, primitive tests)"
    pbrt_bytes = 32 + 8 + 32 * n_node + 36 * n_tri
    sec = ms_per_step * 1e-3
    own_gbs = own_bytes * n / sec / 1e9
    pbrt_gbs = pbrt_bytes * n / sec / 1e9
    algorithmic = {
        "own_layout": {"bytes_per_ray": round(own_bytes, 1), "gbs": round(own_gbs, 1),
                       "frac": round(own_gbs / HBM_PEAK_GBS, 4), "node_visits_per_ray": round(visits, 2),
                       "prim_tests_per_ray": round(prim_tests, 2),
                       "formula": "32 (ray) + 16 (hit) + 64 x node visits + 48 x primitive tests"},
        "pbrt_layout": {"bytes_per_ray": round(pbrt_bytes, 1), "gbs": round(pbrt_gbs, 1),
                        "frac": round(pbrt_gbs / HBM_PEAK_GBS, 4), "node_tests_per_ray": round(n_node, 2),
                        "tri_tests_per_ray": round(n_tri, 2), "source": src,
                        "formula": "32 + 8 + 32 x node tests + 36 x triangle tests (the reference's layout; "
                                   "the caches serve much of it, so this is not a bound)"},
        "gpu_box_tests_per_ray": round(box_tests, 2), "gpu_prim_tests_per_ray": round(prim_tests, 2)}
    if pmc and n == N_RAYS:
        traffic = pmc["traffic_bytes"]
        achieved = traffic / sec / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": round(traffic),
                    "basis": "measured: FETCH_SIZE (calibrated) + WRITE_SIZE per launch of the traversal kernel from "
                             "rocprofv3 --pmc passes in this run, over the HIP-event launch time",
                    "kernel": pmc["kernel"], "fetch_bytes": round(pmc["fetch_bytes"]),
                    "write_bytes": round(pmc["write_bytes"]),
                    "fetch_calibration": {k: (round(v, 4) if isinstance(v, float) else v)
                                          for k, v in pmc["fetch_calibration"].items()},
                    "l2_hit_rate": round(pmc["l2_hit_rate"], 4),
                    "memory_share_of_l2_reads": round(pmc["memory_share_of_l2_reads"], 4),
                    "algorithmic": algorithmic}
    else:
        roofline = {"bound": "hbm", "achieved": round(own_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(own_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                    "basis": "algorithmic bytes of the kernel's own layout (no counter pass in this run%s)"
                             % ((": " + pmc_error) if pmc_error else ""),
                    "algorithmic": algorithmic}
    del rd, hits
    bvh.close()
    torch.cuda.empty_cache()

    extra = {}
    if not args.no_render:
        log("cornell")
        extra["cornell"] = bench_cornell(args.no_cpu)
        log("mlp")
        extra["mlp"] = bench_mlp()
        log("guided vs plain")
        extra["guided"] = bench_guided(no_cpu=args.no_cpu)
        log("guided vs plain, indirectly lit room")
        extra["guided_indirect"] = bench_guided_indirect()
        log("batched")
        extra["batched"] = bench_batched(args.no_cpu, pmc=pmc_batched, builder=args.builder)
        log("combined")
        extra["combined"] = bench_combined(no_cpu=args.no_cpu, builder=args.builder)

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
                   "bvh_builder": args.builder,
                   "parallelism": "single-gpu"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "hit_fraction": round(hit_frac, 4),
        "any_hit": any_hit,
        "bvh": {"inner_nodes": info["n_inner_nodes"], "max_depth": info["max_depth"],
                "device_mb": round(info["device_bytes"] / 2 ** 20, 1), "build_s": round(info["build_seconds"], 2)},
        "wall_s_timed": round(wall, 4),
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
