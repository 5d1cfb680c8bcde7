"""Autotuner for the wavefront renderer's launch shape.

Parameters tuned (see illume_wavefront_set_tuning / illume_wavefront_create):
  * max_paths  — size of the path-slot pool (queue capacity),
  * block      — threads per block of the extend and shade stages (wave shape),
  * regen_min  — regenerate camera paths only while fewer slots than this are
                 live (the compaction/refill threshold, as a fraction of the pool),
  * graph_iters — iterations per replayed hipGraph.

Each candidate renders a reduced version of the real batch (every job with a
few samples per pixel) and is timed by wall clock between device
synchronisations (a batch render reads its queue counters back every graph
replay, so host time is the job's time). The fastest few are timed again on a
larger probe (two renders each), whose fastest is kept: on the 64-scene batch
the 2-spp probe alone picked shapes up to 3% slower on the full render
(candidates within its noise).
Results are cached per (device name, workload size bucket) in a JSON file so a
workload class is tuned once.
"""
import itertools
import json
import math
import os
import time

DEFAULT_SPACE = {
    "max_paths": (1 << 21, 1 << 22, 1 << 23),
    "block": (128, 256),
    "regen_frac": (1.01, 0.5),
    "graph_iters": (4, 8),
}


def workload_key(jobs):
    """Size bucket of a batch: log2 of total paths and of total triangles."""
    import numpy as np
    from .scenes import CAMERA_DTYPE
    paths = 0
    for _, cam, spp, *_ in jobs:
        c = np.ascontiguousarray(cam, CAMERA_DTYPE).reshape(1)
        paths += int(c["width"][0]) * int(c["height"][0]) * int(spp)
    tris = sum(len(sc.arrays["idx"]) for sc, *_ in jobs)
    return "p%d_t%d_j%d" % (int(math.log2(max(paths, 1))), int(math.log2(max(tris, 1))), len(jobs))


def _time_render(wf, jobs, reps=1):
    import torch
    wf.render_batch(jobs)  # warm-up (captures the graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        wf.render_batch(jobs)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def tune(jobs, space=None, cache_path=None, max_depth=8, sample_spp=2, log=None, refine_top=3, refine_spp=8):
    """Pick the fastest launch shape for `jobs` (list of (Scene, camera, spp,
    max_depth, seed)): every candidate on a sample_spp probe, then the
    refine_top fastest on a refine_spp probe. Returns (params dict, table of
    measured candidates)."""
    import torch
    from .render import Wavefront
    space = dict(DEFAULT_SPACE, **(space or {}))
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"
    key = "%s|%s" % (dev, workload_key(jobs))
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            return cache[key]["best"], cache[key]["table"]
    probe = [(sc, cam, min(sample_spp, spp), d, seed) for sc, cam, spp, d, seed, *_ in jobs]
    table = []
    for mp, blk, rf, gi in itertools.product(space["max_paths"], space["block"], space["regen_frac"],
                                            space["graph_iters"]):
        regen = int(mp * rf) + (1 if rf > 1 else 0)
        wf = Wavefront(max_paths=mp, max_depth=max_depth, block=blk, regen_min=min(regen, mp + 1), graph_iters=gi)
        dt = _time_render(wf, probe)
        row = {"max_paths": mp, "block": blk, "regen_frac": rf, "graph_iters": gi, "seconds": dt,
               "iterations": wf.last_iterations()}
        table.append(row)
        if log:
            log(row)
        wf.close()
        del wf
    best = min(table, key=lambda r: r["seconds"])
    if refine_top > 1:
        fine = [(sc, cam, min(refine_spp, spp), d, seed) for sc, cam, spp, d, seed, *_ in jobs]
        for row in sorted(table, key=lambda r: r["seconds"])[:refine_top]:
            wf = make_wavefront(row, max_depth)
            row["refined_seconds"] = _time_render(wf, fine, reps=2)
            if log:
                log(row)
            wf.close()
            del wf
        best = min((r for r in table if "refined_seconds" in r), key=lambda r: r["refined_seconds"])
    if cache_path:
        cache[key] = {"best": best, "table": table}
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return best, table


def make_wavefront(params, max_depth=8):
    from .render import Wavefront
    mp = int(params["max_paths"])
    rf = float(params["regen_frac"])
    regen = min(int(mp * rf) + (1 if rf > 1 else 0), mp + 1)
    return Wavefront(max_paths=mp, max_depth=max_depth, block=int(params["block"]), regen_min=regen,
                     graph_iters=int(params["graph_iters"]))
