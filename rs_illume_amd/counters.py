"""Hardware-counter measurement of the closest-hit traversal, and kernel-trace
summaries, from rocprofv3.

``collect()`` runs the traversal workload (``python -m rs_illume_amd.workloads``:
counter calibration kernels, then closest-hit launches of the 2^20-triangle /
2^24-ray workload) under ``rocprofv3 --pmc`` as child processes, one counter
pass each (a pass holds at most 4 TCC counters: FETCH_SIZE uses 3, WRITE_SIZE
2), and returns per-launch figures for the traversal kernel:

* HBM-side traffic: FETCH_SIZE (L2 read misses to the memory fabric) and
  WRITE_SIZE, in bytes. FETCH_SIZE is calibrated against
  ``illume_probe_random_reads``, which reads a known number of lane-private
  random 64-byte lines (the traversal's node-fetch pattern); the device copy's
  factor for streaming reads is reported beside it.
* L2 behaviour: TCC_HIT / (TCC_HIT + TCC_MISS), and the share of L2 read
  requests that leave the L2 for memory (TCC_EA0_RDREQ_DRAM / TCC_READ). The
  Infinity Cache (MALL) sits on the memory side of that boundary and has no
  counter here, so "memory" includes its hits.

``kernel_stats(csv_dir, pattern, skip)`` summarises a ``--kernel-trace`` run:
per-dispatch durations of the kernels whose name matches, with the first
``skip`` dispatches (warm-up) left out.
"""
import csv
import glob
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)

# kernel-name patterns (demangled names in the rocprofv3 traces)
CLOSEST4 = r"traverse_refill_kernel<false, ?4>"
CLOSEST2 = r"traverse_kernel<false, ?false, ?2"
PROBE = r"probe_kernel"
COPY = r"copyBuffer"
PROBE_BYTES = (1 << 20) * 64 * 64          # illume_probe_random_reads in the workload
COPY_BYTES = 1 << 30                       # the workload's 1 GiB device copy reads this much

PASSES = (
    ("FETCH_SIZE",),
    ("WRITE_SIZE",),
    ("TCC_HIT_sum", "TCC_MISS_sum", "TCC_READ_sum", "TCC_EA0_RDREQ_DRAM_sum"),
)


def _rows(out_dir, suffix):
    rows = []
    for path in glob.glob(os.path.join(out_dir, "**", "*" + suffix), recursive=True):
        with open(path, newline="") as f:
            rows.extend(csv.DictReader(f))
    return rows


def _col(row, *names):
    low = {k.lower(): k for k in row}
    for n in names:
        if n.lower() in low:
            return row[low[n.lower()]]
    raise KeyError(names)


def counter_values(out_dir):
    """{(dispatch id, kernel name): {counter: value}} of a --pmc run (values
    of several rows of one counter in one dispatch are summed)."""
    out = {}
    for r in _rows(out_dir, "counter_collection.csv"):
        key = (int(_col(r, "Dispatch_Id")), _col(r, "Kernel_Name"))
        name = _col(r, "Counter_Name")
        d = out.setdefault(key, {})
        d[name] = d.get(name, 0.0) + float(_col(r, "Counter_Value"))
    return out


def _mean_for(values, pattern, counter):
    v = [d[counter] for (_, k), d in sorted(values.items()) if re.search(pattern, k) and counter in d]
    return (sum(v) / len(v), len(v)) if v else (None, 0)


BATCH = r"render_batch_kernel"


def run_pass(counters, width=4, launches=2, timeout=150, keep_dir=None, batched=False, builder="SAH_GPU"):
    """One rocprofv3 --pmc pass over the traversal workload (or, batched, the
    64-scene batched render); returns the per-dispatch counter values."""
    out_dir = keep_dir or tempfile.mkdtemp(prefix="illume_pmc_", dir="/tmp")
    env = dict(os.environ)
    env["TMPDIR"] = "/tmp"
    env["PYTHONPATH"] = REPO + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
    cmd = ["timeout", "-s", "KILL", str(timeout), "rocprofv3", "--pmc"] + list(counters) + \
          ["--output-format", "csv", "-d", out_dir, "-o", "run", "--", sys.executable, "-m",
           "rs_illume_amd.workloads", "--width", str(width), "--launches", str(launches), "--builder", builder] + \
          (["--batched"] if batched else [])
    r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True)
    vals = counter_values(out_dir)
    if not vals:
        raise RuntimeError("rocprofv3 pass %s gave no counters (rc %d): %s" % (counters, r.returncode,
                                                                              (r.stderr or r.stdout)[-400:]))
    return vals


def collect(width=4, launches=2, log=None, builder="SAH_GPU"):
    """Per-launch counter figures of the closest-hit traversal kernel on the
    2^20-triangle / 2^24-ray workload (see the module docstring)."""
    pattern = CLOSEST4 if width == 4 else CLOSEST2
    got = {}
    for counters in PASSES:
        if log:
            log("rocprofv3 --pmc %s" % " ".join(counters))
        vals = run_pass(counters, width=width, launches=launches, builder=builder)
        for c in counters:
            got[c], n = _mean_for(vals, pattern, c)
            if got[c] is None:
                raise RuntimeError("no %s value for the traversal kernel" % c)
            if c == "FETCH_SIZE":
                got["probe_fetch_kib"], _ = _mean_for(vals, PROBE, c)
                got["copy_fetch_kib"], _ = _mean_for(vals, COPY, c)
                got["dispatches"] = n
        names = sorted({k for (_, k) in vals if re.search(pattern, k)})
        got["kernel"] = names[0] if names else None
    probe_factor = (PROBE_BYTES / 1024.0) / got["probe_fetch_kib"]
    copy_factor = (COPY_BYTES / 1024.0) / got["copy_fetch_kib"] if got.get("copy_fetch_kib") else None
    fetch = got["FETCH_SIZE"] * 1024.0 * probe_factor
    write = got["WRITE_SIZE"] * 1024.0
    hit, miss = got["TCC_HIT_sum"], got["TCC_MISS_sum"]
    return {
        "kernel": got["kernel"],
        "dispatches_per_pass": got["dispatches"],
        "fetch_bytes": fetch,
        "write_bytes": write,
        "traffic_bytes": fetch + write,
        "fetch_size_raw_kib": got["FETCH_SIZE"],
        "write_size_kib": got["WRITE_SIZE"],
        "fetch_calibration": {"factor": probe_factor, "kernel": "illume_probe_random_reads (%d random lane-private "
                              "64-byte lines)" % (PROBE_BYTES // 64), "copy_factor": copy_factor},
        "l2_hit_rate": hit / (hit + miss) if hit + miss else None,
        "l2_requests": hit + miss,
        "l2_reads": got["TCC_READ_sum"],
        "l2_read_misses_to_memory": got["TCC_EA0_RDREQ_DRAM_sum"],
        "memory_share_of_l2_reads": got["TCC_EA0_RDREQ_DRAM_sum"] / got["TCC_READ_sum"] if got["TCC_READ_sum"]
        else None,
    }


def collect_batched(log=None, builder="SAH_GPU"):
    """HBM-side traffic and L2 hit rate of one batched render launch (the
    64 procedural scenes): two passes, FETCH_SIZE + TCC_HIT_sum, then
    WRITE_SIZE + TCC_MISS_sum. FETCH_SIZE is reported raw (no calibration:
    the batched render's reads mix node lines, primitive records and
    per-pixel film atomics)."""
    got = {}
    for counters in (("FETCH_SIZE", "TCC_HIT_sum"), ("WRITE_SIZE", "TCC_MISS_sum")):
        if log:
            log("rocprofv3 --pmc %s (batched render)" % " ".join(counters))
        vals = run_pass(counters, launches=1, timeout=240, batched=True, builder=builder)
        for c in counters:
            got[c], _ = _mean_for(vals, BATCH, c)
            if got[c] is None:
                raise RuntimeError("no %s value for the batched render kernel" % c)
        names = sorted({k for (_, k) in vals if re.search(BATCH, k)})
        got["kernel"] = names[0] if names else None
    hit, miss = got["TCC_HIT_sum"], got["TCC_MISS_sum"]
    fetch, write = got["FETCH_SIZE"] * 1024.0, got["WRITE_SIZE"] * 1024.0
    return {"kernel": got["kernel"], "fetch_bytes": fetch, "write_bytes": write, "traffic_bytes": fetch + write,
            "l2_hit_rate": hit / (hit + miss) if hit + miss else None}


def kernel_stats(out_dir, pattern, skip=0):
    """Durations (ms) of the dispatches of kernels matching `pattern` in a
    --kernel-trace run, in dispatch order, minus the first `skip`:
    {"kernel", "calls", "average_ms", "min_ms", "max_ms", "skipped"}."""
    rows = []
    for r in _rows(out_dir, "kernel_trace.csv"):
        name = _col(r, "Kernel_Name")
        if re.search(pattern, name):
            rows.append((int(_col(r, "Dispatch_Id")), name,
                         (int(_col(r, "End_Timestamp")) - int(_col(r, "Start_Timestamp"))) / 1e6))
    rows.sort()
    kept = rows[skip:]
    if not kept:
        return None
    d = [x[2] for x in kept]
    return {"kernel": kept[0][1], "calls": len(d), "average_ms": sum(d) / len(d), "min_ms": min(d), "max_ms": max(d),
            "skipped": skip}


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="traversal counters (rocprofv3 --pmc passes) or a kernel-trace summary")
    ap.add_argument("--width", type=int, default=4, choices=(2, 4))
    ap.add_argument("--trace-dir", help="summarise this --kernel-trace output directory instead")
    ap.add_argument("--pattern", default=CLOSEST4)
    ap.add_argument("--skip", type=int, default=0)
    a = ap.parse_args()
    if a.trace_dir:
        print(json.dumps(kernel_stats(a.trace_dir, a.pattern, a.skip)))
    else:
        print(json.dumps(collect(a.width, log=lambda m: print(m, file=sys.stderr, flush=True))))
