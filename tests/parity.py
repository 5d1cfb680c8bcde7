"""Parity contract for closest-hit results (shared by the GPU tests).

* prim ids must equal the oracle's for every ray; a mismatch is counted as a
  tie (not a failure) only if the oracle confirms that BOTH primitives are
  genuine hits of that ray and their hit distances are within 2 ulps;
* for equal ids the hit distance must be within 2 ulps (target 0) and the
  barycentrics equal;
* any-hit booleans must be equal.
"""
import numpy as np

from conftest import ulp_diff

MISS = 0xFFFFFFFF
T_ULP_BOUND = 2


def compare_closest(gpu_hits, ref_hits, oscene, rays):
    g = gpu_hits["prim"].astype(np.int64)
    r = ref_hits["prim"].astype(np.int64)
    same = g == r
    both_hit = same & (g != MISS)
    t_ulps = ulp_diff(gpu_hits["t"][both_hit], ref_hits["t"][both_hit])
    b_equal = (np.array_equal(gpu_hits["b1"][both_hit].view(np.uint32), ref_hits["b1"][both_hit].view(np.uint32))
               and np.array_equal(gpu_hits["b2"][both_hit].view(np.uint32),
                                  ref_hits["b2"][both_hit].view(np.uint32)))
    ties, failures = 0, []
    for i in np.nonzero(~same)[0]:
        if g[i] == MISS or r[i] == MISS:
            failures.append(int(i))
            continue
        okg, tg = oscene.prim_intersect(g[i], rays[i])
        okr, tr = oscene.prim_intersect(r[i], rays[i])
        if okg and okr and ulp_diff(np.float32(tg), np.float32(tr)) <= T_ULP_BOUND:
            ties += 1
        else:
            failures.append(int(i))
    return {
        "n": len(g),
        "hits": int((r != MISS).sum()),
        "mismatch": int((~same).sum()),
        "ties": ties,
        "failures": failures,
        "max_t_ulp": int(t_ulps.max()) if t_ulps.size else 0,
        "t_exact_frac": float((t_ulps == 0).mean()) if t_ulps.size else 1.0,
        "bary_equal": bool(b_equal),
    }
