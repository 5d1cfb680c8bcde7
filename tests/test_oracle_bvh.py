"""Pins the oracle BVH (accelerators/bvh.rs) by structural invariants and by
brute-force equivalence, and the oracle as a whole against the committed
golden fixtures."""
import os

import numpy as np
import pytest

from cases import edge_case_rays, small_scene

HERE = os.path.dirname(os.path.abspath(__file__))


def _check_tree(osc, verts, idx, spheres, max_prims):
    bounds, off, npr, axis, ordered = osc.nodes()
    n_prims = len(idx) + len(spheres)
    assert sorted(ordered.tolist()) == list(range(n_prims)), "every primitive exactly once"

    def prim_box(p):
        if p < len(idx):
            v = verts[idx[p]]
            return v.min(0), v.max(0)
        c, r = spheres[p - len(idx), :3], spheres[p - len(idx), 3]
        return c - r, c + r

    seen = np.zeros(len(bounds), bool)
    depth_max = 0
    stack = [(0, 0)]
    while stack:
        i, dep = stack.pop()
        depth_max = max(depth_max, dep)
        seen[i] = True
        lo, hi = bounds[i, :3], bounds[i, 3:]
        if npr[i] > 0:
            for p in ordered[off[i]:off[i] + npr[i]]:
                plo, phi = prim_box(p)
                assert (plo >= lo - 1e-6).all() and (phi <= hi + 1e-6).all()
        else:
            for c in (i + 1, off[i]):   # depth-first: first child at i+1 (bvh.rs:504-537)
                clo, chi = bounds[c, :3], bounds[c, 3:]
                assert (clo >= lo).all() and (chi <= hi).all(), "child box inside parent"
                stack.append((c, dep + 1))
    assert seen.all(), "every node reachable"
    return depth_max


@pytest.mark.parametrize("seed,n_tris,n_spheres,max_prims", [(1, 3000, 0, 4), (2, 2000, 50, 4), (3, 1500, 20, 1),
                                                             (4, 2500, 0, 8)])
def test_bvh_invariants_and_brute_force(oracle, seed, n_tris, n_spheres, max_prims):
    from rs_illume_amd import workloads as W
    v, idx, sph = small_scene(n_tris, n_spheres, seed)
    osc = oracle.OracleScene(v, idx, sph, max_prims)
    _check_tree(osc, v, idx, sph, max_prims)
    rays = np.concatenate([W.sphere_to_cube_rays(3000, seed=seed), edge_case_rays(v, idx, sph, seed)])
    h = osc.intersect(rays)
    hb = osc.intersect(rays, brute=True)
    for i in np.nonzero(h["prim"] != hb["prim"])[0]:
        ok1, t1 = osc.prim_intersect(h["prim"][i], rays[i])
        ok2, t2 = osc.prim_intersect(hb["prim"][i], rays[i])
        assert ok1 and ok2 and t1 == t2, "only exact-distance ties may differ"
    same = h["prim"] == hb["prim"]
    assert np.array_equal(h["t"][same].view(np.uint32), hb["t"][same].view(np.uint32))
    assert np.array_equal(osc.intersect_p(rays), osc.intersect_p(rays, brute=True))
    # any-hit agrees with closest hit for infinite rays
    inf = np.isinf(rays["t_max"])
    assert np.array_equal((h["prim"][inf] != oracle.MISS), osc.intersect_p(rays)[inf].astype(bool))


def test_bvh_sah_leaf_sizes(oracle):
    v, idx, sph = small_scene(4000, 0, 9)
    osc = oracle.OracleScene(v, idx, sph, 4)
    bounds, off, npr, axis, ordered = osc.nodes()
    leaves = npr[npr > 0]
    # pbrt SAH: leaves above max_prims only when all centroids coincide
    assert leaves.max() <= 4
    info = osc.info()
    assert info["nodes"] == 2 * len(leaves) - 1


def test_bvh_stats_counts(oracle):
    from rs_illume_amd import workloads as W
    v, idx, sph = small_scene(3000, 0, 5, median_sigma=0.01)
    osc = oracle.OracleScene(v, idx, sph, 4)
    rays = W.sphere_to_cube_rays(2000, seed=3)
    _, st = osc.intersect(rays, with_stats=True)
    _, stb = osc.intersect(rays, brute=True, with_stats=True)
    assert st[0] > len(rays) and st[1] < 3000 * len(rays) / 10


@pytest.mark.parametrize("name", ["isect_tris_spheres", "isect_tris_only"])
def test_oracle_reproduces_golden_fixture(oracle, name):
    z = np.load(os.path.join(HERE, "golden", name + ".npz"))
    rays = np.ascontiguousarray(z["rays"]).view(oracle.RAY_DTYPE).reshape(-1)
    osc = oracle.OracleScene(z["verts"], z["idx"], z["spheres"], 4)
    h = osc.intersect(rays)
    assert np.array_equal(h["prim"], z["prim"])
    for k in ("t", "b1", "b2"):
        assert np.array_equal(h[k].view(np.uint32), z[k].view(np.uint32)), k
    assert np.array_equal(osc.intersect_p(rays), z["occluded"])
