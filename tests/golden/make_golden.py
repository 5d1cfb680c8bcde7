"""Writes the intersection golden fixtures with the CPU oracle.

    python tests/golden/make_golden.py

Each fixture (.npz, plain arrays, no pickles) holds the inputs — vertices,
indices, spheres, rays — and the oracle's outputs: closest-hit prim ids, hit
distance / barycentric bits and any-hit booleans. The oracle's BVH answers are
cross-checked against its O(N) brute-force loop before anything is written.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [REPO, os.path.join(REPO, "oracle"), os.path.join(REPO, "tests")]

import pyoracle  # noqa: E402
from cases import edge_case_rays, small_scene  # noqa: E402
from rs_illume_amd import workloads as W  # noqa: E402


def make(name, n_tris, n_spheres, n_rays, seed):
    v, idx, sph = small_scene(n_tris, n_spheres, seed)
    rays = np.concatenate([W.sphere_to_cube_rays(n_rays, seed=seed + 100), edge_case_rays(v, idx, sph, seed + 200)])
    sc = pyoracle.OracleScene(v, idx, sph, 4)
    hits = sc.intersect(rays)
    occ = sc.intersect_p(rays)
    brute = sc.intersect(rays, brute=True)
    occ_b = sc.intersect_p(rays, brute=True)
    assert np.array_equal(occ, occ_b), "oracle BVH any-hit disagrees with brute force"
    mism = np.nonzero(hits["prim"] != brute["prim"])[0]
    for i in mism:  # only exact-distance ties may differ
        ok1, t1 = sc.prim_intersect(hits["prim"][i], rays[i])
        ok2, t2 = sc.prim_intersect(brute["prim"][i], rays[i])
        assert ok1 and ok2 and t1 == t2, "oracle BVH closest hit disagrees with brute force"
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, verts=v, idx=idx, spheres=sph, rays=rays.view(np.float32).reshape(-1, 8),
                        prim=hits["prim"], t=hits["t"], b1=hits["b1"], b2=hits["b2"], occluded=occ)
    print("wrote", path, "rays", len(rays), "hits", int((hits["prim"] != pyoracle.MISS).sum()),
          "occluded", int(occ.sum()), "bvh/brute id ties", len(mism))


if __name__ == "__main__":
    make("isect_tris_spheres", 1024, 16, 4096, 21)
    make("isect_tris_only", 3000, 0, 4096, 31)
