"""Synthetic scene and ray generators for the intersection workloads.

The 1M-triangle incoherent-ray workload: triangle centroids uniform in the unit
cube, each vertex offset from its centroid by an isotropic Gaussian whose sigma
is lognormal (median 0.005, log-sigma 0.5); rays start uniformly on a sphere of
radius 2 around the cube centre and aim at uniform random points inside the
cube. Values are drawn in f64 from numpy's PCG64 stream for the given seed and
rounded to f32 once.
"""
import numpy as np

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("t_max", "<f4"), ("time", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<u4"), ("t", "<f4"), ("b1", "<f4"), ("b2", "<f4")])
MISS = 0xFFFFFFFF


def triangle_soup(n_tris, seed=1, median_sigma=0.005, log_sigma=0.5):
    """Return (verts float32 [3n,3], idx uint32 [n,3]) of an unindexed soup."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.random((n_tris, 3))
    sigma = np.exp(np.log(median_sigma) + log_sigma * rng.standard_normal(n_tris))
    off = rng.standard_normal((n_tris, 3, 3)) * sigma[:, None, None]
    v = (c[:, None, :] + off).astype(np.float32).reshape(-1, 3)
    idx = np.arange(3 * n_tris, dtype=np.uint32).reshape(-1, 3)
    return v, idx


def sphere_to_cube_rays(n_rays, seed=2, radius=2.0, chunk=1 << 22):
    """Incoherent rays from a sphere of `radius` around (0.5,0.5,0.5) towards
    uniform points of the unit cube; t_max = +inf, time = 0."""
    rays = np.zeros(n_rays, RAY_DTYPE)
    rng = np.random.Generator(np.random.PCG64(seed))
    for s in range(0, n_rays, chunk):
        m = min(chunk, n_rays - s)
        u = rng.standard_normal((m, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        o = 0.5 + radius * u
        q = rng.random((m, 3))
        d = q - o
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays["o"][s:s + m] = o.astype(np.float32)
        rays["d"][s:s + m] = d.astype(np.float32)
    rays["t_max"] = np.inf
    rays["time"] = 0.0
    return rays


def spheres_in_cube(n, seed=4, median_radius=0.02, log_sigma=0.5):
    """n spheres (cx, cy, cz, r) with centres uniform in the unit cube."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.random((n, 3))
    r = np.exp(np.log(median_radius) + log_sigma * rng.standard_normal(n))
    return np.concatenate([c, r[:, None]], axis=1).astype(np.float32)
