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


def _profile_launches(width=2, launches=2, copy_mib=1024, builder="SAH_GPU"):
    """The profiling workload for hardware counters: the calibrations of
    FETCH_SIZE / WRITE_SIZE against known bytes (a plain device copy of
    copy_mib MiB; a torch gather of 2^22 random 64-byte rows; and
    illume_probe_random_reads: 2^20 lanes x 64 random 64-byte lines read lane
    by lane with float4 loads, the access pattern of node fetches), then
    `launches` closest-hit launches of the 2^20-triangle / 2^24-ray workload."""
    import ctypes
    import torch
    from . import _lib
    from .accel import BVHAccel, rays_to_device
    n = copy_mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    rows = src.view(-1, 16)
    gen = torch.Generator(device="cuda").manual_seed(0)
    pick = torch.randint(0, rows.shape[0], (1 << 22,), device="cuda", generator=gen)
    gathered = rows.index_select(0, pick)
    probe_out = torch.empty(1 << 20, dtype=torch.float32, device="cuda")
    L = _lib.lib()
    L.illume_probe_random_reads.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.c_void_p, ctypes.c_void_p]
    _lib.check(L.illume_probe_random_reads(ctypes.c_void_p(src.data_ptr()), src.numel() * 4, 1 << 20, 64,
                                           ctypes.c_void_p(probe_out.data_ptr()), None), "illume_probe_random_reads")
    torch.cuda.synchronize()
    del gathered, probe_out
    v, idx = triangle_soup(1 << 20, seed=1)
    rays = sphere_to_cube_rays(1 << 24, seed=2)
    bvh = BVHAccel(v, idx, None, 4, split_method=builder, width=width)
    rd = rays_to_device(rays)
    hits = torch.empty((rd.shape[0], 4), dtype=torch.int32, device="cuda")
    for _ in range(launches):
        bvh.intersect(rd, hits)
    torch.cuda.synchronize()
    print("copy %d MiB, %d closest-hit launches (width %d) done" % (copy_mib, launches, width), flush=True)


def _probe_ceiling(sizes_mib=(16, 100, 1024, 8192), lanes=1 << 22, steps=64):
    """Random-read ceiling of the device for the node-fetch access pattern:
    illume_probe_random_reads over buffers of each size, timed with HIP events."""
    import ctypes
    import torch
    from . import _lib
    L = _lib.lib()
    L.illume_probe_random_reads.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.c_void_p, ctypes.c_void_p]
    out = torch.empty(lanes, dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream()
    for mib in sizes_mib:
        buf = torch.zeros(mib << 18, dtype=torch.float32, device="cuda")
        args = (ctypes.c_void_p(buf.data_ptr()), buf.numel() * 4, lanes, steps, ctypes.c_void_p(out.data_ptr()),
                ctypes.c_void_p(st.cuda_stream))
        _lib.check(L.illume_probe_random_reads(*args), "illume_probe_random_reads")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(3):
            _lib.check(L.illume_probe_random_reads(*args), "illume_probe_random_reads")
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 3
        print("buffer %5d MiB: %.3f ms, %.2f TB/s of random 64-byte lines" % (mib, ms, lanes * steps * 64 / ms / 1e9),
              flush=True)
        del buf


def _profile_batched(launches=1, builder="SAH_GPU"):
    """The profiling workload of the batched render: the 64 procedural scenes
    (seeds 1-64, SAH trees from `builder`) in `launches` calls of
    illume_render_batch at their own resolution and spp, max depth 8."""
    import torch
    from . import scenes as S
    from .render import Scene, render_batch
    descs = [S.procedural_scene(seed) for seed in range(1, 65)]
    scenes = [Scene(d, builder=builder) for d, _, _ in descs]
    jobs = [(sc, cam, spp, 8, seed) for seed, (sc, (_, cam, spp)) in enumerate(zip(scenes, descs), 1)]
    film = None
    for _ in range(launches):
        film, _ = render_batch(jobs, film=film)
    torch.cuda.synchronize()
    print("%d batched render launches done" % launches, flush=True)


def _profile_build(builder="SAH_GPU", width=4, repeats=3):
    """The profiling workload of the device BVH builders: `repeats` builds of
    the 2^20-triangle soup's tree."""
    from .accel import BVHAccel
    v, idx = triangle_soup(1 << 20, seed=1)
    for _ in range(repeats):
        b = BVHAccel(v, idx, None, 4, split_method=builder, width=width)
        info = b.info()
        b.close()
        print("%s build (width %d): %.4f s, %d nodes, depth %d" % (builder, width, info["build_seconds"],
                                                                  info["n_inner_nodes"], info["max_depth"]),
              flush=True)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="closest-hit profiling workload")
    ap.add_argument("--width", type=int, default=2, choices=(2, 4))
    ap.add_argument("--launches", type=int, default=2)
    ap.add_argument("--probe", action="store_true", help="time the random-read probe over several buffer sizes")
    ap.add_argument("--batched", action="store_true", help="the 64-scene batched render instead")
    ap.add_argument("--builder", default="SAH_GPU", choices=("SAH", "SAH_GPU", "LBVH"))
    ap.add_argument("--build", action="store_true", help="BVH builds of the 2^20-triangle soup instead")
    a = ap.parse_args()
    if a.probe:
        _probe_ceiling()
    elif a.build:
        _profile_build(a.builder, a.width)
    elif a.batched:
        _profile_batched(a.launches, a.builder)
    else:
        _profile_launches(a.width, a.launches, builder=a.builder)
