"""GPU parity of closest-hit / any-hit traversal against the CPU oracle.

All calls go through the C ABI (librs_illume_amd.so via ctypes). Contract: see
tests/parity.py.
"""
import os
import sys

import numpy as np
import pytest

from cases import edge_case_rays, small_scene
from parity import MISS, compare_closest

HERE = os.path.dirname(os.path.abspath(__file__))
pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu_run(bvh, rays):
    from rs_illume_amd.accel import hits_to_numpy, rays_to_device
    rd = rays_to_device(rays)
    h = hits_to_numpy(bvh.intersect(rd))
    o = bvh.intersect_p(rd).cpu().numpy()
    return h, o


def _assert_parity(res, occ_gpu, occ_ref):
    assert not res["failures"], "closest-hit id mismatches that are not ties: %s" % res["failures"][:20]
    assert res["max_t_ulp"] <= 2, res
    assert res["bary_equal"], res
    assert np.array_equal(occ_gpu, occ_ref), "any-hit mismatch at %s" % np.nonzero(occ_gpu != occ_ref)[0][:20]


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("name", ["isect_tris_spheres", "isect_tris_only"])
def test_golden_fixture(gpu, oracle, name, width):
    from rs_illume_amd.accel import BVHAccel
    _say("case golden", name, "width", width)
    z = np.load(os.path.join(HERE, "golden", name + ".npz"))
    rays = np.ascontiguousarray(z["rays"]).view(oracle.RAY_DTYPE).reshape(-1)
    ref = np.zeros(len(rays), oracle.HIT_DTYPE)
    for k in ("prim", "t", "b1", "b2"):
        ref[k] = z[k]
    bvh = BVHAccel(z["verts"], z["idx"], z["spheres"], 4, width=width)
    h, o = _gpu_run(bvh, rays)
    osc = oracle.OracleScene(z["verts"], z["idx"], z["spheres"], 4)
    res = compare_closest(h, ref, osc, rays)
    _say(res)
    _assert_parity(res, o, z["occluded"])
    assert res["t_exact_frac"] == 1.0


@pytest.mark.parametrize("builder", ["SAH", "LBVH"])
@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("seed,n_tris,n_spheres,leaf", [(101, 16384, 64, 4), (102, 30000, 0, 1), (103, 5000, 500, 8),
                                                        (104, 20000, 16, 15)])
def test_seeded_scene(gpu, oracle, seed, n_tris, n_spheres, leaf, width, builder):
    from rs_illume_amd import workloads as W
    from rs_illume_amd.accel import BVHAccel
    _say("case seeded", seed, n_tris, n_spheres, leaf, "width", width, "builder", builder)
    v, idx, sph = small_scene(n_tris, n_spheres, seed, median_sigma=0.02)
    rays = np.concatenate([W.sphere_to_cube_rays(50000, seed=seed + 1), edge_case_rays(v, idx, sph, seed + 2)])
    bvh = BVHAccel(v, idx, sph, leaf, split_method=builder, width=width)
    h, o = _gpu_run(bvh, rays)
    osc = oracle.OracleScene(v, idx, sph, 4)
    ref = osc.intersect(rays)
    occ = osc.intersect_p(rays)
    res = compare_closest(h, ref, osc, rays)
    _say(res)
    _assert_parity(res, o, occ)


@pytest.mark.parametrize("builder", ["SAH", "LBVH"])
def test_degenerate_scenes(gpu, oracle, builder):
    from rs_illume_amd.accel import BVHAccel
    _say("case degenerate", builder)
    rng = np.random.default_rng(7)
    # many triangles sharing one centroid, zero-area triangles, duplicates
    base = rng.random((3, 3)).astype(np.float32)
    v = np.concatenate([base] * 64 + [np.repeat(rng.random((1, 3)), 3, 0).astype(np.float32)] * 8
                       + [rng.random((30, 3)).astype(np.float32)])
    idx = np.arange(len(v), dtype=np.uint32).reshape(-1, 3)
    from rs_illume_amd import workloads as W
    rays = W.sphere_to_cube_rays(20000, seed=8)
    bvh = BVHAccel(v, idx, None, 4, split_method=builder)
    h, o = _gpu_run(bvh, rays)
    osc = oracle.OracleScene(v, idx, None, 4)
    res = compare_closest(h, osc.intersect(rays), osc, rays)
    _say(res)
    _assert_parity(res, o, osc.intersect_p(rays))


@pytest.mark.parametrize("width", [2, 4])
def test_tiny_and_empty(gpu, oracle, width):
    from rs_illume_amd.accel import BVHAccel
    from rs_illume_amd import workloads as W
    _say("case tiny/empty")
    rays = W.sphere_to_cube_rays(1000, seed=9)
    # single triangle (root is a leaf)
    v = np.array([[0, 0, 0.5], [1, 0, 0.5], [0, 1, 0.5]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    bvh = BVHAccel(v, idx, None, 4, width=width)
    h, o = _gpu_run(bvh, rays)
    osc = oracle.OracleScene(v, idx, None, 4)
    res = compare_closest(h, osc.intersect(rays), osc, rays)
    _assert_parity(res, o, osc.intersect_p(rays))
    assert res["hits"] > 0
    # spheres only
    sph = np.array([[0.5, 0.5, 0.5, 0.25], [0.2, 0.8, 0.5, 0.1]], np.float32)
    bvh = BVHAccel(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32), sph, 4, width=width)
    h, o = _gpu_run(bvh, rays)
    osc = oracle.OracleScene(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32), sph, 4)
    res = compare_closest(h, osc.intersect(rays), osc, rays)
    _assert_parity(res, o, osc.intersect_p(rays))
    assert res["hits"] > 0
    # no primitives at all
    bvh = BVHAccel(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32), None, 4)
    h, o = _gpu_run(bvh, rays)
    assert (h["prim"] == MISS).all() and (o == 0).all()
    # zero rays
    import torch
    bvh.intersect(torch.zeros((0, 8), dtype=torch.float32, device="cuda"))


def test_invalid_arguments_gpu(gpu):
    from rs_illume_amd import IllumeError
    from rs_illume_amd.accel import BVHAccel
    v = np.zeros((3, 3), np.float32)
    with pytest.raises(IllumeError) as e:
        BVHAccel(v, np.array([[0, 1, 3]], np.uint32))
    assert e.value.status == 1
    with pytest.raises(IllumeError) as e:
        BVHAccel(v, np.array([[0, 1, 2]], np.uint32), None, 16)
    assert e.value.status == 4
    with pytest.raises(IllumeError):
        BVHAccel(np.array([[0, 0, np.nan], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32))


@pytest.mark.parametrize("builder", ["SAH", "SAH_GPU", "LBVH"])
@pytest.mark.parametrize("width", [2, 4])
def test_full_size_1m_tris_16m_rays(gpu, oracle, width, builder):
    """The 1M-triangle, 16M-incoherent-ray workload at full size: oracle parity
    on every ray for the four-child SAH trees (the measured configuration
    builds it on the device) and on a strided 1/64 sample for the others, plus
    size-independent properties on all rays."""
    import torch
    from rs_illume_amd import workloads as W
    from rs_illume_amd.accel import BVHAccel, hits_to_numpy, rays_to_device, split_hits
    n_tris, n_rays = 1 << 20, 1 << 24
    _say("case full-size", n_tris, n_rays, "width", width, "builder", builder)
    v, idx = W.triangle_soup(n_tris, seed=1)
    rays = W.sphere_to_cube_rays(n_rays, seed=2)
    bvh = BVHAccel(v, idx, None, 4, split_method=builder, width=width)
    _say("bvh", bvh.info())
    rd = rays_to_device(rays)
    hd = bvh.intersect(rd)
    od = bvh.intersect_p(rd)
    prim, t, b1, b2 = split_hits(hd)
    hit = prim != MISS
    # property: any-hit == closest found something (t_max = inf everywhere)
    assert torch.equal(hit, od.bool()), "any-hit / closest-hit disagree"
    # property: hit distances positive and finite, barycentrics in [0, 1]
    assert bool((t[hit] > 0).all()) and bool(torch.isfinite(t[hit]).all())
    assert bool(((b1[hit] >= 0) & (b1[hit] <= 1) & (b2[hit] >= 0) & (b2[hit] <= 1)).all())
    # property: determinism (bitwise identical on a second launch)
    hd2 = bvh.intersect(rd)
    assert torch.equal(hd, hd2)
    _say("gpu hit fraction", float(hit.float().mean()))
    sample = np.arange(0, n_rays, 1 if (width == 4 and builder in ("SAH", "SAH_GPU")) else 64)
    _say("oracle rays", len(sample))
    osc = oracle.OracleScene(v, idx, None, 4)
    rs = rays[sample]
    ref = osc.intersect(rs)
    occ = osc.intersect_p(rs)
    h = hits_to_numpy(hd)[sample]
    o = od.cpu().numpy()[sample]
    res = compare_closest(h, ref, osc, rs)
    _say(res)
    _assert_parity(res, o, occ)


@pytest.mark.parametrize("width", [2, 4])
def test_batch_sizes_across_wave_chunks(gpu, oracle, width):
    """Batch queries on the four-child tree run as persistent waves, each
    starting on a static run of the batch and taking further 64-ray runs from a
    device queue; the run sizes depend on the batch size and the device's wave
    slots (64 lanes x 4 SIMDs x 8 waves per compute unit). Every batch size —
    below one wave, exact multiples of 64, one ray either side of filling every
    wave slot — gives each ray exactly the hit and occlusion it gets in one
    large batch, and the large batch matches the oracle."""
    import torch
    from rs_illume_amd import workloads as W
    from rs_illume_amd.accel import BVHAccel, hits_to_numpy, rays_to_device
    v, idx, sph = small_scene(16384, 64, 111, median_sigma=0.02)
    bvh = BVHAccel(v, idx, sph, 4, width=width)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    full = 64 * 4 * 8 * cus
    n_max = full + 1
    _say("case batch sizes, width", width, "wave slots x 64 =", full)
    rays = W.sphere_to_cube_rays(n_max, seed=112)
    rd = rays_to_device(rays)
    h_all = bvh.intersect(rd)
    o_all = bvh.intersect_p(rd)
    sample = np.arange(0, n_max, 7)
    osc = oracle.OracleScene(v, idx, sph, 4)
    rs = rays[sample]
    res = compare_closest(hits_to_numpy(h_all)[sample], osc.intersect(rs), osc, rs)
    _say(res)
    _assert_parity(res, o_all.cpu().numpy()[sample], osc.intersect_p(rs))
    for n in (1, 2, 63, 64, 65, 1000, 4097, full // 2 + 3, full - 1, full):
        for off in (0, n_max - n):
            sub = rd[off:off + n].contiguous()
            h = bvh.intersect(sub)
            o = bvh.intersect_p(sub)
            assert torch.equal(h, h_all[off:off + n]), (n, off)
            assert torch.equal(o, o_all[off:off + n]), (n, off)


def test_intersect_in_captured_graph(gpu):
    """The batch calls (their stream-ordered queue allocation included) can be
    captured in a graph: every replay, with eager work and synchronisations
    between replays, gives the eager call's hits and occlusion."""
    import torch
    from rs_illume_amd import workloads as W
    from rs_illume_amd.accel import BVHAccel, rays_to_device
    v, idx, sph = small_scene(16384, 64, 121, median_sigma=0.02)
    _say("case graph capture")
    for width in (2, 4):
        bvh = BVHAccel(v, idx, sph, 4, width=width)
        rd = rays_to_device(W.sphere_to_cube_rays(300000, seed=122))
        h_ref, o_ref = bvh.intersect(rd), bvh.intersect_p(rd)
        hits = torch.empty_like(h_ref)
        occ = torch.empty_like(o_ref)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            bvh.intersect(rd, hits)
            bvh.intersect_p(rd, occ)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            bvh.intersect(rd, hits)
            bvh.intersect_p(rd, occ)
        for k in range(3):
            hits.fill_(-1)
            occ.fill_(7)
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(hits, h_ref), (width, k)
            assert torch.equal(occ, o_ref), (width, k)
            bvh.intersect(rd[:1000].contiguous())
            torch.cuda.synchronize()
        del g
        bvh.close()
