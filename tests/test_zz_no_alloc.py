"""illume_intersect_closest_noalloc: the closest-hit query that allocates no
device memory (its ray-queue counter lives in the library's per-device
scratch array) returns what illume_intersect_closest returns, bit for bit, on
both tree widths, over two calls in a row with the second one smaller (the
counter is reset by every call), and rejects batches above the limit before
any device work."""
import ctypes

import numpy as np
import pytest


def test_noalloc_rejects_oversized_batch_on_host():
    from rs_illume_amd import _lib
    L = _lib.lib()
    st = L.illume_intersect_closest_noalloc(None, None, (1 << 31) + 1, None, None)
    assert st == 4, st
    assert "2^31" in L.illume_last_error().decode()
    assert L.illume_intersect_closest_noalloc(None, None, 1, None, None) == 1


@pytest.mark.gpu
@pytest.mark.parametrize("width", [4, 2])
def test_noalloc_equals_closest_twice_second_smaller(gpu, width):
    import torch
    from rs_illume_amd import _lib
    from rs_illume_amd import workloads as W
    from rs_illume_amd.accel import BVHAccel, rays_to_device
    L = _lib.lib()
    v, idx = W.triangle_soup(20000, seed=11, median_sigma=0.02)
    sph = W.spheres_in_cube(16, seed=12)
    bvh = BVHAccel(v, idx, sph, 4, width=width)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for n in (300000, 70001):
        print("noalloc width %d n %d" % (width, n), flush=True)
        rd = rays_to_device(W.sphere_to_cube_rays(n, seed=13 + n))
        ref = bvh.intersect(rd)
        out = torch.full((n, 4), -7, dtype=torch.int32, device=gpu)
        _lib.check(L.illume_intersect_closest_noalloc(bvh.handle, ctypes.c_void_p(rd.data_ptr()), n,
                                                      ctypes.c_void_p(out.data_ptr()), stream),
                   "illume_intersect_closest_noalloc")
        torch.cuda.synchronize()
        assert torch.equal(out, ref), "width %d n %d: %d rows differ" % (width, n, int((out != ref).any(1).sum()))
        hit = ((ref[:, 0].to(torch.int64) & 0xFFFFFFFF) != 0xFFFFFFFF).float().mean().item()
        assert 0.05 < hit < 1.0
    # the size limit is checked with a real handle too, before any launch
    st = L.illume_intersect_closest_noalloc(bvh.handle, ctypes.c_void_p(rd.data_ptr()), (1 << 31) + 5,
                                            ctypes.c_void_p(out.data_ptr()), stream)
    assert st == 4
    bvh.close()
