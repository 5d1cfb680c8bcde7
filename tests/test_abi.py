"""CPU checks of the C ABI library: it loads, exports every function that
include/illume_hip.h declares, and rejects invalid input before touching a GPU."""
import ctypes

import numpy as np
import pytest


def test_library_exports_every_declared_symbol():
    from rs_illume_amd import _lib
    L = _lib.lib()
    names = _lib.header_functions()
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_status_strings():
    from rs_illume_amd import _lib
    assert _lib.status_string(0) == "ok"
    assert "invalid" in _lib.status_string(1)
    assert _lib.status_string(99) == "unknown status"


def _create(verts, idx, spheres=None, leaf=4):
    from rs_illume_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    sph = np.zeros((0, 4), np.float32) if spheres is None else spheres
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a.size else None
    st = L.illume_bvh_create(p(verts), len(verts), p(idx), len(idx), p(sph), len(sph), leaf, ctypes.byref(h))
    return st, L.illume_last_error().decode()


@pytest.mark.parametrize("case", ["index_out_of_range", "nan_vertex", "bad_sphere", "leaf_too_big", "leaf_zero"])
def test_invalid_scene_rejected_before_device_work(case):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    sph = None
    leaf = 4
    expect = 1
    if case == "index_out_of_range":
        idx = np.array([[0, 1, 3]], np.uint32)
    elif case == "nan_vertex":
        v[1, 2] = np.nan
    elif case == "bad_sphere":
        sph = np.array([[0, 0, 0, -1.0]], np.float32)
    elif case == "leaf_too_big":
        leaf, expect = 16, 4
    elif case == "leaf_zero":
        leaf, expect = 0, 4
    st, msg = _create(v, idx, sph, leaf)
    assert st == expect, (st, msg)
    assert msg


def test_null_arguments():
    from rs_illume_amd import _lib
    L = _lib.lib()
    assert L.illume_bvh_create(None, 0, None, 0, None, 0, 4, None) == 1
    assert L.illume_intersect_closest(None, None, 1, None, None) == 1
    assert L.illume_intersect_any(None, None, 1, None, None) == 1
    assert L.illume_bvh_destroy(None) == 0


def test_ray_count_limit_checked_first():
    """Grid sizes and queue run ends are 32-bit: batches above 2^31 rays are
    refused (ILLUME_ERR_UNSUPPORTED) before any other check or device work."""
    from rs_illume_amd import _lib
    L = _lib.lib()
    big = (1 << 31) + 1
    assert L.illume_intersect_closest(None, None, big, None, None) == 4
    assert "2^31" in L.illume_last_error().decode()
    assert L.illume_intersect_any(None, None, big, None, None) == 4
    assert L.illume_intersect_counts(None, None, big, 0, None, None) == 4
    assert L.illume_intersect_closest(None, None, 1 << 31, None, None) == 1  # at the limit: next check fails


def test_kernels_only_in_hip_sources():
    """Every kernel is compiled for gfx950: kernels may only be defined in .hip
    files, and every source is compiled with the gfx950 device flags."""
    from rs_illume_amd import build_lib
    srcs = build_lib.sources()
    build_lib.check_host_sources(srcs)
    assert "--offload-arch=gfx950" in build_lib.DEVICE
