"""GPU aggregate with the reference's Primitive interface.

``BVHAccel`` mirrors rs_illume's ``BVHAccel`` (accelerators/bvh.rs:24-93) in
its role as the scene aggregate (``impl Primitive for BVHAccel``,
bvh.rs:540-663): ``world_bound``, ``intersect`` (closest hit) and
``intersect_p`` (any hit). Queries are batched: one call traces a device
array of rays. Device memory and streams come from torch; all work runs in
librs_illume_amd.so.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .workloads import HIT_DTYPE, MISS, RAY_DTYPE

SPLIT_METHODS = ("SAH", "LBVH", "SAH_GPU")


def _stream_handle(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def rays_to_device(rays_np, device="cuda"):
    """Structured RAY_DTYPE array -> device tensor (n, 8) float32."""
    a = np.ascontiguousarray(rays_np, dtype=RAY_DTYPE).view(np.float32).reshape(-1, 8)
    return torch.from_numpy(a).to(device)


def hits_to_numpy(hits_dev):
    """Device tensor (n, 4) int32 -> structured HIT_DTYPE array."""
    h = hits_dev.cpu().numpy()
    return np.ascontiguousarray(h).view(HIT_DTYPE).reshape(-1)


class BVHAccel:
    """Binned-SAH BVH over triangles and full spheres, resident in HBM.

    Arguments follow ``BVHAccel::new(prims, max_prims_in_node, split_method)``
    (bvh.rs:32): the primitives are given as arrays — ``verts`` float32 [V,3],
    ``idx`` uint32 [T,3] and optional ``spheres`` float32 [S,4] (cx, cy, cz, r).
    Primitive ids are triangles 0..T-1 then spheres T..T+S-1. ``width``
    selects 64-byte two-child (2) or 128-byte four-child (4) nodes.
    ``split_method`` "SAH" builds the binned-SAH tree on the host, "SAH_GPU"
    the same tree on the device, "LBVH" a linear BVH over Morton codes on the
    device (as the reference's HLBVH option, bvh.rs:249-502); hits are
    identical.
    """

    def __init__(self, verts, idx, spheres=None, max_prims_in_node=4, split_method="SAH", width=2):
        if split_method not in SPLIT_METHODS:
            raise ValueError("split_method must be one of %s" % (SPLIT_METHODS,))
        if not torch.cuda.is_available():
            raise RuntimeError("BVHAccel needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
        self.idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 3)
        sph = np.zeros((0, 4), np.float32) if spheres is None else spheres
        self.spheres = np.ascontiguousarray(sph, dtype=np.float32).reshape(-1, 4)
        self.n_tris = len(self.idx)
        self.n_spheres = len(self.spheres)
        L = _lib.lib()
        torch.cuda.init()
        h = ctypes.c_void_p()
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a.size else None
        if split_method == "SAH_GPU":
            _lib.check(L.illume_bvh_create_gpu_sah(p(self.verts), len(self.verts), p(self.idx), self.n_tris,
                                                   p(self.spheres), self.n_spheres, int(max_prims_in_node),
                                                   int(width), _stream_handle(None), ctypes.byref(h)),
                       "illume_bvh_create_gpu_sah")
        elif split_method == "LBVH":
            _lib.check(L.illume_bvh_create_gpu(p(self.verts), len(self.verts), p(self.idx), self.n_tris,
                                               p(self.spheres), self.n_spheres, int(max_prims_in_node), int(width),
                                               _stream_handle(None), ctypes.byref(h)), "illume_bvh_create_gpu")
        else:
            _lib.check(L.illume_bvh_create_ex(p(self.verts), len(self.verts), p(self.idx), self.n_tris,
                                              p(self.spheres), self.n_spheres, int(max_prims_in_node), int(width),
                                              ctypes.byref(h)), "illume_bvh_create_ex")
        self.width = int(width)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().illume_bvh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self):
        inf = _lib.BvhInfo()
        _lib.check(_lib.lib().illume_bvh_info(self._h, ctypes.byref(inf)), "illume_bvh_info")
        return {"n_prims": inf.n_prims, "n_inner_nodes": inf.n_inner_nodes, "n_leaves": inf.n_leaves,
                "max_depth": inf.max_depth, "device_bytes": inf.device_bytes,
                "build_seconds": inf.build_seconds, "bounds": list(inf.bounds)}

    def world_bound(self):
        """(min xyz, max xyz) of the scene (Primitive::world_bound)."""
        b = self.info()["bounds"]
        return np.array(b[:3], np.float32), np.array(b[3:], np.float32)

    @staticmethod
    def _check_rays(rays):
        if not (rays.is_cuda and rays.dtype == torch.float32 and rays.dim() == 2 and rays.shape[1] == 8
                and rays.is_contiguous()):
            raise ValueError("rays must be a contiguous cuda float32 tensor of shape (n, 8)")
        if rays.shape[0] >= 2 ** 32:
            raise ValueError("at most 2^32-1 rays per call")

    def intersect(self, rays, hits=None, stream=None):
        """Closest hit of each ray -> int32 tensor (n, 4): prim id bits, t, b1,
        b2 (float bits). Use :func:`hits_to_numpy` or :func:`split_hits`."""
        self._check_rays(rays)
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.int32, device=rays.device)
        _lib.check(_lib.lib().illume_intersect_closest(self._h, ctypes.c_void_p(rays.data_ptr()), n,
                                                       ctypes.c_void_p(hits.data_ptr()), _stream_handle(stream)),
                   "illume_intersect_closest")
        return hits

    def intersect_p(self, rays, occluded=None, stream=None):
        """Any hit of each ray -> uint8 tensor (n,): 1 if occluded."""
        self._check_rays(rays)
        n = rays.shape[0]
        if occluded is None:
            occluded = torch.empty((n,), dtype=torch.uint8, device=rays.device)
        _lib.check(_lib.lib().illume_intersect_any(self._h, ctypes.c_void_p(rays.data_ptr()), n,
                                                   ctypes.c_void_p(occluded.data_ptr()), _stream_handle(stream)),
                   "illume_intersect_any")
        return occluded

    def intersect_counts(self, rays, any_hit=False, stream=None):
        """Per-ray (box tests, primitive tests) of the closest- or any-hit
        traversal -> int32 tensor (n, 2)."""
        self._check_rays(rays)
        n = rays.shape[0]
        counts = torch.empty((n, 2), dtype=torch.int32, device=rays.device)
        _lib.check(_lib.lib().illume_intersect_counts(self._h, ctypes.c_void_p(rays.data_ptr()), n, int(any_hit),
                                                      ctypes.c_void_p(counts.data_ptr()), _stream_handle(stream)),
                   "illume_intersect_counts")
        return counts

    # numpy conveniences (host buffers; PCIe copies included)
    def intersect_np(self, rays_np):
        return hits_to_numpy(self.intersect(rays_to_device(rays_np)))

    def intersect_p_np(self, rays_np):
        return self.intersect_p(rays_to_device(rays_np)).cpu().numpy()


SHAPES = {"triangle": 0, "sphere": 1}


def shape_intersect(kind, shape, rays, any_hit=False, hits=None, stream=None):
    """Shape::intersect (or, any_hit, Shape::intersect_p) of one shape
    (core/shape.rs:8-112) for a batch of rays: ``kind`` "triangle" (shape =
    9 floats, p0 p1 p2) or "sphere" (centre, radius). -> int32 (n, 4) hits as
    :meth:`BVHAccel.intersect` gives them, prim id 0 on a hit."""
    BVHAccel._check_rays(rays)
    n = rays.shape[0]
    sh = np.ascontiguousarray(shape, np.float32).reshape(-1)
    if hits is None:
        hits = torch.empty((n, 4), dtype=torch.int32, device=rays.device)
    L = _lib.lib()
    if not getattr(L, "_shape_declared", False):
        L.illume_shape_intersect_batch.restype = ctypes.c_int
        L.illume_shape_intersect_batch.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                   ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
        L._shape_declared = True
    _lib.check(L.illume_shape_intersect_batch(SHAPES[kind], sh.ctypes.data_as(ctypes.c_void_p),
                                              ctypes.c_void_p(rays.data_ptr()), n, 1 if any_hit else 0,
                                              ctypes.c_void_p(hits.data_ptr()), _stream_handle(stream)),
               "illume_shape_intersect_batch")
    return hits


def split_hits(hits_dev):
    """int32 (n,4) device hits -> (prim int64 with MISS as 0xFFFFFFFF, t, b1, b2) tensors."""
    prim = hits_dev[:, 0].to(torch.int64) & 0xFFFFFFFF
    f = hits_dev.view(torch.float32)
    return prim, f[:, 1], f[:, 2], f[:, 3]


__all__ = ["BVHAccel", "rays_to_device", "hits_to_numpy", "split_hits", "shape_intersect", "MISS", "RAY_DTYPE",
           "HIT_DTYPE"]
