"""Reconstruction filters and the filtered film of the GPU path tracer.

``Filter`` wraps ``illume_filter_create`` (the Filter trait, core/filter.rs,
and its box / triangle / Gaussian / Mitchell / Lanczos-sinc implementations
in filters/*.rs, tabulated as Film::new does, core/film.rs:69-81).
``filter_desc`` fills in pbrt-v3's defaults for each name (the ones scene
files use). ``render_filtered`` runs ``illume_render_filtered`` and ``resolve`` the
film's normalisation (Film::write_image, film.rs:187-226).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .accel import _stream_handle

BOX, TRIANGLE, GAUSSIAN, MITCHELL, LANCZOS = range(5)
FILTER_DTYPE = np.dtype([("type", "<i4"), ("radius", "<f4", 2), ("alpha", "<f4"), ("b", "<f4"), ("c", "<f4"),
                         ("tau", "<f4")])
assert FILTER_DTYPE.itemsize == 28
NAMES = {"box": BOX, "triangle": TRIANGLE, "gaussian": GAUSSIAN, "mitchell": MITCHELL, "sinc": LANCZOS,
         "lanczos": LANCZOS}
# pbrt-v3's defaults per filter: radius (xwidth = ywidth) and shape parameters
DEFAULTS = {BOX: dict(radius=0.5), TRIANGLE: dict(radius=2.0), GAUSSIAN: dict(radius=2.0, alpha=2.0),
            MITCHELL: dict(radius=2.0, b=1.0 / 3.0, c=1.0 / 3.0), LANCZOS: dict(radius=4.0, tau=3.0)}


def filter_desc(kind, radius=None, alpha=None, b=None, c=None, tau=None):
    """Filter description record (the C struct illume_filter_desc). ``kind`` is
    a type number or one of box, triangle, gaussian, mitchell, sinc."""
    t = NAMES[kind] if isinstance(kind, str) else int(kind)
    if t not in DEFAULTS:
        raise ValueError("unknown filter %r" % (kind,))
    dflt = dict(radius=0.5, alpha=2.0, b=1.0 / 3.0, c=1.0 / 3.0, tau=3.0)
    dflt.update(DEFAULTS[t])
    r = dflt["radius"] if radius is None else radius
    r = (r, r) if np.isscalar(r) else tuple(r)
    d = np.zeros(1, FILTER_DTYPE)
    d["type"] = t
    d["radius"] = r
    d["alpha"] = dflt["alpha"] if alpha is None else alpha
    d["b"] = dflt["b"] if b is None else b
    d["c"] = dflt["c"] if c is None else c
    d["tau"] = dflt["tau"] if tau is None else tau
    return d


def _declare():
    L = _lib.lib()
    if getattr(L, "_film_declared", False):
        return L
    vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
    from .render import RenderCfg
    L.illume_filter_create.restype = i32
    L.illume_filter_create.argtypes = [vp, ctypes.POINTER(vp)]
    L.illume_filter_destroy.restype = i32
    L.illume_filter_destroy.argtypes = [vp]
    L.illume_filter_table.restype = i32
    L.illume_filter_table.argtypes = [vp, vp]
    L.illume_render_filtered.restype = i32
    L.illume_render_filtered.argtypes = [vp, vp, ctypes.POINTER(RenderCfg), vp, vp, vp, vp]
    L.illume_film_resolve.restype = i32
    L.illume_film_resolve.argtypes = [vp, u64, ctypes.c_float, vp, vp]
    L._film_declared = True
    return L


class Filter:
    """A reconstruction filter resident on the device (its 16 x 16 table)."""

    def __init__(self, kind="box", **params):
        self.desc = filter_desc(kind, **params)
        h = ctypes.c_void_p()
        _lib.check(_declare().illume_filter_create(self.desc.ctypes.data, ctypes.byref(h)), "illume_filter_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def radius(self):
        return tuple(float(v) for v in self.desc["radius"][0])

    def table(self):
        out = np.zeros(256, np.float32)
        _lib.check(_declare().illume_filter_table(self._h, out.ctypes.data), "illume_filter_table")
        return out.reshape(16, 16)

    def close(self):
        if getattr(self, "_h", None):
            _declare().illume_filter_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_filtered(integrator, scene, filt, rgbw=None, spp=None, sample_offset=0, ray_counts=None, stream=None):
    """Adds the filtered film sums of ``spp`` samples per sample pixel to rgbw
    (device float32 (H, W, 4): weighted RGB and weight) and returns it.
    ``integrator`` supplies camera, max_depth, seed and sampler: a
    PathIntegrator (megakernel) or a WavefrontPathIntegrator (its pool, scene
    default and proposal)."""
    from .render import RenderCfg, WavefrontPathIntegrator, _declare_wf, _p
    w, h = integrator.resolution
    if rgbw is None:
        rgbw = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
    if rgbw.shape != (h, w, 4) or rgbw.dtype != torch.float32 or not rgbw.is_cuda or not rgbw.is_contiguous():
        raise ValueError("rgbw must be a contiguous cuda float32 tensor of shape (%d, %d, 4)" % (h, w))
    cfg = RenderCfg(integrator.spp if spp is None else int(spp), integrator.max_depth, int(sample_offset),
                    integrator.flags, integrator.seed)
    rc = ctypes.c_void_p(ray_counts.data_ptr()) if ray_counts is not None else None
    if isinstance(integrator, WavefrontPathIntegrator):
        sc = scene if scene is not None else integrator.scene
        g = integrator.guide.handle if integrator.guide is not None else None
        _lib.check(_declare_wf().illume_wavefront_render_filtered(
            integrator.wf.handle, sc.handle, _p(integrator.camera), ctypes.byref(cfg), filt.handle, g,
            ctypes.byref(integrator.gcfg), ctypes.c_void_p(rgbw.data_ptr()), rc, _stream_handle(stream)),
            "illume_wavefront_render_filtered")
        return rgbw
    _declare()
    _lib.check(_lib.lib().illume_render_filtered(scene.handle, _p(integrator.camera), ctypes.byref(cfg), filt.handle,
                                                 ctypes.c_void_p(rgbw.data_ptr()), rc, _stream_handle(stream)),
               "illume_render_filtered")
    return rgbw


def resolve(rgbw, scale=1.0, stream=None):
    """Film::write_image's normalisation: (H, W, 3) = scale * max(0, rgb / w)
    where w != 0 (unnormalised where w == 0)."""
    if not rgbw.is_cuda or rgbw.dtype != torch.float32 or not rgbw.is_contiguous() or rgbw.shape[-1] != 4:
        raise ValueError("rgbw must be a contiguous cuda float32 tensor (..., 4)")
    out = torch.empty(rgbw.shape[:-1] + (3,), dtype=torch.float32, device=rgbw.device)
    n = rgbw.numel() // 4
    _lib.check(_declare().illume_film_resolve(ctypes.c_void_p(rgbw.data_ptr()), n, float(scale),
                                              ctypes.c_void_p(out.data_ptr()), _stream_handle(stream)),
               "illume_film_resolve")
    return out
