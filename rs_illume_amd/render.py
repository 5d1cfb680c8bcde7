"""GPU path tracer with the reference's Integrator interface.

``Scene`` mirrors rs_illume's ``Scene`` (core/scene.rs:10-53): aggregate +
lights, with ``intersect`` / ``intersect_p``. ``PathIntegrator`` implements
``Integrator::render`` (core/integrator.rs:17-29) with pbrt-v3's path-tracing
``li`` (the reference declares ``SamplerIntegrator::li``, integrator.rs:31-35,
but has no path integrator). All work runs in librs_illume_amd.so.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .accel import _stream_handle
from .scenes import CAMERA_DTYPE, MATERIAL_DTYPE


# illume_render_cfg.flags per sampler: PCG32 streams, or pbrt-v3's Halton or Sobol' sampler
SAMPLERS = {"random": 0, "halton": 1, "sobol": 2}


class RenderCfg(ctypes.Structure):
    _fields_ = [("spp", ctypes.c_uint32), ("max_depth", ctypes.c_uint32), ("sample_offset", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("seed", ctypes.c_uint64)]


class SceneOpts(ctypes.Structure):
    _fields_ = [("sky", ctypes.c_void_p), ("bvh_builder", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def _declare():
    L = _lib.lib()
    if getattr(L, "_render_declared", False):
        return L
    vp, u32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
    L.illume_scene_create.restype = i32
    L.illume_scene_create.argtypes = [vp, u32, vp, u32, vp, u32, vp, vp, vp, u32, ctypes.POINTER(vp)]
    L.illume_scene_create_ex.restype = i32
    L.illume_scene_create_ex.argtypes = [vp, u32, vp, u32, vp, u32, vp, vp, vp, u32, ctypes.POINTER(SceneOpts),
                                         ctypes.POINTER(vp)]
    L.illume_scene_destroy.restype = i32
    L.illume_scene_destroy.argtypes = [vp]
    L.illume_scene_bvh.restype = i32
    L.illume_scene_bvh.argtypes = [vp, ctypes.POINTER(vp)]
    L.illume_render.restype = i32
    L.illume_render.argtypes = [vp, vp, ctypes.POINTER(RenderCfg), vp, vp, vp]
    L.illume_li_batch.restype = i32
    L.illume_li_batch.argtypes = [vp, vp, u32, ctypes.POINTER(RenderCfg), vp, vp, vp]
    L.illume_render_ao.restype = i32
    L.illume_render_ao.argtypes = [vp, vp, ctypes.POINTER(RenderCfg), u32, ctypes.c_int32, vp, vp, vp]
    L._render_declared = True
    return L


# illume_scene_opts.bvh_builder codes
BUILDERS = {"SAH": 0, "LBVH": 1, "SAH_GPU": 2}


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a.size else None


class Scene:
    """Scene::new(aggregate, lights) over a :class:`scenes.SceneDesc` (with its
    analytic sky, if it has one). ``builder``: "SAH" (binned SAH on the host),
    "SAH_GPU" (the same tree built on the device) or "LBVH" (device)."""

    def __init__(self, desc, builder="SAH"):
        if not torch.cuda.is_available():
            raise RuntimeError("Scene needs a ROCm GPU")
        a = desc.arrays() if hasattr(desc, "arrays") else desc
        self.arrays = a
        L = _declare()
        torch.cuda.init()
        h = ctypes.c_void_p()
        if builder not in BUILDERS:
            raise ValueError("builder must be one of %s" % (tuple(BUILDERS),))
        sky = a.get("sky")
        self._sky = None if sky is None else np.ascontiguousarray(sky, np.float32)
        opts = SceneOpts(self._sky.ctypes.data if self._sky is not None else None, BUILDERS[builder], 0)
        _lib.check(L.illume_scene_create_ex(_p(a["verts"]), len(a["verts"]), _p(a["idx"]), len(a["idx"]),
                                            _p(a["spheres"]), len(a["spheres"]), _p(a["prim_material"]),
                                            _p(a["prim_emission"]), _p(a["materials"]), len(a["materials"]),
                                            ctypes.byref(opts), ctypes.byref(h)), "illume_scene_create_ex")
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            _declare().illume_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _bvh(self):
        b = ctypes.c_void_p()
        _lib.check(_declare().illume_scene_bvh(self._h, ctypes.byref(b)), "illume_scene_bvh")
        return b

    def bvh_info(self):
        """illume_bvh_info of the scene's aggregate (node counts, depth, build
        time)."""
        inf = _lib.BvhInfo()
        _lib.check(_lib.lib().illume_bvh_info(self._bvh(), ctypes.byref(inf)), "illume_bvh_info")
        return {"n_prims": inf.n_prims, "n_inner_nodes": inf.n_inner_nodes, "n_leaves": inf.n_leaves,
                "max_depth": inf.max_depth, "device_bytes": inf.device_bytes, "build_seconds": inf.build_seconds}

    def intersect(self, rays, hits=None, stream=None):
        """Scene::intersect (scene.rs:43-47), batched: see BVHAccel.intersect."""
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.int32, device=rays.device)
        _lib.check(_lib.lib().illume_intersect_closest(self._bvh(), ctypes.c_void_p(rays.data_ptr()), n,
                                                       ctypes.c_void_p(hits.data_ptr()), _stream_handle(stream)),
                   "illume_intersect_closest")
        return hits

    def intersect_p(self, rays, occluded=None, stream=None):
        """Scene::intersect_p (scene.rs:49-53), batched."""
        n = rays.shape[0]
        if occluded is None:
            occluded = torch.empty((n,), dtype=torch.uint8, device=rays.device)
        _lib.check(_lib.lib().illume_intersect_any(self._bvh(), ctypes.c_void_p(rays.data_ptr()), n,
                                                   ctypes.c_void_p(occluded.data_ptr()), _stream_handle(stream)),
                   "illume_intersect_any")
        return occluded


class PathIntegrator:
    """Integrator::render with pbrt-v3's PathIntegrator::li (max_depth,
    Russian roulette after 3 bounces, one light sample per vertex with MIS)."""

    def __init__(self, camera, spp=16, max_depth=5, seed=0, sampler="random"):
        self.camera = np.ascontiguousarray(camera, CAMERA_DTYPE).reshape(1)
        self.spp = int(spp)
        self.max_depth = int(max_depth)
        self.seed = int(seed)
        if sampler not in SAMPLERS:
            raise ValueError("sampler must be one of %s" % (tuple(SAMPLERS),))
        self.sampler = sampler

    @property
    def flags(self):
        return SAMPLERS[self.sampler]

    @property
    def resolution(self):
        return int(self.camera["width"][0]), int(self.camera["height"][0])

    def render_sum(self, scene, rgb_sum=None, spp=None, sample_offset=0, ray_counts=None, stream=None):
        """Adds per-pixel radiance sums of `spp` samples (indices
        sample_offset..) to rgb_sum (device float32 (H, W, 3))."""
        w, h = self.resolution
        if rgb_sum is None:
            rgb_sum = torch.zeros((h, w, 3), dtype=torch.float32, device="cuda")
        cfg = RenderCfg(self.spp if spp is None else int(spp), self.max_depth, int(sample_offset), self.flags,
                        self.seed)
        rc = ctypes.c_void_p(ray_counts.data_ptr()) if ray_counts is not None else None
        _lib.check(_declare().illume_render(scene.handle, _p(self.camera), ctypes.byref(cfg),
                                            ctypes.c_void_p(rgb_sum.data_ptr()), rc, _stream_handle(stream)),
                   "illume_render")
        return rgb_sum

    def render(self, scene, stream=None):
        """The image (H, W, 3) float32 device tensor: mean radiance per pixel."""
        s = self.render_sum(scene, stream=stream)
        return s / float(self.spp)

    def li(self, scene, rays, sample=0, ray_counts=None, stream=None):
        """SamplerIntegrator::li (core/integrator.rs:31-35) of each ray of a
        device float32 (n, 8) ray tensor -> (n, 3) radiance; ray i uses the
        sampler stream (seed, i, sample) after a camera sample's five
        dimensions (illume_li_batch). For parity tests; renders use render()."""
        n = rays.shape[0]
        out = torch.empty((n, 3), dtype=torch.float32, device=rays.device)
        cfg = RenderCfg(1, self.max_depth, int(sample), 0, self.seed)
        rc = ctypes.c_void_p(ray_counts.data_ptr()) if ray_counts is not None else None
        _lib.check(_declare().illume_li_batch(scene.handle, ctypes.c_void_p(rays.data_ptr()), n, ctypes.byref(cfg),
                                              ctypes.c_void_p(out.data_ptr()), rc, _stream_handle(stream)),
                   "illume_li_batch")
        return out


class AOIntegrator(PathIntegrator):
    """AOIntegrator (integrators/ao.rs:16-112): ``AOIntegrator::new(cos_sample,
    ns, camera, sampler, pixel_bounds)`` then ``render``. Each camera sample's
    value is the ambient occlusion at its closest hit, estimated with
    ``n_samples`` hemisphere directions (cosine-distributed if ``cos_sample``):
    pi for an unoccluded point, 0 for a miss; grey in all three channels."""

    def __init__(self, camera, spp=16, n_samples=64, cos_sample=True, seed=0, sampler="random"):
        super().__init__(camera, spp, 0, seed, sampler)
        self.n_samples = int(n_samples)
        self.cos_sample = bool(cos_sample)

    def render_sum(self, scene, rgb_sum=None, spp=None, sample_offset=0, ray_counts=None, stream=None):
        w, h = self.resolution
        if rgb_sum is None:
            rgb_sum = torch.zeros((h, w, 3), dtype=torch.float32, device="cuda")
        cfg = RenderCfg(self.spp if spp is None else int(spp), 0, int(sample_offset), self.flags, self.seed)
        rc = ctypes.c_void_p(ray_counts.data_ptr()) if ray_counts is not None else None
        _lib.check(_declare().illume_render_ao(scene.handle, _p(self.camera), ctypes.byref(cfg), self.n_samples,
                                               1 if self.cos_sample else 0, ctypes.c_void_p(rgb_sum.data_ptr()), rc,
                                               _stream_handle(stream)), "illume_render_ao")
        return rgb_sum


GUIDE_PRODUCT = 1  # illume_guide_cfg.flags: proposal x vMF fit of the BSDF lobe


class GuideCfg(ctypes.Structure):
    _fields_ = [("guide_frac", ctypes.c_float), ("train", ctypes.c_uint32), ("train_batch", ctypes.c_uint32),
                ("min_alpha", ctypes.c_float), ("flags", ctypes.c_uint32)]


class CameraC(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float * 3), ("look", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("fov", ctypes.c_float), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32)]


class JobC(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("camera", CameraC), ("cfg", RenderCfg), ("film_offset", ctypes.c_uint64)]


def _camera_c(cam):
    cam = np.ascontiguousarray(cam, CAMERA_DTYPE).reshape(1)
    c = CameraC()
    c.pos[:] = [float(x) for x in cam["pos"][0]]
    c.look[:] = [float(x) for x in cam["look"][0]]
    c.up[:] = [float(x) for x in cam["up"][0]]
    c.fov = float(cam["fov"][0])
    c.width = int(cam["width"][0])
    c.height = int(cam["height"][0])
    return c


def _job_array(jobs, per_px):
    """illume_job array of (Scene, camera, spp, max_depth, seed[, sample_offset[,
    sampler]]) tuples, images laid out one after another (per_px floats per
    pixel): (array, offsets, total floats)."""
    arr = (JobC * len(jobs))()
    offsets, off = [], 0
    for i, jb in enumerate(jobs):
        sc, cam, spp, depth, seed = jb[:5]
        so = jb[5] if len(jb) > 5 else 0
        smp = jb[6] if len(jb) > 6 else "random"
        if smp not in SAMPLERS:
            raise ValueError("sampler must be one of %s" % (tuple(SAMPLERS),))
        c = _camera_c(cam)
        arr[i].scene = sc.handle
        arr[i].camera = c
        arr[i].cfg = RenderCfg(int(spp), int(depth), int(so), SAMPLERS[smp], int(seed))
        arr[i].film_offset = off
        offsets.append(off)
        off += c.width * c.height * per_px
    return arr, offsets, off


def render_batch(jobs, film=None, ray_counts=None, stream=None):
    """Plain renders of many scenes and cameras on the megakernel in one
    launch (illume_render_batch). jobs as for :meth:`Wavefront.render_batch`;
    returns (film, offsets) with every job's radiance sums (H*W*3 floats) at
    its offset."""
    L = _declare_wf()
    arr, offsets, off = _job_array(jobs, 3)
    if film is None:
        film = torch.zeros(off, dtype=torch.float32, device="cuda")
    rc = ctypes.c_void_p(ray_counts.data_ptr()) if ray_counts is not None else None
    _lib.check(L.illume_render_batch(arr, len(jobs), ctypes.c_void_p(film.data_ptr()), rc, _stream_handle(stream)),
               "illume_render_batch")
    return film, offsets


def _declare_wf():
    L = _declare()
    if getattr(L, "_wf_declared", False):
        return L
    vp, u32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
    L.illume_wavefront_create.restype = i32
    L.illume_wavefront_create.argtypes = [u32, u32, ctypes.POINTER(vp)]
    L.illume_wavefront_destroy.restype = i32
    L.illume_wavefront_destroy.argtypes = [vp]
    L.illume_wavefront_set_tuning.restype = i32
    L.illume_wavefront_set_tuning.argtypes = [vp, u32, u32, u32]
    L.illume_wavefront_last_iterations.restype = ctypes.c_uint64
    L.illume_wavefront_last_iterations.argtypes = [vp]
    L.illume_wavefront_render.restype = i32
    L.illume_wavefront_render.argtypes = [vp, vp, vp, ctypes.POINTER(RenderCfg), vp, ctypes.POINTER(GuideCfg), vp, vp,
                                          vp]
    L.illume_wavefront_render_batch.restype = i32
    L.illume_wavefront_render_batch.argtypes = [vp, ctypes.POINTER(JobC), u32, vp, ctypes.POINTER(GuideCfg), vp, vp, vp]
    L.illume_wavefront_render_batch_guides.restype = i32
    L.illume_wavefront_render_batch_guides.argtypes = [vp, ctypes.POINTER(JobC), u32, ctypes.POINTER(vp),
                                                       ctypes.POINTER(GuideCfg), vp, vp, vp]
    L.illume_wavefront_render_batch_filtered.restype = i32
    L.illume_wavefront_render_batch_filtered.argtypes = [vp, ctypes.POINTER(JobC), ctypes.POINTER(vp), u32, vp,
                                                         ctypes.POINTER(GuideCfg), vp, vp, vp]
    L.illume_render_batch.restype = i32
    L.illume_render_batch.argtypes = [ctypes.POINTER(JobC), u32, vp, vp, vp]
    L.illume_wavefront_render_filtered.restype = i32
    L.illume_wavefront_render_filtered.argtypes = [vp, vp, vp, ctypes.POINTER(RenderCfg), vp, vp,
                                                   ctypes.POINTER(GuideCfg), vp, vp, vp]
    L._wf_declared = True
    return L


class Wavefront:
    """Pool of ``max_paths`` path slots in HBM with the wavefront pipeline
    (regenerate, extend, infer, shade) replayed from hipGraphs."""

    def __init__(self, max_paths=1 << 20, max_depth=8, block=256, regen_min=None, graph_iters=8):
        L = _declare_wf()
        torch.cuda.init()
        h = ctypes.c_void_p()
        _lib.check(L.illume_wavefront_create(int(max_paths), int(max_depth), ctypes.byref(h)),
                   "illume_wavefront_create")
        self._h = h
        self.max_paths = int(max_paths)
        self.max_depth = int(max_depth)
        self.tune(block, regen_min, graph_iters)

    def tune(self, block=256, regen_min=None, graph_iters=8):
        # None: regenerate every iteration (the pool clamps to its capacity + 1)
        rm = 0xFFFFFFFF if regen_min is None else int(regen_min)
        _lib.check(_declare_wf().illume_wavefront_set_tuning(self._h, int(block), rm, int(graph_iters)),
                   "illume_wavefront_set_tuning")
        self.params = {"block": int(block), "regen_min": rm, "graph_iters": int(graph_iters)}

    @property
    def handle(self):
        return self._h

    def last_iterations(self):
        return int(_declare_wf().illume_wavefront_last_iterations(self._h))

    def close(self):
        if getattr(self, "_h", None):
            _declare_wf().illume_wavefront_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render_batch(self, jobs, film=None, guide=None, gcfg=None, ray_counts=None, stream=None, filters=None,
                     guides=None):
        """jobs: list of (Scene, camera, spp, max_depth, seed[, sample_offset[,
        sampler]]), sampler one of ``SAMPLERS`` (default "random").
        Returns (film, offsets): film is a flat float32 device tensor holding
        every job's radiance sums (H*W*3 each) at its offset. With
        ``filters`` (one :class:`rs_illume_amd.film.Filter` per job) each
        job's image is its filtered film, H*W*4 floats (weighted RGB and
        weight; :func:`rs_illume_amd.film.resolve` gives the pixels). With
        ``guides`` (one :class:`rs_illume_amd.guide.Guide` or None per job)
        each job has its own proposal network, trained on its own samples
        (illume_wavefront_render_batch_guides)."""
        if filters is not None and len(filters) != len(jobs):
            raise ValueError("one filter per job")
        if guides is not None:
            if len(guides) != len(jobs) or filters is not None or guide is not None:
                raise ValueError("guides: one per job, without filters or a shared guide")
            arr, offsets, off = _job_array(jobs, 3)
            if film is None:
                film = torch.zeros(off, dtype=torch.float32, device="cuda")
            gh = (ctypes.c_void_p * len(jobs))(*[(g.handle.value if g is not None else None) for g in guides])
            gc = gcfg if gcfg is not None else GuideCfg(0.0, 0, 0, 0.0)
            rc = ctypes.c_void_p(ray_counts.data_ptr()) if ray_counts is not None else None
            _lib.check(_declare_wf().illume_wavefront_render_batch_guides(
                self._h, arr, len(jobs), gh, ctypes.byref(gc), ctypes.c_void_p(film.data_ptr()), rc,
                _stream_handle(stream)), "illume_wavefront_render_batch_guides")
            return film, offsets
        arr, offsets, off = _job_array(jobs, 3 if filters is None else 4)
        if film is None:
            film = torch.zeros(off, dtype=torch.float32, device="cuda")
        g = guide.handle if guide is not None else None
        gc = gcfg if gcfg is not None else GuideCfg(0.0, 0, 0, 0.0)
        rc = ctypes.c_void_p(ray_counts.data_ptr()) if ray_counts is not None else None
        if filters is None:
            _lib.check(_declare_wf().illume_wavefront_render_batch(self._h, arr, len(jobs), g, ctypes.byref(gc),
                                                                   ctypes.c_void_p(film.data_ptr()), rc,
                                                                   _stream_handle(stream)),
                       "illume_wavefront_render_batch")
        else:
            fh = (ctypes.c_void_p * len(jobs))(*[f.handle.value for f in filters])
            _lib.check(_declare_wf().illume_wavefront_render_batch_filtered(
                self._h, arr, fh, len(jobs), g, ctypes.byref(gc), ctypes.c_void_p(film.data_ptr()), rc,
                _stream_handle(stream)), "illume_wavefront_render_batch_filtered")
        return film, offsets


class WavefrontPathIntegrator(PathIntegrator):
    """The same Integrator::render as :class:`PathIntegrator`, run on a
    :class:`Wavefront` pool; optionally guided by a
    :class:`rs_illume_amd.guide.Guide` (one-sample MIS with probability
    ``guide_frac``) and training it online from finished paths (``train`` =
    the most Adam steps per graph replay, True = 16, 0/False = frozen). The proposal is used
    only at diffuse vertices and at microfacet vertices with alpha >=
    ``min_alpha``."""

    def __init__(self, scene, camera, spp=16, max_depth=5, seed=0, max_paths=1 << 20, guide=None, guide_frac=0.5,
                 train=False, train_batch=65536, wavefront=None, min_alpha=0.0, sampler="random", product=False):
        super().__init__(camera, spp, max_depth, seed, sampler)
        self.scene = scene
        self.guide = guide
        steps = 16 if train is True else int(train)
        self.gcfg = GuideCfg(float(guide_frac) if guide is not None else 0.0, steps, int(train_batch),
                             float(min_alpha), GUIDE_PRODUCT if product else 0)
        self.wf = wavefront if wavefront is not None else Wavefront(max_paths, max_depth)

    def render_sum(self, scene=None, rgb_sum=None, spp=None, sample_offset=0, ray_counts=None, stream=None):
        w, h = self.resolution
        if rgb_sum is None:
            rgb_sum = torch.zeros((h, w, 3), dtype=torch.float32, device="cuda")
        sc = scene if scene is not None else self.scene
        cfg = RenderCfg(self.spp if spp is None else int(spp), self.max_depth, int(sample_offset), self.flags,
                        self.seed)
        rc = ctypes.c_void_p(ray_counts.data_ptr()) if ray_counts is not None else None
        g = self.guide.handle if self.guide is not None else None
        _lib.check(_declare_wf().illume_wavefront_render(self.wf.handle, sc.handle, _p(self.camera), ctypes.byref(cfg),
                                                         g, ctypes.byref(self.gcfg),
                                                         ctypes.c_void_p(rgb_sum.data_ptr()), rc,
                                                         _stream_handle(stream)), "illume_wavefront_render")
        return rgb_sum

    def render(self, scene=None, stream=None):
        return self.render_sum(scene, stream=stream) / float(self.spp)
