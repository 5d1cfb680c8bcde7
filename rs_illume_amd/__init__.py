"""rs_illume_amd — MI355X-native ray-tracing hot path of rs_illume.

BVH traversal (closest and any hit), triangle/sphere intersection and the
path-tracer bounce loop as hand-written HIP kernels for gfx950 behind the C ABI
in include/illume_hip.h (librs_illume_amd.so). The Python classes mirror the
reference's traits (Primitive::intersect/intersect_p, Integrator::render) over
that ABI.
"""
from . import workloads  # noqa: F401
from ._lib import IllumeError, lib  # noqa: F401


def __getattr__(name):
    # torch-backed classes are imported lazily so that the package (and the
    # ABI loader) import without initialising a GPU.
    if name in ("BVHAccel", "rays_to_device", "hits_to_numpy", "split_hits"):
        from . import accel
        return getattr(accel, name)
    raise AttributeError(name)
