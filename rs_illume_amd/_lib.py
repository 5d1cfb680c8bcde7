"""ctypes binding of librs_illume_amd.so (the C ABI in include/illume_hip.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librs_illume_amd.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HEADER = os.path.join(INCLUDE, "illume_hip.h")

_lib = None


class IllumeError(RuntimeError):
    def __init__(self, status, fn, msg):
        super().__init__("%s failed: status %d (%s): %s" % (fn, status, status_string(status), msg))
        self.status = status


class BvhInfo(ctypes.Structure):
    _fields_ = [("n_prims", ctypes.c_uint64), ("n_inner_nodes", ctypes.c_uint64), ("n_leaves", ctypes.c_uint64),
                ("max_depth", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64), ("build_seconds", ctypes.c_double),
                ("bounds", ctypes.c_float * 6)]


def _declare(L):
    vp, u32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
    sig = {
        "illume_bvh_create": (i32, [vp, u32, vp, u32, vp, u32, i32, ctypes.POINTER(vp)]),
        "illume_bvh_create_ex": (i32, [vp, u32, vp, u32, vp, u32, i32, i32, ctypes.POINTER(vp)]),
        "illume_bvh_create_gpu": (i32, [vp, u32, vp, u32, vp, u32, i32, i32, vp, ctypes.POINTER(vp)]),
        "illume_bvh_create_gpu_sah": (i32, [vp, u32, vp, u32, vp, u32, i32, i32, vp, ctypes.POINTER(vp)]),
        "illume_bvh_destroy": (i32, [vp]),
        "illume_bvh_info": (i32, [vp, ctypes.POINTER(BvhInfo)]),
        "illume_intersect_closest": (i32, [vp, vp, u32, vp, vp]),
        "illume_intersect_any": (i32, [vp, vp, u32, vp, vp]),
        "illume_intersect_closest_noalloc": (i32, [vp, vp, u32, vp, vp]),
        "illume_intersect_counts": (i32, [vp, vp, u32, i32, vp, vp]),
        "illume_status_string": (ctypes.c_char_p, [i32]),
        "illume_last_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("librs_illume_amd.so is not built: run `python -m rs_illume_amd.build_lib` "
                               "(or __graft_entry__.build())")
        # torch's HIP runtime first: the library then binds to the runtime
        # already in the process (the same soname) instead of loading a second
        # one that does not see the devices torch has opened
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def status_string(status):
    return lib().illume_status_string(status).decode()


def check(status, fn):
    if status != 0:
        raise IllumeError(status, fn, lib().illume_last_error().decode())


def header_functions(path=None):
    """Names of the functions declared in one header, or (path None) in every
    header under include/ (the drop-in boundary illume_hip.h and the
    measurement-only illume_probe.h)."""
    paths = [path] if path else sorted(os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE) if f.endswith(".h"))
    names = set()
    for p in paths:
        text = re.sub(r"/\*.*?\*/", "", open(p).read(), flags=re.S)
        names.update(re.findall(r"\b(illume_[a-z_0-9]+)\s*\(", text))
    return sorted(names)
