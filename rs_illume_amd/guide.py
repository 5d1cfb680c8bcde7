"""Learned sampling proposal (not in the reference): an MLP from the encoded
hit position and normal to an 8-lobe von Mises-Fisher mixture over incident
directions, trained online on the path tracer's own radiance samples.

Parameter layout (20,200 floats, shared with the C ABI): for each layer l of
144-64-64-64-40, the weight matrix W_l as [in][out] row-major followed by the
bias b_l. Outputs per lobe j: logit (5j), unnormalised mean direction
(5j+1..5j+3), log concentration (5j+4, clamped to [-4.6, 6.9]).
"""
import ctypes

import numpy as np

DIMS = (144, 64, 64, 64, 40)
N_PARAMS = sum(DIMS[i] * DIMS[i + 1] + DIMS[i + 1] for i in range(4))
RECORD_FLOATS = 10   # normalised position (3), normal (3), direction (3), weight (1)
assert N_PARAMS == 20200


def init_params(seed=0, out_scale=0.1):
    """He-uniform hidden layers, output layer scaled by out_scale, zero biases."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for l in range(4):
        fan_in, fan_out = DIMS[l], DIMS[l + 1]
        bound = np.sqrt(6.0 / fan_in) * (out_scale if l == 3 else 1.0)
        parts.append(rng.uniform(-bound, bound, (fan_in, fan_out)).reshape(-1))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


def offsets():
    """(weight offset, bias offset) of each layer in the flat parameter vector."""
    o, out = 0, []
    for l in range(4):
        w = o
        o += DIMS[l] * DIMS[l + 1]
        out.append((w, o))
        o += DIMS[l + 1]
    return out


class Guide:
    """Device-resident network (fp32 master weights, bf16 MFMA copies, Adam
    state) behind illume_guide_* (include/illume_hip.h)."""

    def __init__(self, params=None, lr=1e-3, seed=0):
        import torch
        from . import _lib
        self._lib = _lib
        L = _lib.lib()
        _declare(L)
        p = init_params(seed) if params is None else np.ascontiguousarray(params, np.float32)
        assert p.size == N_PARAMS
        torch.cuda.init()
        h = ctypes.c_void_p()
        _lib.check(L.illume_guide_create(p.ctypes.data_as(ctypes.c_void_p), ctypes.c_float(lr), ctypes.byref(h)),
                   "illume_guide_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lib().illume_guide_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(N_PARAMS, np.float32)
        self._lib.check(self._lib.lib().illume_guide_get_params(self._h, out.ctypes.data_as(ctypes.c_void_p)),
                        "illume_guide_get_params")
        return out

    def infer(self, inputs, out=None, stream=None):
        """inputs: device float32 (n, 6) -> device float32 (n, 40)."""
        import torch
        from .accel import _stream_handle
        n = inputs.shape[0]
        if out is None:
            out = torch.empty((n, 40), dtype=torch.float32, device=inputs.device)
        self._lib.check(self._lib.lib().illume_guide_infer(self._h, ctypes.c_void_p(inputs.data_ptr()), n,
                                                           ctypes.c_void_p(out.data_ptr()), _stream_handle(stream)),
                        "illume_guide_infer")
        return out

    def grad(self, records, stream=None):
        """Gradient (device float32 (20200,)) of the weighted NLL over records (n, 10)."""
        import torch
        from .accel import _stream_handle
        g = torch.zeros(N_PARAMS, dtype=torch.float32, device=records.device)
        self._lib.check(self._lib.lib().illume_guide_grad(self._h, ctypes.c_void_p(records.data_ptr()),
                                                          records.shape[0], ctypes.c_void_p(g.data_ptr()),
                                                          _stream_handle(stream)), "illume_guide_grad")
        return g

    def stats(self):
        """(Adam steps taken, loss of the last step)."""
        steps, loss = ctypes.c_uint64(), ctypes.c_float()
        self._lib.check(self._lib.lib().illume_guide_stats(self._h, ctypes.byref(steps), ctypes.byref(loss)),
                        "illume_guide_stats")
        return int(steps.value), float(loss.value)

    def train_step(self, records, stream=None):
        """One Adam step on the batch of records (n, 10)."""
        from .accel import _stream_handle
        self._lib.check(self._lib.lib().illume_guide_train_step(self._h, ctypes.c_void_p(records.data_ptr()),
                                                                records.shape[0], _stream_handle(stream)),
                        "illume_guide_train_step")


def _declare(L):
    if getattr(L, "_guide_declared", False):
        return
    vp, u32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
    L.illume_guide_create.restype = i32
    L.illume_guide_create.argtypes = [vp, ctypes.c_float, ctypes.POINTER(vp)]
    L.illume_guide_destroy.restype = i32
    L.illume_guide_destroy.argtypes = [vp]
    L.illume_guide_get_params.restype = i32
    L.illume_guide_get_params.argtypes = [vp, vp]
    L.illume_guide_infer.restype = i32
    L.illume_guide_infer.argtypes = [vp, vp, u32, vp, vp]
    L.illume_guide_grad.restype = i32
    L.illume_guide_grad.argtypes = [vp, vp, u32, vp, vp]
    L.illume_guide_stats.restype = i32
    L.illume_guide_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_float)]
    L.illume_guide_train_step.restype = i32
    L.illume_guide_train_step.argtypes = [vp, vp, u32, vp]
    L._guide_declared = True
