// illume_guide_* entry points: device allocation of the proposal network,
// inference, gradients and Adam steps.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include "../../include/illume_hip.h"
#include "guide_kernels.h"

namespace illume {
int abi_fail(int status, const std::string &msg);
int abi_hip_fail(hipError_t e, const char *what);
}  // namespace illume

using namespace illume;

struct illume_guide {
    GuideDev dev;
    std::vector<void *> allocs;
    float *grad = nullptr, *wsum = nullptr, *loss = nullptr;
    float *partial = nullptr;  // per-block gradients, weight sums and losses of a training step
    float lr = 1e-3f;
    int step = 0;
};

namespace illume {
const GuideDev *guide_dev(const illume_guide *g) { return &g->dev; }
}  // namespace illume

namespace {
const int kDin[4] = {144, 64, 64, 64}, kDout[4] = {64, 64, 64, 40};

template <typename T>
hipError_t dalloc(illume_guide *g, T **p, size_t count) {
    hipError_t e = hipMalloc(p, count * sizeof(T));
    if (e == hipSuccess) {
        g->allocs.push_back(*p);
        e = hipMemset(*p, 0, count * sizeof(T));
    }
    return e;
}
}  // namespace

extern "C" {

int illume_guide_destroy(illume_guide *g) {
    if (!g) return ILLUME_OK;
    for (void *p : g->allocs) hipFree(p);
    delete g;
    return ILLUME_OK;
}

int illume_guide_create(const float *init_params, float lr, illume_guide **out) {
    if (!out || !init_params) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (!(lr > 0.f)) return abi_fail(ILLUME_ERR_INVALID_ARG, "learning rate must be > 0");
    // Non-finite parameters are accepted: the renderer then finds every
    // decoded mixture invalid and samples the BSDF alone at every vertex.
    illume_guide *g = new (std::nothrow) illume_guide();
    if (!g) return abi_fail(ILLUME_ERR_OUT_OF_MEMORY, "host allocation failed");
    g->lr = lr;
    GuideDev &d = g->dev;
    hipError_t e = hipSuccess;
    uint32_t off = 0;
    for (int l = 0; l < 4 && e == hipSuccess; ++l) {
        d.din[l] = kDin[l];
        d.dout[l] = kDout[l];
        d.w_off[l] = off;
        off += kDin[l] * kDout[l];
        d.b_off[l] = off;
        off += kDout[l];
    }
    if (e == hipSuccess) e = dalloc(g, &d.img, 1);
    if (e == hipSuccess) e = dalloc(g, &d.params, kGuideParams);
    if (e == hipSuccess) e = dalloc(g, &d.m, kGuideParams);
    if (e == hipSuccess) e = dalloc(g, &d.v, kGuideParams);
    if (e == hipSuccess) e = dalloc(g, &g->grad, kGuideParams);
    if (e == hipSuccess) e = dalloc(g, &g->wsum, 64);
    if (e == hipSuccess) e = dalloc(g, &g->loss, 64);
    if (e == hipSuccess) e = dalloc(g, &g->partial, kGradPartialFloats);
    if (e == hipSuccess) e = hipMemcpy(d.params, init_params, kGuideParams * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = guide_launch_sync(d, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        illume_guide_destroy(g);
        return abi_hip_fail(e, "guide allocation");
    }
    *out = g;
    return ILLUME_OK;
}

int illume_guide_get_params(const illume_guide *g, float *host_out) {
    if (!g || !host_out) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    hipError_t e = hipMemcpy(host_out, g->dev.params, kGuideParams * sizeof(float), hipMemcpyDeviceToHost);
    return e == hipSuccess ? ILLUME_OK : abi_hip_fail(e, "hipMemcpy(params)");
}

int illume_guide_infer(const illume_guide *g, const float *inputs, uint32_t n, float *out, void *stream) {
    if (!g || (n && (!inputs || !out))) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    hipError_t e = guide_launch_infer(g->dev, inputs, n, out, (hipStream_t)stream);
    return e == hipSuccess ? ILLUME_OK : abi_hip_fail(e, "guide infer launch");
}

int illume_guide_grad(const illume_guide *g, const float *records, uint32_t n, float *grad_out, void *stream) {
    if (!g || !grad_out || (n && !records)) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    hipError_t e = guide_launch_grad(g->dev, records, n, 1, grad_out, g->wsum, nullptr, g->partial, (hipStream_t)stream);
    return e == hipSuccess ? ILLUME_OK : abi_hip_fail(e, "guide grad launch");
}

}  // extern "C"

namespace illume {
// One Adam step on n records read with the given stride (in records).
int guide_train_strided(illume_guide *g, const float *records, uint32_t n, uint32_t stride, hipStream_t st) {
    if (n == 0) return ILLUME_OK;
    const hipError_t e = guide_launch_train_step(g->dev, records, n, stride, g->lr, ++g->step, g->loss, g->partial, st);
    return e == hipSuccess ? ILLUME_OK : abi_hip_fail(e, "guide train launch");
}
}  // namespace illume

extern "C" {

int illume_guide_train_step(illume_guide *g, const float *records, uint32_t n, void *stream) {
    if (!g || (n && !records)) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    if (n == 0) return ILLUME_OK;
    const hipError_t e = guide_launch_train_step(g->dev, records, n, 1, g->lr, ++g->step, g->loss, g->partial,
                                                 (hipStream_t)stream);
    return e == hipSuccess ? ILLUME_OK : abi_hip_fail(e, "guide train launch");
}

int illume_guide_stats(const illume_guide *g, uint64_t *steps, float *last_loss) {
    if (!g || !steps || !last_loss) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    *steps = (uint64_t)g->step;
    *last_loss = 0.f;
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess && g->step > 0) e = hipMemcpy(last_loss, g->loss, sizeof(float), hipMemcpyDeviceToHost);
    return e == hipSuccess ? ILLUME_OK : abi_hip_fail(e, "guide stats");
}

}  // extern "C"
