// Device state of the learned proposal and its launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace illume {

constexpr int kGuideParams = 20200;

struct GuideDev {
    float *params;  // fp32 master weights, flat [W1 b1 W2 b2 W3 b3 W4 b4], W_l as [in][out]
    float *m, *v;   // Adam moments
    __bf16 *wt[4];  // n-major copies [out_pad][in_pad] (forward B operands)
    __bf16 *wk[4];  // k-major copies [in][out_pad] (backward B operands)
    float *bias[4]; // biases, fp32, out_pad entries (zero padded)
    uint32_t w_off[4], b_off[4];
    int din[4], dout[4], wt_ld[4], wk_ld[4];
};

__device__ void write_bf16_copy(const GuideDev &g, uint32_t i, float p);

hipError_t guide_launch_infer(const GuideDev &g, const float *in, uint32_t n, float *out, hipStream_t st);
// Row count read on the device from *n_ptr (at most max_n).
hipError_t guide_launch_infer_count(const GuideDev &g, const float *in, const uint32_t *n_ptr, uint32_t max_n,
                                    float *out, hipStream_t st);
// Gradient of the weighted NLL over n records into grad (zeroed first); wsum
// (device float) receives the weight sum, loss (optional) the loss. Record r
// is read at rec + 10 * stride * r (stride >= 1 trains on an even subset).
// partial: scratch of kGradPartialFloats floats (per-block gradients).
constexpr size_t kGradPartialFloats = (size_t)256 * kGuideParams;
hipError_t guide_launch_grad(const GuideDev &g, const float *rec, uint32_t n, uint32_t stride, float *grad,
                             float *wsum, float *loss, float *partial, hipStream_t st);
hipError_t guide_launch_adam(const GuideDev &g, const float *grad, float lr, int step, hipStream_t st);
// Rewrites the bf16 copies and biases from the fp32 master weights.
hipError_t guide_launch_sync(const GuideDev &g, hipStream_t st);

}  // namespace illume
