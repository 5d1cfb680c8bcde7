// Device state of the learned proposal and its launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace illume {

constexpr int kGuideParams = 20200;

// The bf16 weights and fp32 biases in the layout the kernels stage in LDS, so
// that a block stages them with one contiguous copy: forward A operands
// w0..w3 ([out_pad][kGuideWaLd[l]], inputs in register order, guide_wa_pos),
// backward A operands b1..b3 (W_l as [in][72], outputs in register order,
// guide_wb_pos), biases (64 per layer, zero padded).
struct alignas(16) GuideWeightImage {
    __bf16 w0[64 * 168];
    __bf16 w1[64 * 72];
    __bf16 w2[64 * 72];
    __bf16 w3[48 * 72];
    __bf16 b1[64 * 72];
    __bf16 b2[64 * 72];
    __bf16 b3[64 * 72];
    float bias[4][64];
};
static_assert(sizeof(GuideWeightImage) % 16 == 0, "16-byte staging");

struct GuideDev {
    float *params;  // fp32 master weights, flat [W1 b1 W2 b2 W3 b3 W4 b4], W_l as [in][out]
    float *m, *v;   // Adam moments
    GuideWeightImage *img;  // bf16 copies and biases
    uint32_t w_off[4], b_off[4];
    int din[4], dout[4];
};

// Inference copies: row strides (input slots padded by 8 so that the 16 rows
// an A fragment reads fall in different LDS banks) and rows (outputs, padded
// to 16).
constexpr int kGuideWaLd[4] = {168, 72, 72, 72};
constexpr int kGuideWaRows[4] = {64, 64, 64, 48};

// Slot of input k of layer l in the inference kernel's register order.
// Layer 0 (encoding feature f = 24 i + 2 o + c, 0..143: input i, octave o,
// c = 1 for the cosine): lane group g = o / 3 computes octaves 3g..3g+2 of
// every input of its row, feature f at slot s = 12 (o % 3) + 2 i + c, which
// lands in k-block s / 8, element s % 8, i.e. k = 32 (s / 8) + 8 g + s % 8
// (slots 36..39 of each group are zero). Layers 1..3 (hidden unit
// d, the output row of the previous layer's MFMA tile m = d / 16, row
// r = d % 16): the lane that holds it is in group r / 4, element r % 4, and
// tiles 2j, 2j+1 fill k-block j, so k = 32 (m / 2) + 8 (r / 4) + 4 (m % 2) + r % 4.
__host__ __device__ inline int guide_wa_pos(int l, int k) {
    if (l == 0) {
        const int i = k / 24, o = (k % 24) / 2, c = k % 2;
        const int g = o / 3, s = 12 * (o % 3) + 2 * i + c;
        return 32 * (s / 8) + 8 * g + s % 8;
    }
    const int m = k / 16, r = k % 16;
    return 32 * (m / 2) + 8 * (r / 4) + 4 * (m % 2) + r % 4;
}

// Slot of output o of layer l (1..3) in the backward products' K order: the
// output layer's gradient is read in natural order, hidden-layer gradients
// arrive in the register order of guide_wa_pos.
__host__ __device__ inline int guide_wb_pos(int l, int o) { return l == 3 ? o : guide_wa_pos(l + 1, o); }

__device__ void write_bf16_copy(const GuideDev &g, uint32_t i, float p);

hipError_t guide_launch_infer(const GuideDev &g, const float *in, uint32_t n, float *out, hipStream_t st);
// Row count read on the device from *n_ptr (at most max_n).
hipError_t guide_launch_infer_count(const GuideDev &g, const float *in, const uint32_t *n_ptr, uint32_t max_n,
                                    float *out, hipStream_t st);
// Several networks at once: network k (of n_nets, device array nets) infers
// the counts[k * count_stride] (device) rows starting at row k * stride of
// slots and out; row r's inputs are path slot slots[r]'s in stage (position
// xyz in stage[2 p], normal xyz in stage[2 p + 1]). The outputs are stored as
// fp16 (what the renderer's shading reads).
hipError_t guide_launch_infer_nets(const GuideDev *nets, uint32_t n_nets, const float4 *stage, const uint32_t *slots,
                                   const uint32_t *counts, uint32_t count_stride, uint32_t stride, __half *out,
                                   hipStream_t st);
// Training scratch (partial): one gradient row per block of the fused pass
// (at most kTrainBlocksMax blocks), then each block's weight sum and loss.
constexpr uint32_t kTrainBlocksMax = 256;
constexpr size_t kGradPartialFloats = (size_t)kTrainBlocksMax * kGuideParams + 2 * kTrainBlocksMax;
// Gradient of the weighted NLL over n records, normalised by their total
// weight, into grad; wsum (optional, device float) receives the total weight,
// loss (optional) the normalised loss. Record r is read at rec + 10 * stride
// * r (stride >= 1 trains on an even subset).
hipError_t guide_launch_grad(const GuideDev &g, const float *rec, uint32_t n, uint32_t stride, float *grad,
                             float *wsum, float *loss, float *partial, hipStream_t st);
// One Adam step (learning rate lr, step number step >= 1) on the gradient of
// those n records; loss (optional) as above.
hipError_t guide_launch_train_step(const GuideDev &g, const float *rec, uint32_t n, uint32_t stride, float lr,
                                   int step, float *loss, float *partial, hipStream_t st);
// Rewrites the bf16 copies and biases from the fp32 master weights.
hipError_t guide_launch_sync(const GuideDev &g, hipStream_t st);

}  // namespace illume
