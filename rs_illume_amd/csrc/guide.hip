// Learned sampling proposal: 144-64-64-64-40 ReLU MLP from the frequency
// encoding of (normalised hit position, normal) to an 8-lobe vMF mixture.
// Inference and training (forward, backward, weight gradients, Adam) run on
// MFMA with bf16 weights and activations and fp32 accumulation
// (v_mfma_f32_16x16x32_bf16).
//
// Every product is computed transposed (weights as the A operand, rows of the
// batch as the B operand), with the hidden units ordered so that an output
// tile's registers are already the next product's B operand: activations and
// their gradients stay in registers through all layers. Weights are staged in
// LDS once per block. A training step is one fused pass (forward, backward and
// the weight-gradient products per 32-row tile, the weight gradients kept in
// registers per block) and one reduction that applies Adam to the fp32 master
// weights and rewrites the bf16 copies.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include "guide_kernels.h"

namespace illume {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;
constexpr int kEnc = 144, kOut = 40, kLobes = 8;
constexpr uint32_t kInferGrid = 768;   // inference blocks at most (3 per CU: 47 KB of LDS each)

// Copies BYTES (a multiple of 16) from global memory to LDS, THREADS threads,
// each issuing its loads eight at a time before the matching stores.
template <int THREADS, int BYTES>
__device__ __forceinline__ void stage_to_lds(void *dst, const void *src) {
    constexpr int N = BYTES / 16, PER = (N + THREADS - 1) / THREADS;
    const uint4 *s = reinterpret_cast<const uint4 *>(src);
    uint4 *d = reinterpret_cast<uint4 *>(dst);
#pragma unroll
    for (int k0 = 0; k0 < PER; k0 += 8) {
        uint4 r[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int i = (int)threadIdx.x + THREADS * (k0 + k);
            if (k0 + k < PER) r[k] = s[i < N ? i : N - 1];  // in bounds either way; stored only if i < N
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int i = (int)threadIdx.x + THREADS * (k0 + k);
            if (k0 + k < PER && i < N) d[i] = r[k];
        }
    }
}

__device__ __forceinline__ floatx4 mfma(bf16x8 a, bf16x8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

// Frequency encoding: input i (position xyz in [0,1], normal xyz), octave j:
// sin(2^j pi v) at feature 24 i + 2 j, cos at 24 i + 2 j + 1 (layer 0's
// 160 inputs hold 16 zeros besides). The angle is reduced exactly (2^j pi v =
// 2 pi fract(2^(j-1) v): scaling by a power of two and taking the fraction,
// v_fract_f32, are exact, except that the fraction of a tiny negative number
// is capped just below 1 rather than rounded to it) and the hardware
// v_sin/v_cos, which take their argument in revolutions, do the rest.

// Inference computes the transposed products Y^T = W^T X^T: the weights are
// the MFMA A operand (16 output units x 32 inputs) and the rows of the batch
// the B operand (32 inputs x 16 rows). An output tile then sits in the lanes
// exactly as the next layer's B operand needs it (lane = row, lane group =
// a run of 4 output units), given the input order of guide_wa_pos, so the
// activations never leave registers; the encoding is computed straight into
// B-operand registers. Each block stages all weights (47 KB, padded rows) in
// LDS once; its waves then loop over 32-row tiles (two 16-row MFMA columns
// share every A fragment) without further barriers.
constexpr int kInfWaves = 4;

struct alignas(16) InferWeights {
    __bf16 w0[64 * 168];
    __bf16 w1[64 * 72];
    __bf16 w2[64 * 72];
    __bf16 w3[48 * 72];
    float bias[4][64];
};
static_assert(sizeof(InferWeights) % 16 == 0, "16-byte staging");
// the image's forward weights are its first bytes, in InferWeights' order
constexpr int kInferWBytes = (int)offsetof(InferWeights, bias);
static_assert(offsetof(GuideWeightImage, b1) == kInferWBytes, "forward weights first in the image");

__device__ __forceinline__ bf16x8 lds_wa(const __bf16 *w, int ld, int m, int j, int lane) {
    return *reinterpret_cast<const bf16x8 *>(w + (16 * m + (lane & 15)) * ld + 32 * j + (lane >> 4) * 8);
}

// The inputs of a row come from a RowSource: rows of a float array, or the
// proposal inputs of path slots gathered through a row -> slot list.
struct RowsDirect {
    const float *__restrict__ in;
    uint32_t stride;  // floats per row; the inputs are its first six
    __device__ void load(uint32_t row, float *v) const {
#pragma unroll
        for (int i = 0; i < 6; ++i) v[i] = in[(size_t)row * stride + i];
    }
};
struct RowsGather {
    const uint32_t *__restrict__ slots;  // row -> path slot
    const float4 *__restrict__ stage;    // [2 p]: position xyz, [2 p + 1]: normal xyz
    __device__ void load(uint32_t row, float *v) const {
        const uint32_t p = slots[row];
        const float4 a = stage[2 * (size_t)p], c = stage[2 * (size_t)p + 1];
        v[0] = a.x;
        v[1] = a.y;
        v[2] = a.z;
        v[3] = c.x;
        v[4] = c.y;
        v[5] = c.z;
    }
};

// Encoding of one row into layer 0's B operand: lane group g computes
// octaves 3g .. 3g + 2 of all six inputs, octave 3g + d of input i at slot
// s = 12 d + 2 i (sine) and s + 1 (cosine); slots 36..39 are zero. Octave 3g
// comes from the hardware sin / cos of the exactly reduced angle, the next two
// from the double-angle formulas sin 2a = 2 sin a cos a and cos 2a =
// cos^2 a - sin^2 a (one fused multiply-add): their error, a few float ulps,
// is far below the bf16 rounding of the operand, and a row costs 12
// transcendentals per lane instead of 40. Rows past n (n >= 1) encode row
// n - 1's inputs: inference stores no output for them and training gives them
// weight 0, so their gradient is zero. encode_v encodes a row from its six
// inputs v, encode_b loads them first.
__device__ __forceinline__ void encode_v(const float *v, int g, bf16x8 *b) {
    float e[40];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float x = v[i];
        const float r = ldexpf(x, 3 * g - 1);
        const float rev = __builtin_amdgcn_fractf(r);
        float sv = __builtin_amdgcn_sinf(rev), cv = __builtin_amdgcn_cosf(rev);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (d > 0) {
                const float p = sv * cv;
                cv = fmaf(cv, cv, -(sv * sv));
                sv = p + p;
            }
            e[12 * d + 2 * i] = sv;
            e[12 * d + 2 * i + 1] = cv;
        }
    }
#pragma unroll
    for (int q = 36; q < 40; ++q) e[q] = 0.f;
#pragma unroll
    for (int q = 0; q < 40; ++q) b[q / 8][q % 8] = to_bf16(e[q]);
}

template <typename Src>
__device__ __forceinline__ void encode_b(const Src &src, uint32_t row, uint32_t n, int g, bf16x8 *b) {
    float v[6];
    src.load(row < n ? row : n - 1, v);
    encode_v(v, g, b);
}

// acc (output tiles of one 16-row column, bias included) -> ReLU, bf16 -> the
// next layer's B operand (tile m fills k-block m / 2, elements 4 (m % 2) + e).
// The ReLU acts on the bf16 values as 16-bit integers (v_pk_max_i16 against
// zero, two values per instruction): rounding keeps the sign, so it equals the
// ReLU before rounding except that a NaN with its sign bit clear passes through.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short short2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <int MT>
__device__ __forceinline__ void relu_to_b(const floatx4 *acc, bf16x8 *b) {
    static_assert(MT % 2 == 0, "whole k-blocks");
#pragma unroll
    for (int j = 0; j < MT / 2; ++j) {
        u32x4 w;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            // dword q of k-block j: tile 2 j + q / 2, elements 2 (q % 2) and the next
            const int m = 2 * j + q / 2, e = 2 * (q % 2);
            const bf16x2 p = __builtin_convertvector((float2v){acc[m][e], acc[m][e + 1]}, bf16x2);
            const short2v r = __builtin_elementwise_max(__builtin_bit_cast(short2v, p), (short2v)0);
            w[q] = __builtin_bit_cast(unsigned, r);
        }
        b[j] = __builtin_bit_cast(bf16x8, w);
    }
}

// acc[nn][m] = W^T tile m (16 outputs) x the two 16-row columns nn of b, the
// accumulation started from the layer's bias (BIAS: bias[16 m + 4 g + e] is
// this lane's element e of tile m) or from zero.
template <int KB, int MT, bool BIAS = false>
__device__ __forceinline__ void layer_t(const __bf16 *w, int ld, const bf16x8 (*b)[KB], floatx4 (*acc)[4], int lane,
                                        const float *bias = nullptr) {
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        const floatx4 c0 = BIAS ? *reinterpret_cast<const floatx4 *>(bias + 16 * m + 4 * (lane >> 4))
                                : floatx4{0.f, 0.f, 0.f, 0.f};
        acc[0][m] = c0;
        acc[1][m] = c0;
#pragma unroll
        for (int j = 0; j < KB; ++j) {
            const bf16x8 a = lds_wa(w, ld, m, j, lane);
            acc[0][m] = mfma(a, b[0][j], acc[0][m]);
            acc[1][m] = mfma(a, b[1][j], acc[1][m]);
        }
    }
}

// Four consecutive outputs of a row.
__device__ __forceinline__ void put4(float *o, float4 y) { *reinterpret_cast<float4 *>(o) = y; }
__device__ __forceinline__ void put4(__half *o, float4 y) {
    const uint32_t a = (uint32_t)__half_as_ushort(__float2half(y.x)) | (uint32_t)__half_as_ushort(__float2half(y.y)) << 16;
    const uint32_t b = (uint32_t)__half_as_ushort(__float2half(y.z)) | (uint32_t)__half_as_ushort(__float2half(y.w)) << 16;
    *reinterpret_cast<uint2 *>(o) = make_uint2(a, b);
}

// n rows of src -> out [n][40] (float or fp16) by the blocks blockIdx.x of
// one grid row; n from n_ptr (at most max_n) when n_ptr is non-null.
template <typename OutT, typename Src>
__device__ __forceinline__ void infer_rows(const GuideDev &g, const Src &src, const uint32_t *__restrict__ n_ptr,
                                           uint32_t n_arg, OutT *__restrict__ out) {
    __shared__ InferWeights W;
    const uint32_t n = n_ptr ? min(*n_ptr, n_arg) : n_arg;
    const uint32_t n_iter = (n + 31) / 32;
    if (blockIdx.x * kInfWaves >= n_iter) return;  // the whole block has no rows
    stage_to_lds<kInfWaves * kWave, kInferWBytes>(W.w0, g.img->w0);
    stage_to_lds<kInfWaves * kWave, (int)sizeof(W.bias)>(W.bias, g.img->bias);
    __syncthreads();
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave, grp = lane >> 4;
    for (uint32_t it = blockIdx.x * kInfWaves + wave; it < n_iter; it += gridDim.x * kInfWaves) {
        // Re-read the weight fragments from LDS in every iteration: kept in
        // registers across iterations they would take 168 VGPRs.
        __asm__ volatile("" ::: "memory");
        const uint32_t base = it * 32;
        bf16x8 b0[2][5], b1[2][2];
        floatx4 acc[2][4];
        encode_b(src, base + (lane & 15), n, grp, b0[0]);
        encode_b(src, base + 16 + (lane & 15), n, grp, b0[1]);
        layer_t<5, 4, true>(W.w0, 168, b0, acc, lane, W.bias[0]);
        relu_to_b<4>(acc[0], b1[0]);
        relu_to_b<4>(acc[1], b1[1]);
        layer_t<2, 4, true>(W.w1, 72, b1, acc, lane, W.bias[1]);
        relu_to_b<4>(acc[0], b1[0]);
        relu_to_b<4>(acc[1], b1[1]);
        layer_t<2, 4, true>(W.w2, 72, b1, acc, lane, W.bias[2]);
        relu_to_b<4>(acc[0], b1[0]);
        relu_to_b<4>(acc[1], b1[1]);
        layer_t<2, 3, true>(W.w3, 72, b1, acc, lane, W.bias[3]);
#pragma unroll
        for (int nn = 0; nn < 2; ++nn) {
            const uint32_t row = base + 16 * nn + (lane & 15);
            if (row >= n) continue;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const int o = 16 * m + 4 * grp;
                if (o >= kOut) continue;
                const float4 y = make_float4(acc[nn][m][0], acc[nn][m][1], acc[nn][m][2], acc[nn][m][3]);
                put4(out + (size_t)row * kOut + o, y);
            }
        }
    }
}

__global__ __launch_bounds__(kInfWaves *kWave) void guide_infer_kernel(GuideDev g, const float *__restrict__ in,
                                                                       const uint32_t *__restrict__ n_ptr,
                                                                       uint32_t n_arg, float *__restrict__ out) {
    infer_rows(g, RowsDirect{in, 6}, n_ptr, n_arg, out);
}

// One grid row per network k (blockIdx.y): its rows start at row k * stride
// of slots and out, their count is counts[k]; row r's inputs are those of
// path slot slots[r] in stage.
__global__ __launch_bounds__(kInfWaves *kWave) void guide_infer_nets_kernel(const GuideDev *__restrict__ nets,
                                                                            const float4 *__restrict__ stage,
                                                                            const uint32_t *__restrict__ slots,
                                                                            const uint32_t *__restrict__ counts,
                                                                            uint32_t count_stride, uint32_t stride,
                                                                            __half *__restrict__ out) {
    const uint32_t k = blockIdx.y;
    const GuideDev g = nets[k];
    infer_rows(g, RowsGather{slots + (size_t)k * stride, stage}, counts + (size_t)k * count_stride, stride,
               out + (size_t)kOut * k * stride);
}

// d(-log q)/d outputs for one row (mixture decode as in the sampler), on the
// hardware exp2 / log2 / reciprocal / square root: the loss's gradient needs
// no more than float rounding, and the oracle's gradient agrees within its
// bf16 tolerance.
__device__ __forceinline__ float hexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896340736f); }
__device__ __forceinline__ float hrcp(float x) { return __builtin_amdgcn_rcpf(x); }
// 1 - exp(-2k) for k >= 0, with the series where the difference cancels
__device__ __forceinline__ float hom2k(float k) { return k < 1e-3f ? 2.f * k * (1.f - k) : 1.f - hexp(-2.f * k); }
__device__ void nll_grad_row(const float *y, const float *w, float scale, float *gy, float *loss) {
    float mx = y[0];
    for (int j = 1; j < kLobes; ++j) mx = fmaxf(mx, y[5 * j]);
    float pi[kLobes], s = 0.f;
    for (int j = 0; j < kLobes; ++j) {
        pi[j] = hexp(y[5 * j] - mx);
        s += pi[j];
    }
    const float inv_s = hrcp(s);
    float v[kLobes], mu[kLobes][3], kap[kLobes], imn[kLobes], q = 0.f;
    bool clamped[kLobes];
    for (int j = 0; j < kLobes; ++j) {
        pi[j] *= inv_s;
        const float a = y[5 * j + 1], b = y[5 * j + 2], c = y[5 * j + 3];
        const float len2 = a * a + b * b + c * c;
        const float inv = len2 > 0.f ? __builtin_amdgcn_rsqf(len2) : 0.f;
        imn[j] = inv;
        mu[j][0] = a * inv;
        mu[j][1] = b * inv;
        mu[j][2] = len2 > 0.f ? c * inv : 1.f;
        const float r = y[5 * j + 4];
        clamped[j] = r < -4.6f || r > 6.9f;
        kap[j] = hexp(fminf(fmaxf(r, -4.6f), 6.9f));
        const float cdot = mu[j][0] * w[0] + mu[j][1] * w[1] + mu[j][2] * w[2];
        v[j] = kap[j] * hrcp(6.28318530717958647692f * hom2k(kap[j])) * hexp(kap[j] * (cdot - 1.f));
        q += pi[j] * v[j];
    }
    *loss = -scale * __builtin_amdgcn_logf(q) * 0.69314718055994530942f;
    const float iq = hrcp(q);
    for (int j = 0; j < kLobes; ++j) {
        const float gam = pi[j] * v[j] * iq;
        gy[5 * j] = scale * (pi[j] - gam);
        const float k = kap[j];
        const float cdot = mu[j][0] * w[0] + mu[j][1] * w[1] + mu[j][2] * w[2];
        const float dlk = hrcp(k) - 2.f * hexp(-2.f * k) * hrcp(hom2k(k)) + (cdot - 1.f);
        gy[5 * j + 4] = clamped[j] ? 0.f : scale * (-gam * dlk * k);
        for (int a = 0; a < 3; ++a) gy[5 * j + 1 + a] = scale * (-gam * k * (w[a] - cdot * mu[j][a]) * imn[j]);
    }
}

// The same gradient with a row split over two lanes of a wave: lane r and lane
// r + 32 take lobes 4 h .. 4 h + 3 (h = lane / 32) of the row whose outputs
// are y, combine the mixture's maximum, normaliser and density with one
// exchange each, and overwrite their lobes' outputs in y with the gradient
// (each lane reads and writes only its own lobes). Returns the row's loss
// (both lanes). Every lane of the wave calls it.
__device__ __forceinline__ float nll_grad_half(float *y, const float *w, float scale, int lane) {
    const int h = lane >> 5;
    float *yh = y + 20 * h;
    float mx = yh[0];
#pragma unroll
    for (int j = 1; j < 4; ++j) mx = fmaxf(mx, yh[5 * j]);
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    float pi[4], s = 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        pi[j] = hexp(yh[5 * j] - mx);
        s += pi[j];
    }
    s += __shfl_xor(s, 32);
    const float inv_s = hrcp(s);
    float v[4], mu[4][3], kap[4], imn[4], q = 0.f;
    bool clamped[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        pi[j] *= inv_s;
        const float a = yh[5 * j + 1], b = yh[5 * j + 2], c = yh[5 * j + 3];
        const float len2 = a * a + b * b + c * c;
        const float inv = len2 > 0.f ? __builtin_amdgcn_rsqf(len2) : 0.f;
        imn[j] = inv;
        mu[j][0] = a * inv;
        mu[j][1] = b * inv;
        mu[j][2] = len2 > 0.f ? c * inv : 1.f;
        const float r = yh[5 * j + 4];
        clamped[j] = r < -4.6f || r > 6.9f;
        kap[j] = hexp(fminf(fmaxf(r, -4.6f), 6.9f));
        const float cdot = mu[j][0] * w[0] + mu[j][1] * w[1] + mu[j][2] * w[2];
        v[j] = kap[j] * hrcp(6.28318530717958647692f * hom2k(kap[j])) * hexp(kap[j] * (cdot - 1.f));
        q += pi[j] * v[j];
    }
    q += __shfl_xor(q, 32);
    const float iq = hrcp(q);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float gam = pi[j] * v[j] * iq;
        const float k = kap[j];
        const float cdot = mu[j][0] * w[0] + mu[j][1] * w[1] + mu[j][2] * w[2];
        const float dlk = hrcp(k) - 2.f * hexp(-2.f * k) * hrcp(hom2k(k)) + (cdot - 1.f);
        const bool live = scale > 0.f;  // a row of weight 0 contributes exactly nothing
        yh[5 * j] = live ? scale * (pi[j] - gam) : 0.f;
        yh[5 * j + 4] = (clamped[j] || !live) ? 0.f : scale * (-gam * dlk * k);
#pragma unroll
        for (int a = 0; a < 3; ++a)
            yh[5 * j + 1 + a] = live ? scale * (-gam * k * (w[a] - cdot * mu[j][a]) * imn[j]) : 0.f;
    }
    return -scale * __builtin_amdgcn_logf(q) * 0.69314718055994530942f;
}

// ------------------------------------------------------------------ training
// One fused pass per Adam step (guide_train_kernel), then one reduction that
// also applies Adam (guide_reduce_adam_kernel).
//
// A block of kTrainWaves waves takes a contiguous run of 32-row tiles, one
// tile per wave per round. Each wave runs its tile's forward pass as in
// inference (activations in registers), the NLL gradient of its rows (lanes
// 0..31, one row each, through the wave's LDS) and the backward products
// dX^T = W G^T with the weights as A operands (wb copies) so that every
// gradient lands in the registers the next backward product reads. The
// activations X_1..X_3 and the pre-activation gradients G_3..G_0 are staged
// transposed ([unit][32 rows] bf16) in the wave's LDS, one G at a time; after
// each G the block synchronises and every wave adds the weight-gradient
// products dW_l = X_l^T G_l (MFMA with K = the 32 rows of a tile) of all the
// round's tiles into the dW tiles it owns, which stay in registers for the
// whole run (X_0, the encoding, is recomputed from the rows' staged inputs).
// Nothing per tile goes to HBM: a block writes one partial gradient row, its
// weight sum and its loss; the reduction adds the rows, normalises by the
// total weight and applies Adam to the fp32 master weights and their bf16
// copies.
constexpr int kTrainWaves = 8;  // two per SIMD
constexpr int kTrainThreads = kTrainWaves * kWave;

// Staged tensors are [unit][row], kStageLd rows per unit.
constexpr int kStageLd = 32;
struct alignas(16) TrainWaveLds {
    __bf16 x[64 * kStageLd];     // the wave's X_l while dW_l is formed (X_3, then X_2, then X_1)
    union {
        float y[32][48];           // outputs of the tile, then their gradients
        __bf16 g[64 * kStageLd];   // the current G_l
    } yg;
    float in[32][6];               // the rows' network inputs (zero past n)
};
struct TrainLds {
    GuideWeightImage w;
    TrainWaveLds wave[kTrainWaves];
};
static_assert(sizeof(TrainLds) <= 160 * 1024, "training LDS within one CU's 160 KB");
static_assert(kGradPartialFloats >= (size_t)kTrainBlocksMax * kGuideParams + 2 * kTrainBlocksMax,
              "one partial row per block and its weight sum and loss");

// LDS written by some lanes of a wave and read by others of the same wave.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ReLU pattern of a hidden layer for this lane's row of column nn: bit
// 16 nn + 4 m + e set when unit 16 m + 4 g + e is positive.
__device__ __forceinline__ uint32_t relu_bits(const bf16x8 *h, int nn) {
    uint32_t bits = 0;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if ((float)h[m / 2][4 * (m % 2) + e] > 0.f) bits |= 1u << (16 * nn + 4 * m + e);
    return bits;
}

// Gradient of a hidden layer's output (acc, units 16 m + 4 g + e of this
// lane's row in column nn) masked by that layer's ReLU -> the next backward
// product's B operand.
__device__ __forceinline__ void mask_to_b(const floatx4 *acc, uint32_t bits, int nn, bf16x8 *b) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float v = ((bits >> (16 * nn + 4 * m + e)) & 1u) ? acc[m][e] : 0.f;
            b[m / 2][4 * (m % 2) + e] = to_bf16(v);
        }
}

// A 64-unit tensor held in register order (v[nn] for the two 16-row columns)
// -> dst [unit][row] in LDS.
__device__ __forceinline__ void stage_units(const bf16x8 (*v)[2], __bf16 *dst, int lane) {
    const int g = lane >> 4, r = lane & 15;
#pragma unroll
    for (int nn = 0; nn < 2; ++nn)
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int e = 0; e < 4; ++e) dst[(16 * m + 4 * g + e) * kStageLd + 16 * nn + r] = v[nn][m / 2][4 * (m % 2) + e];
}

// The lane's element offset in a staged fragment, (r16 row, 8 grp column):
// an opaque value per call, so that the fragments' addresses are formed where
// they are read instead of being kept live (and spilled) across the round.
__device__ __forceinline__ int frag_off(int lane) {
    int o = (lane & 15) * kStageLd + 8 * (lane >> 4);
    __asm__ volatile("" : "+v"(o));
    return o;
}

// dW tile (16 inputs it x 16 outputs u) of layer l >= 1 += X_l^T G_l over the
// 32 rows of the round's tiles; colsum += the rows' G_l of output column u
// (the bias gradient) when want_bias.
__device__ __forceinline__ floatx4 dw_tile(const TrainLds &L, int it, int u, floatx4 acc, float &colsum,
                                           bool want_bias, int lane) {
    const int fo = frag_off(lane);
#pragma unroll
    for (int ti = 0; ti < kTrainWaves; ++ti) {
        const TrainWaveLds &T = L.wave[ti];
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(T.x + 16 * it * kStageLd + fo);
        const bf16x8 b = *reinterpret_cast<const bf16x8 *>(T.yg.g + 16 * u * kStageLd + fo);
        acc = mfma(a, b, acc);
        if (want_bias)
#pragma unroll
            for (int e = 0; e < 8; ++e) colsum += (float)b[e];
    }
    return acc;
}
// Two dW tiles of one input tile (it, u0) and (it, u1): each A fragment read
// once for both products.
__device__ __forceinline__ void dw_tile2(const TrainLds &L, int it, int u0, int u1, floatx4 &acc0, floatx4 &acc1,
                                         float &colsum0, float &colsum1, bool want_bias, int lane) {
    const int fo = frag_off(lane);
#pragma unroll
    for (int ti = 0; ti < kTrainWaves; ++ti) {
        const TrainWaveLds &T = L.wave[ti];
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(T.x + 16 * it * kStageLd + fo);
        const bf16x8 b0 = *reinterpret_cast<const bf16x8 *>(T.yg.g + 16 * u0 * kStageLd + fo);
        const bf16x8 b1 = *reinterpret_cast<const bf16x8 *>(T.yg.g + 16 * u1 * kStageLd + fo);
        acc0 = mfma(a, b0, acc0);
        acc1 = mfma(a, b1, acc1);
        if (want_bias)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                colsum0 += (float)b0[e];
                colsum1 += (float)b1[e];
            }
    }
}

// Layer 0's A operand: encoding feature 16 it + r16 of rows 8 grp .. 8 grp + 7
// of a tile, from its staged inputs tin[32][6]. A cosine is the sine a quarter
// revolution on (one transcendental per element, not both and a select).
__device__ __forceinline__ bf16x8 enc_a(const float (*tin)[6], int it, int lane) {
    const int grp = lane >> 4, r16 = lane & 15;
    const int f = 16 * it + r16, i = f / 24, o = (f % 24) / 2;
    const float quarter = (f & 1) ? 0.25f : 0.f;
    bf16x8 a;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float r = ldexpf(tin[8 * grp + e][i], o - 1);
        const float rev = __builtin_amdgcn_fractf(r);
        a[e] = to_bf16(__builtin_amdgcn_sinf(rev + quarter));
    }
    return a;
}

// Writes a dW tile (and, with want_bias, the column sums as bias gradients)
// of layer l to the partial row.
__device__ __forceinline__ void put_dw(const GuideDev &g, float *prow, int l, int it, int u, floatx4 acc,
                                       float colsum, bool want_bias, int lane) {
    const int grp = lane >> 4, r16 = lane & 15;
    const int din = g.din[l], dout = g.dout[l], out = 16 * u + r16;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int in = 16 * it + 4 * grp + e;
        if (in < din && out < dout) prow[g.w_off[l] + in * dout + out] = acc[e];
    }
    if (want_bias) {
        colsum += __shfl_xor(colsum, 16);
        colsum += __shfl_xor(colsum, 32);
        if (grp == 0 && out < dout) prow[g.b_off[l] + out] = colsum;
    }
}

// A round's record data for one lane of a wave's 32-row tile at row base:
// the sample weights of row base + (lane % 32) (zero past n), the inputs of
// row base + lane for the staged copy (lanes < 32; zero past n), and the
// encoding's inputs of rows base + r16 and base + 16 + r16 (past n: row n - 1).
struct RoundIn {
    float w[4];
    float in[6];
    float v[2][6];
};
__device__ __forceinline__ void load_round(const float *__restrict__ rec, uint32_t n, uint32_t stride_f, uint32_t base,
                                           int lane, RoundIn &ri) {
    const uint32_t rw = base + (uint32_t)(lane & 31);
#pragma unroll
    for (int c = 0; c < 4; ++c) ri.w[c] = rw < n ? rec[(size_t)rw * stride_f + 6 + c] : 0.f;
    const uint32_t ri_row = base + (uint32_t)lane;
#pragma unroll
    for (int i = 0; i < 6; ++i) ri.in[i] = (lane < 32 && ri_row < n) ? rec[(size_t)ri_row * stride_f + i] : 0.f;
#pragma unroll
    for (int nn = 0; nn < 2; ++nn) {
        const uint32_t row = base + 16 * nn + (uint32_t)(lane & 15);
        RowsDirect{rec, stride_f}.load(row < n ? row : n - 1, ri.v[nn]);
    }
}

// Block b trains on tiles [b * per_block, (b + 1) * per_block) of the n
// records (record r at rec + stride_f * r floats); partial row b gets its
// gradient (rows weighted by w, not yet normalised), wl[2 b], wl[2 b + 1] its
// weight sum and loss. dW tiles of wave w (11, 44 accumulator registers, so
// that two waves share a SIMD): layers 1..3, input tile w % 4 and output tiles
// w / 4 and w / 4 + 2 (< 3 for layer 3); layer 0, input tile w with every
// output tile, and input tile 8 (features 128..143; tile 9 is padding) with
// output tile w % 4 over the rows of waves 4 (w / 4) .. 4 (w / 4) + 3, the two
// halves added at the end. The owners of input tile 0 sum the bias gradients.
__global__ __launch_bounds__(kTrainThreads, 1) void guide_train_kernel(GuideDev g, const float *__restrict__ rec,
                                                                      uint32_t n, uint32_t stride_f,
                                                                      uint32_t per_block, float *__restrict__ partial,
                                                                      float *__restrict__ wl) {
    __shared__ TrainLds L;
    __shared__ float red[2][kTrainWaves];
    const uint32_t t_begin = blockIdx.x * per_block, t_end = t_begin + per_block;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);  // uniform: scalar registers
    // the first round's record data, loaded before the weights are staged so
    // that its latency overlaps the staging
    RoundIn first;
    load_round(rec, n, stride_f, (t_begin + wave) * 32, (int)threadIdx.x % kWave, first);
    stage_to_lds<kTrainThreads, (int)sizeof(GuideWeightImage)>(&L.w, g.img);
    __syncthreads();
    TrainWaveLds &S = L.wave[wave];
    const int itw = wave % 4, uo = wave / 4;  // layers 1..3: input tile, first output tile
    const bool bias_hi = itw == 0;            // sums the bias gradients of its output tiles (layers 1..3)
    const bool bias_0 = wave == 0;            // ... and of layer 0
    floatx4 d3[2], d2[2], d1[2], d0[4], d8;
    float c3[2], c2[2], c1[2], c0[4];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        d3[j] = d2[j] = d1[j] = floatx4{0.f, 0.f, 0.f, 0.f};
        c3[j] = c2[j] = c1[j] = 0.f;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        d0[u] = floatx4{0.f, 0.f, 0.f, 0.f};
        c0[u] = 0.f;
    }
    d8 = floatx4{0.f, 0.f, 0.f, 0.f};
    float wave_loss = 0.f, wave_wsum = 0.f;
    for (uint32_t t0 = t_begin; t0 < t_end; t0 += kTrainWaves) {
        // lane-derived values (the encoding's feature choices, fragment
        // addresses) are formed inside the round: hoisted out of it they
        // would be kept live through the whole round and spilled
        int lane = (int)threadIdx.x % kWave;
        __asm__ volatile("" : "+v"(lane));
        const int grp = lane >> 4, r16 = lane & 15;
        const uint32_t base = (t0 + wave) * 32;
        RoundIn ri;
        if (t0 == t_begin)
            ri = first;
        else
            load_round(rec, n, stride_f, base, lane, ri);
        // the rows' sampled directions and weights, needed after the forward
        // pass (lanes r and r + 32: row r)
        const float *w_row = ri.w;
        // ---- forward (rows past n repeat row n - 1 with weight 0: their gradient is zero)
        if (lane < 32)
#pragma unroll
            for (int i = 0; i < 6; ++i) S.in[lane][i] = ri.in[i];
        uint32_t mask1, mask2, mask3;
        bf16x8 h1[2][2], h2[2][2];  // X_1, X_2 until their weight gradients
        {
            bf16x8 x0[2][5], h[2][2];
            floatx4 acc[2][4];
            encode_v(ri.v[0], grp, x0[0]);
            encode_v(ri.v[1], grp, x0[1]);
            layer_t<5, 4, true>(L.w.w0, 168, x0, acc, lane, L.w.bias[0]);
            relu_to_b<4>(acc[0], h1[0]);
            relu_to_b<4>(acc[1], h1[1]);
            mask1 = relu_bits(h1[0], 0) | relu_bits(h1[1], 1);
            layer_t<2, 4, true>(L.w.w1, 72, h1, acc, lane, L.w.bias[1]);
            relu_to_b<4>(acc[0], h2[0]);
            relu_to_b<4>(acc[1], h2[1]);
            mask2 = relu_bits(h2[0], 0) | relu_bits(h2[1], 1);
            layer_t<2, 4, true>(L.w.w2, 72, h2, acc, lane, L.w.bias[2]);
            relu_to_b<4>(acc[0], h[0]);
            relu_to_b<4>(acc[1], h[1]);
            mask3 = relu_bits(h[0], 0) | relu_bits(h[1], 1);
            stage_units(h, S.x, lane);  // X_3
            layer_t<2, 3, true>(L.w.w3, 72, h, acc, lane, L.w.bias[3]);
#pragma unroll
            for (int nn = 0; nn < 2; ++nn)
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const int o = 16 * m + 4 * grp;
                    *reinterpret_cast<float4 *>(&S.yg.y[16 * nn + r16][o]) =
                        make_float4(acc[nn][m][0], acc[nn][m][1], acc[nn][m][2], acc[nn][m][3]);
                }
        }
        wave_lds_sync();
        // ---- output gradient: lanes r and r + 32 share row base + r (a row of
        // weight 0, or past n, gets a zero gradient)
        {
            float *yr = S.yg.y[lane & 31];
            const float wgt = w_row[3];
            const float row_loss = nll_grad_half(yr, w_row, wgt > 0.f ? wgt : 0.f, lane);
            if (lane < 32 && wgt > 0.f) {
                wave_wsum += wgt;
                wave_loss += row_loss;
            }
            if (lane >= 32)
#pragma unroll
                for (int o = kOut; o < 48; ++o) yr[o] = 0.f;
        }
        wave_lds_sync();
        // G_3 as the first backward product's B operand (natural output order,
        // K padded to 64) and transposed for the weight gradients
        bf16x8 g3[2][2], gz[2][2], gt[4];
        floatx4 acc[2][4];
        // (the eight outputs of a k-block read as two 16-byte loads; outputs
        // 48..63 are the padding)
#pragma unroll
        for (int nn = 0; nn < 2; ++nn)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int k0 = 32 * j + 8 * grp;
                float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
                if (k0 < 48) {
                    const float4 *src = reinterpret_cast<const float4 *>(&S.yg.y[16 * nn + r16][k0]);
                    a = src[0];
                    b = src[1];
                }
                g3[nn][j][0] = to_bf16(a.x);
                g3[nn][j][1] = to_bf16(a.y);
                g3[nn][j][2] = to_bf16(a.z);
                g3[nn][j][3] = to_bf16(a.w);
                g3[nn][j][4] = to_bf16(b.x);
                g3[nn][j][5] = to_bf16(b.y);
                g3[nn][j][6] = to_bf16(b.z);
                g3[nn][j][7] = to_bf16(b.w);
            }
        // G_3^T: lane (part p, unit u) gathers rows 8 p .. 8 p + 7 of output u;
        // a 16-lane group takes one p and 16 consecutive units, so its reads
        // fall in 16 consecutive banks
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int unit = (lane & 15) + 16 * q, part = (lane >> 4) & 3;
#pragma unroll
            for (int e = 0; e < 8; ++e) gt[q][e] = to_bf16(unit < 48 ? S.yg.y[part * 8 + e][unit] : 0.f);
        }
        wave_lds_sync();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int unit = (lane & 15) + 16 * q, part = (lane >> 4) & 3;
            *reinterpret_cast<bf16x8 *>(S.yg.g + unit * kStageLd + part * 8) = gt[q];
        }
        // Each G_{l-1} = (W_l G_l^T) * relu'(h_l) needs only this wave's G_l,
        // so it is formed before the barrier that opens dW_l and its products
        // run while the other waves arrive.
        // ---- G_2 (registers), dW_3
        layer_t<2, 4>(L.w.b3, 72, g3, acc, lane);
        mask_to_b(acc[0], mask3, 0, gz[0]);
        mask_to_b(acc[1], mask3, 1, gz[1]);
        __syncthreads();
        if (uo == 0)
            dw_tile2(L, itw, 0, 2, d3[0], d3[1], c3[0], c3[1], bias_hi, lane);
        else
            d3[0] = dw_tile(L, itw, 1, d3[0], c3[0], bias_hi, lane);
        __syncthreads();
        // ---- G_2 and X_2 staged, G_1 (registers), dW_2
        stage_units(gz, S.yg.g, lane);
        stage_units(h2, S.x, lane);
        layer_t<2, 4>(L.w.b2, 72, gz, acc, lane);
        mask_to_b(acc[0], mask2, 0, gz[0]);
        mask_to_b(acc[1], mask2, 1, gz[1]);
        __syncthreads();
        dw_tile2(L, itw, uo, uo + 2, d2[0], d2[1], c2[0], c2[1], bias_hi, lane);
        __syncthreads();
        // ---- G_1 and X_1 staged, G_0 (registers), dW_1
        stage_units(gz, S.yg.g, lane);
        stage_units(h1, S.x, lane);
        layer_t<2, 4>(L.w.b1, 72, gz, acc, lane);
        mask_to_b(acc[0], mask1, 0, gz[0]);
        mask_to_b(acc[1], mask1, 1, gz[1]);
        __syncthreads();
        dw_tile2(L, itw, uo, uo + 2, d1[0], d1[1], c1[0], c1[1], bias_hi, lane);
        __syncthreads();
        // ---- G_0 staged, dW_0 (A = the encoding of the round's rows)
        stage_units(gz, S.yg.g, lane);
        __syncthreads();
        const int fo = frag_off(lane);
#pragma unroll
        for (int ti = 0; ti < kTrainWaves; ++ti) {
            const TrainWaveLds &T = L.wave[ti];
            bf16x8 b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                b[u] = *reinterpret_cast<const bf16x8 *>(T.yg.g + 16 * u * kStageLd + fo);
                if (bias_0)
#pragma unroll
                    for (int e = 0; e < 8; ++e) c0[u] += (float)b[u][e];
            }
            const bf16x8 a = enc_a(T.in, wave, lane);
#pragma unroll
            for (int u = 0; u < 4; ++u) d0[u] = mfma(a, b[u], d0[u]);
            if (ti / 4 == uo) {
                const bf16x8 a8 = enc_a(T.in, 8, lane);
                const int u8 = itw;
                d8 = mfma(a8, u8 == 0 ? b[0] : u8 == 1 ? b[1] : u8 == 2 ? b[2] : b[3], d8);
            }
        }
        __syncthreads();
    }
    // ---- the block's partial gradient row, weight sum and loss
    const int lane = (int)threadIdx.x % kWave;
    float *prow = partial + (size_t)blockIdx.x * kGuideParams;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int u = uo + 2 * j;
        if (u < 3) put_dw(g, prow, 3, itw, u, d3[j], c3[j], bias_hi, lane);
        put_dw(g, prow, 2, itw, u, d2[j], c2[j], bias_hi, lane);
        put_dw(g, prow, 1, itw, u, d1[j], c1[j], bias_hi, lane);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) put_dw(g, prow, 0, wave, u, d0[u], c0[u], bias_0, lane);
    // input tile 8 of layer 0: waves 4..7 hand their halves to waves 0..3
    floatx4 *half8 = reinterpret_cast<floatx4 *>(S.x);
    if (uo == 1) half8[lane] = d8;
    __syncthreads();
    if (uo == 0) {
        const floatx4 o = reinterpret_cast<const floatx4 *>(L.wave[wave + 4].x)[lane];
        put_dw(g, prow, 0, 8, itw, d8 + o, 0.f, false, lane);
    }
    float l = wave_loss, w = wave_wsum;
    for (int off = 32; off > 0; off >>= 1) {
        l += __shfl_xor(l, off);
        w += __shfl_xor(w, off);
    }
    if (lane == 0) {
        red[0][wave] = w;
        red[1][wave] = l;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float ws = 0.f, ls = 0.f;
        for (int k = 0; k < kTrainWaves; ++k) {
            ws += red[0][k];
            ls += red[1][k];
        }
        wl[2 * blockIdx.x] = ws;
        wl[2 * blockIdx.x + 1] = ls;
    }
}

// Parameter i's gradient: the sum of the blocks' partial rows over the total
// weight. A block takes 64 consecutive parameters as 16 float4 columns; its
// thread (column c, row group r) loads rows r, r + 16, ... of column c (all
// loads in flight at once, 256 contiguous bytes of a row per 16 threads), and
// an LDS reduction adds the row groups. With adam, one Adam step (Kingma &
// Ba 2015) on the fp32 master weight and its bf16 copies; grad (optional)
// keeps the gradient. Block 0 writes the total weight to *wsum and the
// normalised loss to *loss (each optional). rows <= kTrainBlocksMax.
constexpr int kRedGroups = 16;
static_assert(kTrainBlocksMax <= 16 * kRedGroups, "16 rows per thread cover every partial row");
static_assert(kGuideParams % 4 == 0, "float4 columns");
__global__ __launch_bounds__(256) void guide_reduce_adam_kernel(GuideDev g, const float *__restrict__ partial,
                                                                const float *__restrict__ wl, uint32_t rows,
                                                                float *__restrict__ grad, float *__restrict__ wsum,
                                                                float *__restrict__ loss, int adam, float lr, float c1,
                                                                float c2) {
    __shared__ float4 part[kRedGroups][16];
    __shared__ float tot[2];
    constexpr uint32_t kCols = kGuideParams / 4;
    // every load is issued before the first result is needed: this thread's
    // Adam state, the weight sums and losses, the partial rows
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    const bool owner = threadIdx.x < 64 && i < (uint32_t)kGuideParams;
    float m0 = 0.f, v0 = 0.f, p0 = 0.f;
    if (owner && adam) {
        m0 = g.m[i];
        v0 = g.v[i];
        p0 = g.params[i];
    }
    if (threadIdx.x < 64) {
        // total weight (lanes 0..31) and loss (lanes 32..63) over the rows
        const int which = threadIdx.x / 32;
        float w[kTrainBlocksMax / 32];
#pragma unroll
        for (int k = 0; k < (int)kTrainBlocksMax / 32; ++k) {
            const uint32_t r = threadIdx.x % 32 + 32 * k;
            w[k] = r < rows ? wl[2 * r + which] : 0.f;
        }
        float v = 0.f;
#pragma unroll
        for (int k = 0; k < (int)kTrainBlocksMax / 32; ++k) v += w[k];
        for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off, 32);
        if (threadIdx.x % 32 == 0) tot[which] = v;
    }
    const int c = threadIdx.x % 16, rg = threadIdx.x / 16;
    const uint32_t col = blockIdx.x * 16 + c;
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    if (col < kCols) {
        const float4 *src = reinterpret_cast<const float4 *>(partial) + col;
        float4 v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t r = rg + kRedGroups * k;
            v[k] = r < rows ? src[(size_t)r * kCols] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            s.x += v[k].x;
            s.y += v[k].y;
            s.z += v[k].z;
            s.w += v[k].w;
        }
    }
    part[rg][c] = s;
    __syncthreads();
    const float ws = tot[0];
    const float inv = ws > 0.f ? 1.f / ws : 0.f;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (wsum) *wsum = ws;
        if (loss) *loss = tot[1] * inv;
    }
    if (owner) {
        const int cc = threadIdx.x / 4, e = threadIdx.x % 4;
        float sum = 0.f;
#pragma unroll
        for (int r = 0; r < kRedGroups; ++r) {
            const float4 q = part[r][cc];
            sum += e == 0 ? q.x : e == 1 ? q.y : e == 2 ? q.z : q.w;
        }
        const float gi = sum * inv;
        if (grad) grad[i] = gi;
        if (adam) {
            const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
            const float m = b1 * m0 + (1.f - b1) * gi;
            const float v = b2 * v0 + (1.f - b2) * gi * gi;
            g.m[i] = m;
            g.v[i] = v;
            const float pnew = p0 - lr * (m / c1) / (sqrtf(v / c2) + eps);
            g.params[i] = pnew;
            write_bf16_copy(g, i, pnew);
        }
    }
}

__global__ void zero_kernel(float *__restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

__global__ void guide_sync_kernel(GuideDev g) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (uint32_t)kGuideParams) write_bf16_copy(g, i, g.params[i]);
}

}  // namespace

__device__ void write_bf16_copy(const GuideDev &g, uint32_t i, float p) {
    GuideWeightImage &w = *g.img;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const uint32_t wo = g.w_off[l], bo = g.b_off[l];
        const int dout = g.dout[l];
        if (i >= wo && i < bo) {
            const uint32_t k = (i - wo) / dout, o = (i - wo) % dout;
            __bf16 *wa = l == 0 ? w.w0 : l == 1 ? w.w1 : l == 2 ? w.w2 : w.w3;
            wa[o * kGuideWaLd[l] + guide_wa_pos(l, (int)k)] = (__bf16)p;
            if (l > 0) {
                __bf16 *wb = l == 1 ? w.b1 : l == 2 ? w.b2 : w.b3;
                wb[k * 72 + guide_wb_pos(l, (int)o)] = (__bf16)p;
            }
            return;
        }
        if (i >= bo && i < bo + (uint32_t)dout) {
            w.bias[l][i - bo] = p;
            return;
        }
    }
}

hipError_t guide_launch_infer(const GuideDev &g, const float *in, uint32_t n, float *out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint32_t blocks = min((n + 32 * kInfWaves - 1) / (32 * kInfWaves), kInferGrid);
    guide_infer_kernel<<<blocks, kInfWaves * kWave, 0, st>>>(g, in, nullptr, n, out);
    return hipGetLastError();
}

hipError_t guide_launch_infer_count(const GuideDev &g, const float *in, const uint32_t *n_ptr, uint32_t max_n,
                                    float *out, hipStream_t st) {
    if (max_n == 0) return hipSuccess;
    const uint32_t blocks = min((max_n + 32 * kInfWaves - 1) / (32 * kInfWaves), kInferGrid);
    guide_infer_kernel<<<blocks, kInfWaves * kWave, 0, st>>>(g, in, n_ptr, max_n, out);
    return hipGetLastError();
}

hipError_t guide_launch_infer_nets(const GuideDev *nets, uint32_t n_nets, const float4 *stage, const uint32_t *slots,
                                   const uint32_t *counts, uint32_t count_stride, uint32_t stride, __half *out,
                                   hipStream_t st) {
    if (n_nets == 0 || stride == 0) return hipSuccess;
    const uint32_t blocks = min((stride + 32 * kInfWaves - 1) / (32 * kInfWaves), kInferGrid);
    guide_infer_nets_kernel<<<dim3(blocks, n_nets), kInfWaves * kWave, 0, st>>>(nets, stage, slots, counts,
                                                                                 count_stride, stride, out);
    return hipGetLastError();
}

namespace {
// The fused pass over n records and the reduction (with or without Adam).
hipError_t train_pass(const GuideDev &g, const float *rec, uint32_t n, uint32_t stride, float *grad, float *wsum,
                      float *loss, float *partial, int adam, float lr, int step, hipStream_t st) {
    const uint32_t stride_f = 10u * (stride ? stride : 1u);
    const uint32_t n_tiles = (n + 31) / 32;
    // per block: a whole number of rounds (one tile per wave), at most
    // kTrainBlocksMax blocks
    uint32_t per_block = (n_tiles + kTrainBlocksMax - 1) / kTrainBlocksMax;
    per_block = (per_block + kTrainWaves - 1) / kTrainWaves * kTrainWaves;
    const uint32_t blocks = (n_tiles + per_block - 1) / per_block;
    float *wl = partial + (size_t)kTrainBlocksMax * kGuideParams;
    guide_train_kernel<<<blocks, kTrainThreads, 0, st>>>(g, rec, n, stride_f, per_block, partial, wl);
    const float c1 = 1.f - powf(0.9f, (float)step), c2 = 1.f - powf(0.999f, (float)step);
    guide_reduce_adam_kernel<<<(kGuideParams + 63) / 64, 256, 0, st>>>(g, partial, wl, blocks, grad, wsum, loss, adam,
                                                                        lr, c1, c2);
    return hipGetLastError();
}
}  // namespace

hipError_t guide_launch_grad(const GuideDev &g, const float *rec, uint32_t n, uint32_t stride, float *grad,
                             float *wsum, float *loss, float *partial, hipStream_t st) {
    if (n == 0) {
        zero_kernel<<<(kGuideParams + 255) / 256, 256, 0, st>>>(grad, kGuideParams);
        if (wsum) zero_kernel<<<1, 64, 0, st>>>(wsum, 1);
        if (loss) zero_kernel<<<1, 64, 0, st>>>(loss, 1);
        return hipGetLastError();
    }
    return train_pass(g, rec, n, stride, grad, wsum, loss, partial, 0, 0.f, 1, st);
}

hipError_t guide_launch_train_step(const GuideDev &g, const float *rec, uint32_t n, uint32_t stride, float lr,
                                   int step, float *loss, float *partial, hipStream_t st) {
    if (n == 0) return hipSuccess;
    return train_pass(g, rec, n, stride, nullptr, nullptr, loss, partial, 1, lr, step, st);
}

hipError_t guide_launch_sync(const GuideDev &g, hipStream_t st) {
    guide_sync_kernel<<<(kGuideParams + 255) / 256, 256, 0, st>>>(g);
    return hipGetLastError();
}

}  // namespace illume
