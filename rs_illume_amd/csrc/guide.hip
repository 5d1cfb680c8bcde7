// Learned sampling proposal: 144-64-64-64-40 ReLU MLP from the frequency
// encoding of (normalised hit position, normal) to an 8-lobe vMF mixture.
// Inference and training (forward, backward, Adam) run on MFMA with bf16
// weights and activations and fp32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// Every product is computed transposed (weights as the A operand, rows of the
// batch as the B operand), with the hidden units ordered so that an output
// tile's registers are already the next product's B operand: activations and
// their gradients stay in registers through all layers. Weights are staged in
// LDS once per block. Training runs in two passes: per 32-row tile the
// forward and backward products (storing activations and gradients
// transposed), then the weight gradients as MFMA products with K = rows over
// chunks of tiles, then one reduction; Adam updates the fp32 master weights
// and rewrites the bf16 copies.
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
static_assert(kGradPartialFloats >= (size_t)128 * kGuideParams, "one partial row per gradient split");

__device__ __forceinline__ floatx4 mfma(bf16x8 a, bf16x8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

// Frequency encoding: input i (position xyz in [0,1], normal xyz), octave j:
// sin(2^j pi v) at feature 24 i + 2 j, cos at 24 i + 2 j + 1; features
// 144..159 are zero. The angle is reduced exactly (2^j pi v =
// 2 pi fract(2^(j-1) v): scaling by a power of two and taking the fraction
// are exact) and the hardware v_sin/v_cos, which take their argument in
// revolutions, do the rest.

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

__device__ __forceinline__ bf16x8 lds_wa(const __bf16 *w, int ld, int m, int j, int lane) {
    return *reinterpret_cast<const bf16x8 *>(w + (16 * m + (lane & 15)) * ld + 32 * j + (lane >> 4) * 8);
}

// Encoding of one row into layer 0's B operand: lane group g computes
// features 40 g .. 40 g + 39 (input i = f / 24, octave o = (f % 24) / 2,
// sin at even f, cos at odd f); features 144.. are zero.
__device__ __forceinline__ void encode_b(const float *__restrict__ in, uint32_t stride, uint32_t row, uint32_t n, int g,
                                         bf16x8 *b) {
    float v[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) v[i] = row < n ? in[(size_t)row * stride + i] : 0.f;
#pragma unroll
    for (int s = 0; s < 40; s += 2) {
        const int f = 40 * g + s;
        float sv = 0.f, cv = 0.f;
        if (f < kEnc) {
            const int i = f / 24, o = (f % 24) / 2;
            const float x = i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : i == 3 ? v[3] : i == 4 ? v[4] : v[5];
            const float r = ldexpf(x, o - 1);
            const float rev = r - floorf(r);
            sv = __builtin_amdgcn_sinf(rev);
            cv = __builtin_amdgcn_cosf(rev);
        }
        b[s / 8][s % 8] = to_bf16(sv);
        b[s / 8][s % 8 + 1] = to_bf16(cv);
    }
}

// acc (output tiles of one 16-row column) -> bias, ReLU, bf16 -> the next
// layer's B operand (tile m fills k-block m / 2, elements 4 (m % 2) + e).
template <int MT>
__device__ __forceinline__ void relu_to_b(const floatx4 *acc, const float *bias, int g, bf16x8 *b) {
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float h = acc[m][e] + bias[16 * m + 4 * g + e];
            b[m / 2][4 * (m % 2) + e] = to_bf16(h > 0.f ? h : 0.f);
        }
}

template <int KB, int MT>
__device__ __forceinline__ void layer_t(const __bf16 *w, int ld, const bf16x8 (*b)[KB], floatx4 (*acc)[4], int lane) {
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        acc[0][m] = floatx4{0.f, 0.f, 0.f, 0.f};
        acc[1][m] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < KB; ++j) {
            const bf16x8 a = lds_wa(w, ld, m, j, lane);
            acc[0][m] = mfma(a, b[0][j], acc[0][m]);
            acc[1][m] = mfma(a, b[1][j], acc[1][m]);
        }
    }
}

// n rows of in [n][6] -> out [n][40] by the blocks blockIdx.x of one grid
// row; n from n_ptr (at most max_n) when n_ptr is non-null.
__device__ __forceinline__ void infer_rows(const GuideDev &g, const float *__restrict__ in,
                                           const uint32_t *__restrict__ n_ptr, uint32_t n_arg,
                                           float *__restrict__ out) {
    __shared__ InferWeights W;
    const uint32_t n = n_ptr ? min(*n_ptr, n_arg) : n_arg;
    const uint32_t n_iter = (n + 31) / 32;
    if (blockIdx.x * kInfWaves >= n_iter) return;  // the whole block has no rows
    {
        const __bf16 *src[4] = {g.wa[0], g.wa[1], g.wa[2], g.wa[3]};
        __bf16 *dst[4] = {W.w0, W.w1, W.w2, W.w3};
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const int n16 = kGuideWaRows[l] * kGuideWaLd[l] / 8;
            for (int i = threadIdx.x; i < n16; i += kInfWaves * kWave)
                reinterpret_cast<uint4 *>(dst[l])[i] = reinterpret_cast<const uint4 *>(src[l])[i];
        }
        for (int i = threadIdx.x; i < 4 * 64; i += kInfWaves * kWave) {
            const int l = i / 64, o = i % 64;
            W.bias[l][o] = o < kGuideWaRows[l] ? g.bias[l][o] : 0.f;
        }
    }
    __syncthreads();
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave, grp = lane >> 4;
    for (uint32_t it = blockIdx.x * kInfWaves + wave; it < n_iter; it += gridDim.x * kInfWaves) {
        // Re-read the weight fragments from LDS in every iteration: kept in
        // registers across iterations they would take 168 VGPRs.
        __asm__ volatile("" ::: "memory");
        const uint32_t base = it * 32;
        bf16x8 b0[2][5], b1[2][2];
        floatx4 acc[2][4];
        encode_b(in, 6, base + (lane & 15), n, grp, b0[0]);
        encode_b(in, 6, base + 16 + (lane & 15), n, grp, b0[1]);
        layer_t<5, 4>(W.w0, 168, b0, acc, lane);
        relu_to_b<4>(acc[0], W.bias[0], grp, b1[0]);
        relu_to_b<4>(acc[1], W.bias[0], grp, b1[1]);
        layer_t<2, 4>(W.w1, 72, b1, acc, lane);
        relu_to_b<4>(acc[0], W.bias[1], grp, b1[0]);
        relu_to_b<4>(acc[1], W.bias[1], grp, b1[1]);
        layer_t<2, 4>(W.w2, 72, b1, acc, lane);
        relu_to_b<4>(acc[0], W.bias[2], grp, b1[0]);
        relu_to_b<4>(acc[1], W.bias[2], grp, b1[1]);
        layer_t<2, 3>(W.w3, 72, b1, acc, lane);
#pragma unroll
        for (int nn = 0; nn < 2; ++nn) {
            const uint32_t row = base + 16 * nn + (lane & 15);
            if (row >= n) continue;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const int o = 16 * m + 4 * grp;
                if (o >= kOut) continue;
                const float4 y = make_float4(acc[nn][m][0] + W.bias[3][o], acc[nn][m][1] + W.bias[3][o + 1],
                                             acc[nn][m][2] + W.bias[3][o + 2], acc[nn][m][3] + W.bias[3][o + 3]);
                *reinterpret_cast<float4 *>(out + (size_t)row * kOut + o) = y;
            }
        }
    }
}

__global__ __launch_bounds__(kInfWaves *kWave) void guide_infer_kernel(GuideDev g, const float *__restrict__ in,
                                                                       const uint32_t *__restrict__ n_ptr,
                                                                       uint32_t n_arg, float *__restrict__ out) {
    infer_rows(g, in, n_ptr, n_arg, out);
}

// One grid row per network k (blockIdx.y): its rows start at row k * stride
// of in and out, their count is counts[k].
__global__ __launch_bounds__(kInfWaves *kWave) void guide_infer_nets_kernel(const GuideDev *__restrict__ nets,
                                                                            const float *__restrict__ in,
                                                                            const uint32_t *__restrict__ counts,
                                                                            uint32_t count_stride, uint32_t stride,
                                                                            float *__restrict__ out) {
    const uint32_t k = blockIdx.y;
    const GuideDev g = nets[k];
    infer_rows(g, in + 6 * (size_t)k * stride, counts + (size_t)k * count_stride, stride,
               out + (size_t)kOut * k * stride);
}

// d(-log q)/d outputs for one row (mixture decode as in the sampler).
__device__ void nll_grad_row(const float *y, const float *w, float scale, float *gy, float *loss) {
    float mx = y[0];
    for (int j = 1; j < kLobes; ++j) mx = fmaxf(mx, y[5 * j]);
    float pi[kLobes], s = 0.f;
    for (int j = 0; j < kLobes; ++j) {
        pi[j] = expf(y[5 * j] - mx);
        s += pi[j];
    }
    float v[kLobes], mu[kLobes][3], kap[kLobes], mn[kLobes], q = 0.f;
    bool clamped[kLobes];
    for (int j = 0; j < kLobes; ++j) {
        pi[j] /= s;
        const float a = y[5 * j + 1], b = y[5 * j + 2], c = y[5 * j + 3];
        const float len = sqrtf(a * a + b * b + c * c);
        mn[j] = len;
        const float inv = len > 0.f ? 1.f / len : 0.f;
        mu[j][0] = a * inv;
        mu[j][1] = b * inv;
        mu[j][2] = len > 0.f ? c * inv : 1.f;
        const float r = y[5 * j + 4];
        clamped[j] = r < -4.6f || r > 6.9f;
        kap[j] = expf(fminf(fmaxf(r, -4.6f), 6.9f));
        const float cdot = mu[j][0] * w[0] + mu[j][1] * w[1] + mu[j][2] * w[2];
        v[j] = kap[j] / (6.28318530717958647692f * (-expm1f(-2.f * kap[j]))) * expf(kap[j] * (cdot - 1.f));
        q += pi[j] * v[j];
    }
    *loss = -scale * logf(q);
    for (int j = 0; j < kLobes; ++j) {
        const float gam = pi[j] * v[j] / q;
        gy[5 * j] = scale * (pi[j] - gam);
        const float k = kap[j];
        const float cdot = mu[j][0] * w[0] + mu[j][1] * w[1] + mu[j][2] * w[2];
        const float dlk = 1.f / k - 2.f * expf(-2.f * k) / (-expm1f(-2.f * k)) + (cdot - 1.f);
        gy[5 * j + 4] = clamped[j] ? 0.f : scale * (-gam * dlk * k);
        const float inv = mn[j] > 0.f ? 1.f / mn[j] : 0.f;
        for (int a = 0; a < 3; ++a) gy[5 * j + 1 + a] = scale * (-gam * k * (w[a] - cdot * mu[j][a]) * inv);
    }
}

// ------------------------------------------------------------------ training
// Phase A (guide_train_tile_kernel), per 32-row tile and per wave: the
// forward pass as in inference, keeping the activations in registers; the
// outputs go to the wave's LDS, where lanes 0..31 turn their row into the
// NLL gradient; the backward products dX^T = W G^T then run with the weights
// as A operands (wb copies, K in register order) so that every gradient again
// lands in the registers that the next backward product reads. Activations
// X_1..X_3 and pre-activation gradients G_0..G_3 are stored transposed,
// [unit][32 rows] bf16, for phase B.
// Phase B (guide_train_dw_kernel): dW_l = X_l^T G_l as MFMA products with
// K = rows, split over chunks of tiles (one partial row per chunk), X_0 (the
// encoding) recomputed from the records; the bias gradients are the column
// sums of G_l. guide_reduce_kernel adds the partial rows.
constexpr int kTrainWaves = 6;
constexpr uint32_t kTrainGrid = 256;   // phase A blocks at most (one per CU: 140 KB of LDS each)
constexpr uint32_t kTrainSplits = 128;  // phase B row chunks
constexpr int kDwItems = 9 + 4 + 4 + 4;  // (layer, 16-unit input tile) pairs

struct alignas(16) TrainWeights {
    __bf16 w0[64 * 168];
    __bf16 w1[64 * 72];
    __bf16 w2[64 * 72];
    __bf16 w3[48 * 72];
    __bf16 b1[64 * 72];
    __bf16 b2[64 * 72];
    __bf16 b3[64 * 72];
    float bias[4][64];
};
struct alignas(16) TrainWaveLds {
    float y[32][48];     // outputs of the tile, then their gradients
    __bf16 tr[64 * 40];  // transpose staging, [unit][rows] with rows padded to 40
};
struct TrainLds {
    TrainWeights w;
    TrainWaveLds wave[kTrainWaves];
};
static_assert(sizeof(TrainLds) <= 160 * 1024, "training LDS within one CU's 160 KB");

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

// Stores a 64-unit tensor held in register order (v[n] for the two 16-row
// columns) as [unit][32 rows] bf16 at dst, through the wave's LDS.
__device__ __forceinline__ void store_units(TrainWaveLds &S, const bf16x8 (*v)[2], __bf16 *dst, int lane) {
    const int g = lane >> 4, r = lane & 15;
#pragma unroll
    for (int nn = 0; nn < 2; ++nn)
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int e = 0; e < 4; ++e) S.tr[(16 * m + 4 * g + e) * 40 + 16 * nn + r] = v[nn][m / 2][4 * (m % 2) + e];
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = lane + 64 * q, unit = c / 4, part = c % 4;
        reinterpret_cast<uint4 *>(dst)[c] = *reinterpret_cast<const uint4 *>(S.tr + unit * 40 + part * 8);
    }
    wave_lds_sync();
}

__global__ __launch_bounds__(kTrainWaves *kWave) void guide_train_tile_kernel(
    GuideDev g, const float *__restrict__ rec, uint32_t n, uint32_t stride_f, float *__restrict__ wsum,
    __bf16 *__restrict__ tiles, float *__restrict__ loss) {
    __shared__ TrainLds L;
    const uint32_t n_tiles = (n + 31) / 32;
    if (blockIdx.x * kTrainWaves >= n_tiles) return;  // the whole block has no rows
    {
        const __bf16 *src[7] = {g.wa[0], g.wa[1], g.wa[2], g.wa[3], g.wb[1], g.wb[2], g.wb[3]};
        __bf16 *dst[7] = {L.w.w0, L.w.w1, L.w.w2, L.w.w3, L.w.b1, L.w.b2, L.w.b3};
        const int n16[7] = {64 * 168 / 8, 64 * 72 / 8, 64 * 72 / 8, 48 * 72 / 8, 64 * 72 / 8, 64 * 72 / 8, 64 * 72 / 8};
#pragma unroll
        for (int l = 0; l < 7; ++l)
            for (int i = threadIdx.x; i < n16[l]; i += kTrainWaves * kWave)
                reinterpret_cast<uint4 *>(dst[l])[i] = reinterpret_cast<const uint4 *>(src[l])[i];
        for (int i = threadIdx.x; i < 4 * 64; i += kTrainWaves * kWave) {
            const int l = i / 64, o = i % 64;
            L.w.bias[l][o] = o < kGuideWaRows[l] ? g.bias[l][o] : 0.f;
        }
    }
    __syncthreads();
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave, grp = lane >> 4, r16 = lane & 15;
    TrainWaveLds &S = L.wave[wave];
    float wave_loss = 0.f, wave_wsum = 0.f;
    for (uint32_t t = blockIdx.x * kTrainWaves + wave; t < n_tiles; t += gridDim.x * kTrainWaves) {
        __asm__ volatile("" ::: "memory");
        const uint32_t base = t * 32;
        __bf16 *tile = tiles + (size_t)t * (kGuideTileBytes / 2);
        uint32_t mask1, mask2, mask3;
        {
            bf16x8 x0[2][5], h[2][2];
            floatx4 acc[2][4];
            encode_b(rec, stride_f, base + r16, n, grp, x0[0]);
            encode_b(rec, stride_f, base + 16 + r16, n, grp, x0[1]);
            layer_t<5, 4>(L.w.w0, 168, x0, acc, lane);
            relu_to_b<4>(acc[0], L.w.bias[0], grp, h[0]);
            relu_to_b<4>(acc[1], L.w.bias[0], grp, h[1]);
            mask1 = relu_bits(h[0], 0) | relu_bits(h[1], 1);
            store_units(S, h, tile + 0 * 64 * 32, lane);
            layer_t<2, 4>(L.w.w1, 72, h, acc, lane);
            relu_to_b<4>(acc[0], L.w.bias[1], grp, h[0]);
            relu_to_b<4>(acc[1], L.w.bias[1], grp, h[1]);
            mask2 = relu_bits(h[0], 0) | relu_bits(h[1], 1);
            store_units(S, h, tile + 1 * 64 * 32, lane);
            layer_t<2, 4>(L.w.w2, 72, h, acc, lane);
            relu_to_b<4>(acc[0], L.w.bias[2], grp, h[0]);
            relu_to_b<4>(acc[1], L.w.bias[2], grp, h[1]);
            mask3 = relu_bits(h[0], 0) | relu_bits(h[1], 1);
            store_units(S, h, tile + 2 * 64 * 32, lane);
            layer_t<2, 3>(L.w.w3, 72, h, acc, lane);
#pragma unroll
            for (int nn = 0; nn < 2; ++nn)
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const int o = 16 * m + 4 * grp;
                    *reinterpret_cast<float4 *>(&S.y[16 * nn + r16][o]) =
                        make_float4(acc[nn][m][0] + L.w.bias[3][o], acc[nn][m][1] + L.w.bias[3][o + 1],
                                    acc[nn][m][2] + L.w.bias[3][o + 2], acc[nn][m][3] + L.w.bias[3][o + 3]);
                }
        }
        wave_lds_sync();
        // output gradient: lane r < 32 owns row base + r
        if (lane < 32) {
            float gy[kOut], row_loss = 0.f;
            for (int o = 0; o < kOut; ++o) gy[o] = 0.f;
            const uint32_t row = base + lane;
            if (row < n) {
                const float *r = rec + (size_t)row * stride_f;
                const float wgt = r[9];
                if (wgt > 0.f) {
                    const float w3[3] = {r[6], r[7], r[8]};
                    nll_grad_row(S.y[lane], w3, wgt, gy, &row_loss);
                    wave_wsum += wgt;
                }
            }
            wave_loss += row_loss;
            for (int o = 0; o < 48; ++o) S.y[lane][o] = o < kOut ? gy[o] : 0.f;
        }
        wave_lds_sync();
        // G_3: B operand (natural output order, K padded to 64) and its tile
        bf16x8 g3[2][2], gz[2][2];
        floatx4 acc[2][4];
#pragma unroll
        for (int nn = 0; nn < 2; ++nn)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = 32 * j + 8 * grp + e;
                    g3[nn][j][e] = to_bf16(k < 48 ? S.y[16 * nn + r16][k] : 0.f);
                }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = lane + 64 * q, unit = c / 4, part = c % 4;
            bf16x8 v;
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = to_bf16(unit < 48 ? S.y[part * 8 + e][unit] : 0.f);
            reinterpret_cast<bf16x8 *>(tile + 6 * 64 * 32)[c] = v;
        }
        // backward: G_2 = (W_3 G_3^T) * relu'(h3), then G_1, G_0 likewise
        layer_t<2, 4>(L.w.b3, 72, g3, acc, lane);
        mask_to_b(acc[0], mask3, 0, gz[0]);
        mask_to_b(acc[1], mask3, 1, gz[1]);
        store_units(S, gz, tile + 5 * 64 * 32, lane);
        layer_t<2, 4>(L.w.b2, 72, gz, acc, lane);
        mask_to_b(acc[0], mask2, 0, gz[0]);
        mask_to_b(acc[1], mask2, 1, gz[1]);
        store_units(S, gz, tile + 4 * 64 * 32, lane);
        layer_t<2, 4>(L.w.b1, 72, gz, acc, lane);
        mask_to_b(acc[0], mask1, 0, gz[0]);
        mask_to_b(acc[1], mask1, 1, gz[1]);
        store_units(S, gz, tile + 3 * 64 * 32, lane);
    }
    float l = wave_loss, w = wave_wsum;
    for (int off = 32; off > 0; off >>= 1) {
        l += __shfl_xor(l, off);
        w += __shfl_xor(w, off);
    }
    if (lane == 0) {
        if (loss && l != 0.f) atomicAdd(loss, l);
        if (w != 0.f) atomicAdd(wsum, w);
    }
}

// One wave per (split, layer, input tile): dW over the split's tiles for all
// of the layer's output tiles, written to partial row `split`; input tile 0
// also writes the bias gradients (column sums of G).
__global__ __launch_bounds__(kWave) void guide_train_dw_kernel(GuideDev g, const float *__restrict__ rec, uint32_t n,
                                                               uint32_t stride_f, const __bf16 *__restrict__ tiles,
                                                               uint32_t n_tiles, uint32_t chunk,
                                                               float *__restrict__ partial) {
    const int item = blockIdx.x % kDwItems;
    const uint32_t split = blockIdx.x / kDwItems;
    const int l = item < 9 ? 0 : 1 + (item - 9) / 4;
    const int it = item < 9 ? item : (item - 9) % 4;
    const int lane = threadIdx.x, grp = lane >> 4, r16 = lane & 15;
    const int n_out = l == 3 ? 3 : 4;
    const uint32_t t0 = split * chunk, t1 = min(n_tiles, t0 + chunk);
    floatx4 acc[4];
    float colsum[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        acc[u] = floatx4{0.f, 0.f, 0.f, 0.f};
        colsum[u] = 0.f;
    }
#pragma unroll 4
    for (uint32_t t = t0; t < t1; ++t) {
        const __bf16 *tile = tiles + (size_t)t * (kGuideTileBytes / 2);
        bf16x8 a;
        if (l == 0) {
            // encoding feature f of rows 8 grp .. 8 grp + 7 of the tile
            const int f = 16 * it + r16, i = f / 24, o = (f % 24) / 2;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t row = t * 32 + 8 * grp + e;
                const float v = row < n ? rec[(size_t)row * stride_f + i] : 0.f;
                const float r = ldexpf(v, o - 1);
                const float rev = r - floorf(r);
                a[e] = to_bf16((f & 1) ? __builtin_amdgcn_cosf(rev) : __builtin_amdgcn_sinf(rev));
            }
        } else {
            a = reinterpret_cast<const bf16x8 *>(tile + ((l - 1) * 64 + 16 * it + r16) * 32)[grp];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (u >= n_out) break;
            const bf16x8 b = reinterpret_cast<const bf16x8 *>(tile + ((3 + l) * 64 + 16 * u + r16) * 32)[grp];
            acc[u] = mfma(a, b, acc[u]);
            if (it == 0)
#pragma unroll
                for (int e = 0; e < 8; ++e) colsum[u] += (float)b[e];
        }
    }
    float *prow = partial + (size_t)split * kGuideParams;
    const int din = g.din[l], dout = g.dout[l];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (u >= n_out) break;
        const int out = 16 * u + r16;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int in = 16 * it + 4 * grp + e;
            if (in < din && out < dout) prow[g.w_off[l] + in * dout + out] = acc[u][e];
        }
        if (it == 0) {
            float c = colsum[u];
            c += __shfl_xor(c, 16);
            c += __shfl_xor(c, 32);
            if (grp == 0 && out < dout) prow[g.b_off[l] + out] = c;
        }
    }
}

// grad[i] = (sum over the splits' rows of partial[s][i]) / wsum: 32
// parameters per block, 8 threads per parameter each summing every eighth
// row, then an LDS reduction. The tile pass weighted rows by w, not w / sum w;
// the normalisation (of the loss too) happens here.
__global__ __launch_bounds__(256) void guide_reduce_kernel(const float *__restrict__ partial, uint32_t rows,
                                                           const float *__restrict__ wsum, float *__restrict__ grad,
                                                           float *__restrict__ loss) {
    __shared__ float part[8][32];
    const int p = threadIdx.x % 32, q = threadIdx.x / 32;
    const uint32_t i = blockIdx.x * 32 + p;
    float s = 0.f;
    if (i < (uint32_t)kGuideParams)
        for (uint32_t r = q; r < rows; r += 8) s += partial[(size_t)r * kGuideParams + i];
    part[q][p] = s;
    __syncthreads();
    const float ws = *wsum;
    const float inv = ws > 0.f ? 1.f / ws : 0.f;
    if (q == 0 && i < (uint32_t)kGuideParams) {
        float t = 0.f;
        for (int k = 0; k < 8; ++k) t += part[k][p];
        grad[i] = t * inv;
    }
    if (loss && blockIdx.x == 0 && threadIdx.x == 0) *loss *= inv;
}


__global__ void zero_kernel(float *__restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// Adam (Kingma & Ba 2015) on the fp32 master weights, then the bf16 copies.
__global__ void guide_adam_kernel(GuideDev g, const float *__restrict__ grad, float lr, float c1, float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)kGuideParams) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gi = grad[i];
    const float m = b1 * g.m[i] + (1.f - b1) * gi;
    const float v = b2 * g.v[i] + (1.f - b2) * gi * gi;
    g.m[i] = m;
    g.v[i] = v;
    const float p = g.params[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    g.params[i] = p;
    write_bf16_copy(g, i, p);
}

__global__ void guide_sync_kernel(GuideDev g) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (uint32_t)kGuideParams) write_bf16_copy(g, i, g.params[i]);
}

}  // namespace

__device__ void write_bf16_copy(const GuideDev &g, uint32_t i, float p) {
    for (int l = 0; l < 4; ++l) {
        const uint32_t wo = g.w_off[l], bo = g.b_off[l];
        const int din = g.din[l], dout = g.dout[l];
        if (i >= wo && i < bo) {
            const uint32_t k = (i - wo) / dout, o = (i - wo) % dout;
            g.wt[l][o * g.wt_ld[l] + k] = (__bf16)p;   // n-major [out][in_pad]
            g.wk[l][k * g.wk_ld[l] + o] = (__bf16)p;   // k-major [in][out_pad]
            g.wa[l][o * kGuideWaLd[l] + guide_wa_pos(l, (int)k)] = (__bf16)p;
            if (l > 0) g.wb[l][k * 72 + guide_wb_pos(l, (int)o)] = (__bf16)p;
            return;
        }
        if (i >= bo && i < bo + (uint32_t)dout) {
            g.bias[l][i - bo] = p;
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

hipError_t guide_launch_infer_nets(const GuideDev *nets, uint32_t n_nets, const float *in, const uint32_t *counts,
                                   uint32_t count_stride, uint32_t stride, float *out, hipStream_t st) {
    if (n_nets == 0 || stride == 0) return hipSuccess;
    const uint32_t blocks = min((stride + 32 * kInfWaves - 1) / (32 * kInfWaves), kInferGrid);
    guide_infer_nets_kernel<<<dim3(blocks, n_nets), kInfWaves * kWave, 0, st>>>(nets, in, counts, count_stride, stride,
                                                                                 out);
    return hipGetLastError();
}

hipError_t guide_launch_grad(const GuideDev &g, const float *rec, uint32_t n, uint32_t stride, float *grad,
                             float *wsum, float *loss, float *partial, void *tiles, hipStream_t st) {
    zero_kernel<<<1, 64, 0, st>>>(wsum, 1);
    if (loss) zero_kernel<<<1, 64, 0, st>>>(loss, 1);
    if (n == 0) {
        zero_kernel<<<(kGuideParams + 255) / 256, 256, 0, st>>>(grad, kGuideParams);
        return hipGetLastError();
    }
    const uint32_t stride_f = 10u * (stride ? stride : 1u);
    const uint32_t n_tiles = (n + 31) / 32;
    const uint32_t blocks = min((n_tiles + kTrainWaves - 1) / kTrainWaves, kTrainGrid);
    guide_train_tile_kernel<<<blocks, kTrainWaves * kWave, 0, st>>>(g, rec, n, stride_f, wsum, (__bf16 *)tiles,
                                                                    loss);
    const uint32_t chunk = (n_tiles + kTrainSplits - 1) / kTrainSplits;
    const uint32_t splits = (n_tiles + chunk - 1) / chunk;
    guide_train_dw_kernel<<<splits * kDwItems, kWave, 0, st>>>(g, rec, n, stride_f, (const __bf16 *)tiles, n_tiles,
                                                                chunk, partial);
    guide_reduce_kernel<<<(kGuideParams + 31) / 32, 256, 0, st>>>(partial, splits, wsum, grad, loss);
    return hipGetLastError();
}

hipError_t guide_launch_adam(const GuideDev &g, const float *grad, float lr, int step, hipStream_t st) {
    const float c1 = 1.f - powf(0.9f, (float)step), c2 = 1.f - powf(0.999f, (float)step);
    guide_adam_kernel<<<(kGuideParams + 255) / 256, 256, 0, st>>>(g, grad, lr, c1, c2);
    return hipGetLastError();
}

hipError_t guide_launch_sync(const GuideDev &g, hipStream_t st) {
    guide_sync_kernel<<<(kGuideParams + 255) / 256, 256, 0, st>>>(g);
    return hipGetLastError();
}

}  // namespace illume
