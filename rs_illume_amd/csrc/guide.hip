// Learned sampling proposal: 144-64-64-64-40 ReLU MLP from the frequency
// encoding of (normalised hit position, normal) to an 8-lobe vMF mixture.
// Inference and training (forward, backward, Adam) run on MFMA with bf16
// weights and activations and fp32 accumulation
// (v_mfma_f32_16x16x32_bf16).
//
// One wave owns a tile of 32 rows (two 16-row MFMA tiles). Activations of
// every layer stay in the wave's LDS; weights are read as MFMA B operands
// straight from HBM/L2 (40 KB per copy, reused by every wave). Weight
// gradients are the tile's A^T * dZ products, added to one fp32 gradient
// vector with float atomics; Adam then updates the fp32 master weights and
// rewrites both bf16 copies (n-major for forward B operands, k-major for the
// backward dA = dZ * W^T products).
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include "guide_kernels.h"

namespace illume {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kRows = 32;    // rows per wave
constexpr int kWave = 64;
constexpr int kEnc = 144, kEncPad = 160, kHid = 64, kOut = 40, kOutPad = 48, kLobes = 8;
constexpr uint32_t kGradBlocks = 256;  // target grid of the gradient kernel
constexpr int kGradWaves = 2;          // waves per gradient block (LDS: 2 x 36 KB + 79 KB)
static_assert(kGradPartialFloats == (size_t)kGradBlocks * kGuideParams, "one partial row per gradient block");

__device__ __forceinline__ floatx4 mfma(bf16x8 a, bf16x8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

// A operand of a row-major [rows][ld] bf16 matrix in LDS: rows r0..r0+15,
// k0..k0+31. Lane l holds row r0 + (l & 15), k = k0 + (l >> 4) * 8 + 0..7.
__device__ __forceinline__ bf16x8 lds_a(const __bf16 *m, int ld, int r0, int k0, int lane) {
    return *reinterpret_cast<const bf16x8 *>(m + (r0 + (lane & 15)) * ld + k0 + (lane >> 4) * 8);
}
// B operand from an n-major [n][ld] bf16 matrix (global): columns n0..n0+15,
// k0..k0+31. Lane l holds column n0 + (l & 15), k = k0 + (l >> 4) * 8 + 0..7.
__device__ __forceinline__ bf16x8 glb_b(const __bf16 *__restrict__ w, int ld, int n0, int k0, int lane) {
    return *reinterpret_cast<const bf16x8 *>(w + (n0 + (lane & 15)) * ld + k0 + (lane >> 4) * 8);
}
// Transposed operand from a row-major [rows][ld] LDS matrix: element (i, k) =
// m[k][i]; i = i0 + (lane & 15), k = k0 + (lane >> 4) * 8 + 0..7.
__device__ __forceinline__ bf16x8 lds_t(const __bf16 *m, int ld, int i0, int k0, int lane) {
    bf16x8 v;
    const int i = i0 + (lane & 15), k = k0 + (lane >> 4) * 8;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = m[(k + e) * ld + i];
    return v;
}

struct WaveLds {
    __bf16 a0[kRows * kEncPad];
    __bf16 a1[kRows * kHid], a2[kRows * kHid], a3[kRows * kHid];
    float y[kRows * kOutPad];
    __bf16 dz[2][kRows * kHid];
};

// Frequency encoding: input i (position xyz in [0,1], normal xyz), octave j:
// sin(2^j pi v) at 24 i + 2 j, cos at 24 i + 2 j + 1; columns 144..159 zero.
// The angle is reduced exactly (2^j pi v = 2 pi fract(2^(j-1) v): scaling by
// a power of two and taking the fraction are exact) and the hardware
// v_sin/v_cos, which take their argument in revolutions, do the rest.
__device__ void encode_tile(const float *__restrict__ in, int in_stride, uint32_t row0, uint32_t n, __bf16 *a0,
                            int lane) {
    for (int idx = lane; idx < kRows * kEncPad; idx += kWave) {
        const int r = idx / kEncPad, f = idx % kEncPad;
        float val = 0.f;
        const uint32_t row = row0 + r;
        if (f < kEnc && row < n) {
            const int i = f / 24, j = (f % 24) / 2;
            const float v = in[(size_t)row * in_stride + i];
            const float r = ldexpf(v, j - 1);
            const float rev = r - floorf(r);
            val = (f & 1) ? __builtin_amdgcn_cosf(rev) : __builtin_amdgcn_sinf(rev);
        }
        a0[idx] = to_bf16(val);
    }
}

// out_tile = act(A[32 x K] * W[K x N] + b): W given n-major bf16 [N][K].
// Every B fragment of the layer (N/16 x K/32 of them, 4 VGPRs each) is loaded
// before the first MFMA, so the wave waits for global memory once per layer
// rather than once per product.
template <int K, int N, bool RELU>
__device__ void dense(const __bf16 *a, const __bf16 *__restrict__ wt, const float *__restrict__ b, __bf16 *out_bf,
                      float *out_f, int lane) {
    constexpr int KS = K / 32, NT = N / 16;
    bf16x8 bw[NT][KS];
    float bias[NT];
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) bw[ct][ks] = glb_b(wt, K, ct * 16, ks * 32, lane);
        bias[ct] = b[ct * 16 + (lane & 15)];
    }
#pragma unroll
    for (int rt = 0; rt < kRows / 16; ++rt) {
        bf16x8 af[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) af[ks] = lds_a(a, K, rt * 16, ks * 32, lane);
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) acc = mfma(af[ks], bw[ct][ks], acc);
            const int col = ct * 16 + (lane & 15);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = rt * 16 + (lane >> 4) * 4 + e;
                float v = acc[e] + bias[ct];
                if (RELU) {
                    v = v > 0.f ? v : 0.f;
                    out_bf[row * N + col] = to_bf16(v);
                } else {
                    out_f[row * N + col] = v;
                }
            }
        }
    }
}

// Forward pass of one tile keeping every layer's activations (for the
// backward pass). A wave with active == false only takes part in the block's
// barriers.
__device__ void forward_tile(const GuideDev &g, const float *__restrict__ in, int in_stride, uint32_t row0,
                             uint32_t n, WaveLds &s, int lane, bool active) {
    if (active) encode_tile(in, in_stride, row0, n, s.a0, lane);
    __syncthreads();
    if (active) dense<kEncPad, kHid, true>(s.a0, g.wt[0], g.bias[0], s.a1, nullptr, lane);
    __syncthreads();
    if (active) dense<kHid, kHid, true>(s.a1, g.wt[1], g.bias[1], s.a2, nullptr, lane);
    __syncthreads();
    if (active) dense<kHid, kHid, true>(s.a2, g.wt[2], g.bias[2], s.a3, nullptr, lane);
    __syncthreads();
    if (active) dense<kHid, kOutPad, false>(s.a3, g.wt[3], g.bias[3], nullptr, s.y, lane);
    __syncthreads();
}

// Inference keeps only what the next layer reads: the encoding buffer (10 KB)
// and one hidden buffer (4 KB); hidden layers alternate between them and the
// fp32 outputs (6 KB) reuse the encoding buffer. 14 KB per wave lets 11 waves
// share a CU's LDS.
struct alignas(16) InferLds {
    __bf16 a0[kRows * kEncPad];
    __bf16 h[kRows * kHid];
};
static_assert(kRows * kOutPad * sizeof(float) <= sizeof(InferLds::a0), "outputs fit the encoding buffer");

__device__ const float *forward_infer(const GuideDev &g, const float *__restrict__ in, uint32_t row0, uint32_t n,
                                      InferLds &s, int lane) {
    encode_tile(in, 6, row0, n, s.a0, lane);
    __syncthreads();
    dense<kEncPad, kHid, true>(s.a0, g.wt[0], g.bias[0], s.h, nullptr, lane);
    __syncthreads();
    dense<kHid, kHid, true>(s.h, g.wt[1], g.bias[1], s.a0, nullptr, lane);
    __syncthreads();
    dense<kHid, kHid, true>(s.a0, g.wt[2], g.bias[2], s.h, nullptr, lane);
    __syncthreads();
    float *y = reinterpret_cast<float *>(s.a0);
    dense<kHid, kOutPad, false>(s.h, g.wt[3], g.bias[3], nullptr, y, lane);
    __syncthreads();
    return y;
}

__global__ __launch_bounds__(kWave) void guide_infer_kernel(GuideDev g, const float *__restrict__ in, uint32_t n,
                                                            float *__restrict__ out) {
    __shared__ InferLds s;
    const int lane = threadIdx.x;
    const uint32_t row0 = blockIdx.x * kRows;
    const float *y = forward_infer(g, in, row0, n, s, lane);
    for (int idx = lane; idx < kRows * kOut; idx += kWave) {
        const int r = idx / kOut, c = idx % kOut;
        if (row0 + r < n) out[(size_t)(row0 + r) * kOut + c] = y[r * kOutPad + c];
    }
}

// Same, with the row count read from device memory (queues filled by an
// earlier kernel); blocks beyond the count exit at once.
__global__ __launch_bounds__(kWave) void guide_infer_count_kernel(GuideDev g, const float *__restrict__ in,
                                                                  const uint32_t *__restrict__ n_ptr, uint32_t max_n,
                                                                  float *__restrict__ out) {
    __shared__ InferLds s;
    const uint32_t n = min(*n_ptr, max_n);
    const uint32_t row0 = blockIdx.x * kRows;
    if (row0 >= n) return;
    const int lane = threadIdx.x;
    const float *y = forward_infer(g, in, row0, n, s, lane);
    for (int idx = lane; idx < kRows * kOut; idx += kWave) {
        const int r = idx / kOut, c = idx % kOut;
        if (row0 + r < n) out[(size_t)(row0 + r) * kOut + c] = y[r * kOutPad + c];
    }
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

// dW[in][out] += A^T dZ over the tile's rows; A [32][KA] (bf16, LDS),
// dZ [32][NZ] (bf16, LDS); only in < in_valid and out < out_valid are stored.
// gw is the block's fp32 accumulator in LDS, shared by the block's waves
// (LDS float atomics).
template <int KA, int NZ>
__device__ void accumulate_dw(const __bf16 *a, const __bf16 *dz, float *gw, int in_valid, int out_valid, int lane) {
    for (int it = 0; it < (in_valid + 15) / 16; ++it) {
        for (int ot = 0; ot < (out_valid + 15) / 16; ++ot) {
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(lds_t(a, KA, it * 16, 0, lane), lds_t(dz, NZ, ot * 16, 0, lane), acc);
            const int o = ot * 16 + (lane & 15);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = it * 16 + (lane >> 4) * 4 + e;
                if (i < in_valid && o < out_valid) atomicAdd(gw + i * out_valid + o, acc[e]);
            }
        }
    }
}

// dA[32][64] = dZ[32][KZ] * W^T with W k-major bf16 [64(in)][KZ(out)];
// dZ_prev = dA * (A_prev > 0) -> bf16; the bias gradient of the previous
// layer (column sums of dZ_prev) is added to gb_prev (LDS). As in dense(),
// all B fragments are loaded first.
template <int KZ>
__device__ void backprop(const __bf16 *dz, const __bf16 *__restrict__ wk, const __bf16 *a_prev, __bf16 *dz_prev,
                         float *gb_prev, int lane) {
    constexpr int KS = KZ / 32, NT = kHid / 16;
    bf16x8 bw[NT][KS];
#pragma unroll
    for (int ct = 0; ct < NT; ++ct)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) bw[ct][ks] = glb_b(wk, KZ, ct * 16, ks * 32, lane);
#pragma unroll
    for (int rt = 0; rt < kRows / 16; ++rt) {
        bf16x8 af[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) af[ks] = lds_a(dz, KZ, rt * 16, ks * 32, lane);
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) acc = mfma(af[ks], bw[ct][ks], acc);
            const int col = ct * 16 + (lane & 15);
            float colsum = 0.f;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = rt * 16 + (lane >> 4) * 4 + e;
                const float v = (float)a_prev[row * kHid + col] > 0.f ? acc[e] : 0.f;
                dz_prev[row * kHid + col] = to_bf16(v);
                colsum += v;
            }
            atomicAdd(gb_prev + col, colsum);
        }
    }
}

// records: [n][10] = position (3, normalised), normal (3), direction (3), weight
// (row r at rec + stride_f * r). kGradWaves waves per block; each wave walks
// `tiles` 32-row tiles with its own activations in LDS, and the block
// accumulates the gradient of every parameter in one fp32 LDS array
// (parameter order, LDS float atomics), stored once at the end as the block's
// row of `partial` ([blocks][params]); guide_reduce_kernel sums the rows.
// No global atomics: device-scope float atomics on 20k addresses from every
// block serialise at the memory side. Every wave runs the same number of
// tile iterations (tiles past n are skipped between the barriers), so each
// __syncthreads is reached by all waves.
struct GradLds {
    WaveLds s[kGradWaves];
    float dw[kGuideParams];
};
static_assert(sizeof(GradLds) <= 160 * 1024, "gradient kernel LDS within one CU's 160 KB");

__global__ __launch_bounds__(kGradWaves *kWave) void guide_grad_kernel(
    GuideDev g, const float *__restrict__ rec, uint32_t n, uint32_t stride_f, uint32_t tiles,
    const float *__restrict__ wsum, float *__restrict__ partial, float *__restrict__ loss) {
    __shared__ GradLds L;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    WaveLds &s = L.s[wave];
    for (int i = threadIdx.x; i < kGuideParams; i += kGradWaves * kWave) L.dw[i] = 0.f;
    const float ws = *wsum;
    float block_loss = 0.f;
    for (uint32_t t = 0; t < tiles; ++t) {
        const uint32_t row0 = ((blockIdx.x * tiles + t) * kGradWaves + wave) * kRows;
        const bool active = row0 < n;
        __syncthreads();
        forward_tile(g, rec, (int)stride_f, row0, n, s, lane, active);
        // output gradient: lane r < 32 owns row r
        if (active) {
            float gy[kOut], row_loss = 0.f;
            const uint32_t row = row0 + lane;
            const float wgt = (lane < kRows && row < n) ? rec[(size_t)row * stride_f + 9] : 0.f;
            if (lane < kRows && row < n && wgt > 0.f && ws > 0.f) {
                const float *r = rec + (size_t)row * stride_f;
                const float w3[3] = {r[6], r[7], r[8]};
                nll_grad_row(s.y + lane * kOutPad, w3, wgt / ws, gy, &row_loss);
            } else {
                for (int o = 0; o < kOut; ++o) gy[o] = 0.f;
            }
            block_loss += row_loss;
            if (lane < kRows)
                for (int o = 0; o < kHid; ++o) s.dz[0][lane * kHid + o] = to_bf16(o < kOut ? gy[o] : 0.f);
            // bias of the output layer: column sums over the tile's rows
            float *gb4 = L.dw + g.b_off[3];
            for (int o = 0; o < kOut; ++o) {
                float v = gy[o];
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
                if (lane == 0 && v != 0.f) atomicAdd(gb4 + o, v);
            }
        }
        __syncthreads();
        // layer 4: dW4 = A3^T dY, dZ3 = (dY W4^T) * (A3 > 0)
        if (active) {
            accumulate_dw<kHid, kHid>(s.a3, s.dz[0], L.dw + g.w_off[3], kHid, kOut, lane);
            backprop<kHid>(s.dz[0], g.wk[3], s.a3, s.dz[1], L.dw + g.b_off[2], lane);
        }
        __syncthreads();
        if (active) {
            accumulate_dw<kHid, kHid>(s.a2, s.dz[1], L.dw + g.w_off[2], kHid, kHid, lane);
            backprop<kHid>(s.dz[1], g.wk[2], s.a2, s.dz[0], L.dw + g.b_off[1], lane);
        }
        __syncthreads();
        if (active) {
            accumulate_dw<kHid, kHid>(s.a1, s.dz[0], L.dw + g.w_off[1], kHid, kHid, lane);
            backprop<kHid>(s.dz[0], g.wk[1], s.a1, s.dz[1], L.dw + g.b_off[0], lane);
        }
        __syncthreads();
        if (active) accumulate_dw<kEncPad, kHid>(s.a0, s.dz[1], L.dw + g.w_off[0], kEnc, kHid, lane);
    }
    if (loss) {
        float l = block_loss;
        for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
        if (lane == 0 && l != 0.f) atomicAdd(loss, l);
    }
    __syncthreads();
    float *dst = partial + (size_t)blockIdx.x * kGuideParams;
    for (int i = threadIdx.x; i < kGuideParams; i += kGradWaves * kWave) dst[i] = L.dw[i];
}

// grad[i] = sum over the blocks' rows of partial[b][i] (coalesced over i).
__global__ void guide_reduce_kernel(const float *__restrict__ partial, uint32_t blocks, float *__restrict__ grad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)kGuideParams) return;
    float s = 0.f;
    for (uint32_t b = 0; b < blocks; ++b) s += partial[(size_t)b * kGuideParams + i];
    grad[i] = s;
}

__global__ void guide_wsum_kernel(const float *__restrict__ rec, uint32_t n, uint32_t stride_f,
                                  float *__restrict__ wsum) {
    float s = 0.f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float w = rec[(size_t)i * stride_f + 9];
        if (w > 0.f) s += w;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0 && s != 0.f) atomicAdd(wsum, s);
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
    guide_infer_kernel<<<(n + kRows - 1) / kRows, kWave, 0, st>>>(g, in, n, out);
    return hipGetLastError();
}

hipError_t guide_launch_infer_count(const GuideDev &g, const float *in, const uint32_t *n_ptr, uint32_t max_n,
                                    float *out, hipStream_t st) {
    if (max_n == 0) return hipSuccess;
    guide_infer_count_kernel<<<(max_n + kRows - 1) / kRows, kWave, 0, st>>>(g, in, n_ptr, max_n, out);
    return hipGetLastError();
}

hipError_t guide_launch_grad(const GuideDev &g, const float *rec, uint32_t n, uint32_t stride, float *grad,
                             float *wsum, float *loss, float *partial, hipStream_t st) {
    zero_kernel<<<1, 64, 0, st>>>(wsum, 1);
    if (loss) zero_kernel<<<1, 64, 0, st>>>(loss, 1);
    if (n == 0) {
        zero_kernel<<<(kGuideParams + 255) / 256, 256, 0, st>>>(grad, kGuideParams);
        return hipGetLastError();
    }
    const uint32_t stride_f = 10u * (stride ? stride : 1u);
    guide_wsum_kernel<<<256, 256, 0, st>>>(rec, n, stride_f, wsum);
    // about one block per CU: each block holds 117 KB of LDS
    const uint32_t n_tiles = (n + kRows - 1) / kRows;
    const uint32_t tiles = (n_tiles + kGradBlocks * kGradWaves - 1) / (kGradBlocks * kGradWaves);  // per wave
    const uint32_t blocks = (n_tiles + tiles * kGradWaves - 1) / (tiles * kGradWaves);
    guide_grad_kernel<<<blocks, kGradWaves * kWave, 0, st>>>(g, rec, n, stride_f, tiles, wsum, partial, loss);
    guide_reduce_kernel<<<(kGuideParams + 255) / 256, 256, 0, st>>>(partial, blocks, grad);
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
