// Low-discrepancy samplers on the device, in the reference's float32
// operation order:
//  * Halton: the radical-inverse primitives of core/low_discrepancy.rs
//    (radical_inverse 422-439/466-, scrambled 441-463, inverse_radical_inverse
//    196-205) and pbrt-v3's HaltonSampler pixel-to-index map
//    (GetIndexForSample / SampleDimension, PBR book 3e 7.4);
//  * Sobol': sobol_sample_float and sobol_interval_to_index
//    (low_discrepancy.rs:354-420) over the generator matrices of the first
//    1024 dimensions, and pbrt-v3's SobolSampler (PBR book 3e 7.7).
// One device table per sampler kind: a 4-word header (word 0 = kind), then
// the kind's data.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace illume {

constexpr uint32_t kSamplerHalton = 1, kSamplerSobol = 2;
constexpr uint32_t kTableHeader = 4;    // words
constexpr uint32_t kHaltonDims = 1000;  // PRIME_TABLE_SIZE (low_discrepancy.rs:18)
constexpr int kHaltonMaxRes = 128;      // pbrt-v3 HaltonSampler kMaxResolution
constexpr uint32_t kSobolDims = 1024;   // NUM_SOBOL_DIMENSIONS (sobolmatrices.rs:5)
constexpr uint32_t kSobolCols = 52;     // SOBOL_MATRIX_SIZE (sobolmatrices.rs:6)
constexpr uint32_t kSobolMaxM = 25;     // rows of VD_C_SOBOL_MATRICES

// Halton table: header, [prime, prefix sum of primes] for the 1000 bases,
// then the digit permutations (uint16) of every base back to back.
__device__ __forceinline__ const uint16_t *halton_perms(const uint32_t *table) {
    return reinterpret_cast<const uint16_t *>(table + kTableHeader + 2 * kHaltonDims);
}
// Sobol' table: header, the 32-bit generator columns [1024][52], then the
// interval-to-index matrices as uint64: [25][52] (index bit 2m + c to pixel)
// and [26][52] (pixel bit c to index bits).
__device__ __forceinline__ const uint32_t *sobol_columns(const uint32_t *table) { return table + kTableHeader; }
__device__ __forceinline__ const uint64_t *sobol_vdc(const uint32_t *table) {
    return reinterpret_cast<const uint64_t *>(table + kTableHeader + kSobolDims * kSobolCols);
}
__device__ __forceinline__ const uint64_t *sobol_vdc_inv(const uint32_t *table) {
    return sobol_vdc(table) + kSobolMaxM * kSobolCols;
}

// Per-launch sampler parameters for a sample-bounds rectangle.
struct SamplerLaunch {
    const uint32_t *table;  // null: the random streams
    uint32_t kind;
    // Halton pixel map (HaltonSampler::new)
    int32_t exp0, exp1;
    int64_t scale0, scale1, stride, inv0, inv1;
    // Sobol' pixel grid (SobolSampler::new): bounds origin, power-of-two resolution
    int32_t x0, y0;
    uint32_t log2_res;
    float res;
};

// a / base and a % base, with a 32-bit path for indices below 2^32.
__device__ __forceinline__ void divmod(uint64_t a, uint32_t base, uint64_t &q, uint32_t &r) {
    if ((a >> 32) == 0) {
        const uint32_t a32 = (uint32_t)a, q32 = a32 / base;
        q = q32;
        r = a32 - q32 * base;
    } else {
        q = a / base;
        r = (uint32_t)(a - q * base);
    }
}

__device__ __forceinline__ float radical_inverse_base2(uint64_t a) {
    return (float)__brevll(a) * 0x1p-64f;  // low_discrepancy.rs:468-474
}

__device__ __forceinline__ float radical_inverse_prime(uint32_t base, uint64_t a) {
    const float inv_base = 1.f / (float)base;
    uint64_t reversed = 0;
    float inv_base_n = 1.f;
    while (a != 0) {
        uint64_t next;
        uint32_t digit;
        divmod(a, base, next, digit);
        reversed = reversed * base + digit;
        inv_base_n *= inv_base;
        a = next;
    }
    return fminf((float)reversed * inv_base_n, 0x1.fffffep-1f);
}

__device__ __forceinline__ float scrambled_radical_inverse(uint32_t base, const uint16_t *__restrict__ perm,
                                                           uint64_t a) {
    const float inv_base = 1.f / (float)base;
    uint64_t reversed = 0;
    float inv_base_n = 1.f;
    while (a != 0) {
        uint64_t next;
        uint32_t digit;
        divmod(a, base, next, digit);
        reversed = reversed * base + perm[digit];
        inv_base_n *= inv_base;
        a = next;
    }
    return fminf(inv_base_n * ((float)reversed + inv_base * (float)perm[0] / (1.f - inv_base)), 0x1.fffffep-1f);
}

__device__ __forceinline__ uint64_t inverse_radical_inverse(uint32_t base, uint64_t inverse, int n_digits) {
    uint64_t index = 0;
    for (int i = 0; i < n_digits; ++i) {
        const uint64_t digit = inverse % base;
        inverse /= base;
        index = index * base + digit;
    }
    return index;
}

__device__ __forceinline__ int64_t mod_pos(int64_t a, int64_t b) {
    const int64_t r = a - (a / b) * b;
    return r < 0 ? r + b : r;
}

// HaltonSampler::GetIndexForSample: the global index of sample `sample_num` of pixel (px, py).
__device__ __forceinline__ int64_t halton_index(const SamplerLaunch &h, int px, int py, int64_t sample_num) {
    int64_t offset = 0;
    if (h.stride > 1) {
        const uint64_t d0 = inverse_radical_inverse(2, (uint64_t)mod_pos(px, kHaltonMaxRes), h.exp0);
        const uint64_t d1 = inverse_radical_inverse(3, (uint64_t)mod_pos(py, kHaltonMaxRes), h.exp1);
        offset = (int64_t)d0 * (h.stride / h.scale0) * h.inv0;
        offset += (int64_t)d1 * (h.stride / h.scale1) * h.inv1;
        offset %= h.stride;
    }
    return offset + sample_num * h.stride;
}

// sobol_sample_float with scramble 0: XOR of the columns of the index's set bits.
__device__ __forceinline__ float sobol_value(const uint32_t *__restrict__ table, uint64_t a, uint32_t dim) {
    const uint32_t *col = sobol_columns(table) + dim * kSobolCols;
    uint32_t v = 0;
    while (a != 0) {
        v ^= col[__builtin_ctzll(a)];
        a &= a - 1;
    }
    return fminf((float)v * 0x1p-32f, 0x1.fffffep-1f);
}

// sobol_interval_to_index: the index of sample `frame` in pixel (px, py) of a
// 2^m x 2^m grid.
__device__ __forceinline__ uint64_t sobol_index(const uint32_t *__restrict__ table, uint32_t m, uint64_t frame,
                                                int px, int py) {
    if (m == 0) return 0;  // low_discrepancy.rs:355-357
    const uint64_t *vdc = sobol_vdc(table) + (m - 1) * kSobolCols;
    const uint64_t *inv = sobol_vdc_inv(table) + (m - 1) * kSobolCols;
    uint64_t index = frame << (2 * m), delta = 0;
    for (uint64_t f = frame; f != 0; f &= f - 1) delta ^= vdc[__builtin_ctzll(f)];
    uint64_t b = ((((uint64_t)(uint32_t)px) << m) | (uint64_t)(uint32_t)py) ^ delta;
    for (; b != 0; b &= b - 1) index ^= inv[__builtin_ctzll(b)];
    return index;
}

// SampleDimension for dimensions >= 2 of either kind; past the kind's
// dimensions a splitmix64 hash of (index, dim). Not inlined: the path code
// that draws samples is shared with renderers that only use the random
// sampler, whose register allocation must not carry these loops.
__device__ __noinline__ float lowdisc_dim(const uint32_t *__restrict__ table, uint64_t index, uint32_t dim) {
    if (table[0] == kSamplerHalton) {
        if (dim < kHaltonDims) {
            const uint32_t base = table[kTableHeader + 2 * dim], off = table[kTableHeader + 2 * dim + 1];
            return scrambled_radical_inverse(base, halton_perms(table) + off, index);
        }
    } else if (dim < kSobolDims) {
        return sobol_value(table, index, dim);
    }
    uint64_t z = index * 0x9E3779B97F4A7C15ull + dim;  // splitmix64
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return fminf((float)(uint32_t)(z >> 32) * 0x1p-32f, 0x1.fffffep-1f);
}

}  // namespace illume
