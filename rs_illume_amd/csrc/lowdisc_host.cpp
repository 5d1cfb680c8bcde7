// Host side of the low-discrepancy samplers: the device tables (built once,
// uploaded once per device and kept until the process ends) and the
// per-resolution sampler parameters.
//  * Halton: the prime bases and their digit permutations
//    (compute_radical_inverse_permutations, core/low_discrepancy.rs:1505-1529,
//    with the default-seeded PCG32 of core/rng.rs and pbrt-v3's unbiased
//    bounded draw); the pixel map of pbrt-v3 HaltonSampler::new.
//  * Sobol': the generator columns of the first 1024 dimensions, expanded
//    from the Joe-Kuo primitive polynomials and initial direction numbers
//    (sobol_init.h) by their recurrence, and the interval-to-index matrices of
//    every grid size 2^m (the 2m x 2m binary map from the low index bits to
//    the packed pixel (x << m) | y, and its inverse).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>
#include "kernels.h"
#include "lowdisc.h"
#include "sobol_init.h"

namespace illume {

namespace {

struct HostPcg {
    uint64_t state = 0x853c49e6748fea9bULL, inc = 0xda3e39cb94b95bdbULL;  // PCG32 defaults (rng.rs:9-10)
    uint32_t operator()() {
        const uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    // uniform in [0, b): draws below (2^32 - b) mod b are rejected
    uint32_t below(uint32_t b) {
        const uint32_t threshold = (0u - b) % b;
        for (;;) {
            const uint32_t r = (*this)();
            if (r >= threshold) return r % b;
        }
    }
};

std::vector<uint32_t> make_halton_table() {
    std::vector<uint32_t> primes;
    primes.reserve(kHaltonDims);
    std::vector<bool> composite(8000, false);
    for (uint32_t c = 2; primes.size() < kHaltonDims; ++c) {
        if (composite[c]) continue;
        primes.push_back(c);
        for (uint32_t m = c * c; m < composite.size(); m += c) composite[m] = true;
    }
    uint32_t total = 0;
    for (uint32_t p : primes) total += p;
    std::vector<uint32_t> table(kTableHeader + 2 * kHaltonDims + (total + 1) / 2, 0u);
    table[0] = kSamplerHalton;
    uint16_t *perms = reinterpret_cast<uint16_t *>(table.data() + kTableHeader + 2 * kHaltonDims);
    HostPcg rng;
    uint32_t off = 0;
    for (uint32_t i = 0; i < kHaltonDims; ++i) {
        const uint32_t b = primes[i];
        table[kTableHeader + 2 * i] = b;
        table[kTableHeader + 2 * i + 1] = off;
        uint16_t *p = perms + off;
        for (uint32_t j = 0; j < b; ++j) p[j] = (uint16_t)j;
        for (uint32_t j = 0; j < b; ++j) std::swap(p[j], p[j + rng.below(b - j)]);
        off += b;
    }
    return table;
}

// direction numbers m_1..m_n of dimension j as 64-bit columns m_k 2^(64-k)
std::vector<uint64_t> sobol_columns64(uint32_t j, uint32_t n) {
    std::vector<uint64_t> m(n);
    if (j == 0) {
        for (uint32_t k = 0; k < n; ++k) m[k] = 1;
    } else {
        const uint32_t poly = kSobolPoly[j];
        const int deg = 31 - __builtin_clz(poly);
        for (int k = 0; k < deg && k < (int)n; ++k) m[k] = kSobolInit[j][k];
        // m_k = 2 a_1 m_(k-1) ^ 4 a_2 m_(k-2) ^ ... ^ 2^deg m_(k-deg) ^ m_(k-deg)
        for (int k = deg; k < (int)n; ++k) {
            uint64_t v = m[k - deg] ^ (m[k - deg] << deg);
            for (int i = 1; i < deg; ++i)
                if ((poly >> (deg - i)) & 1u) v ^= m[k - i] << i;
            m[k] = v;
        }
    }
    for (uint32_t k = 0; k < n; ++k) m[k] <<= 63 - k;
    return m;
}

std::vector<uint32_t> make_sobol_table() {
    const size_t words = kTableHeader + kSobolDims * kSobolCols + 2 * (kSobolMaxM + 1) * kSobolCols * 2;
    std::vector<uint32_t> table(words, 0u);
    table[0] = kSamplerSobol;
    uint32_t *cols = table.data() + kTableHeader;
    for (uint32_t j = 0; j < kSobolDims; ++j) {
        const std::vector<uint64_t> c = sobol_columns64(j, kSobolCols);
        for (uint32_t k = 0; k < kSobolCols; ++k) cols[j * kSobolCols + k] = (uint32_t)(c[k] >> 32);
    }
    uint64_t *vdc = reinterpret_cast<uint64_t *>(cols + kSobolDims * kSobolCols);
    uint64_t *inv = vdc + kSobolMaxM * kSobolCols;
    const std::vector<uint64_t> d1 = sobol_columns64(1, kSobolCols);
    for (uint32_t m = 1; m <= kSobolMaxM + 1; ++m) {
        const uint32_t n = 2 * m;
        if (m <= kSobolMaxM)
            for (uint32_t c = 0; c + n < kSobolCols; ++c) vdc[(m - 1) * kSobolCols + c] = d1[n + c] >> (64 - m);
        // map[b] = packed pixel moved by index bit b; invert by Gauss-Jordan on (map | identity)
        std::vector<uint64_t> a(n), id(n);  // a[r]: packed-pixel bit r over index bits; id[r]: over pixel bits
        for (uint32_t b = 0; b < n; ++b) {
            const uint64_t packed = ((b < m ? 1ull << (m - 1 - b) : 0ull) << m) | (d1[b] >> (64 - m));
            for (uint32_t r = 0; r < n; ++r)
                if ((packed >> r) & 1u) a[r] |= 1ull << b;
        }
        for (uint32_t r = 0; r < n; ++r) id[r] = 1ull << r;
        for (uint32_t c = 0; c < n; ++c) {
            uint32_t p = c;
            while (p < n && !((a[p] >> c) & 1u)) ++p;
            if (p == n) return {};  // singular: cannot happen for Sobol' dimensions 0 and 1
            std::swap(a[p], a[c]);
            std::swap(id[p], id[c]);
            for (uint32_t r = 0; r < n; ++r)
                if (r != c && ((a[r] >> c) & 1u)) {
                    a[r] ^= a[c];
                    id[r] ^= id[c];
                }
        }
        // index bit r of the preimage of pixel bit c is bit c of id[r]
        for (uint32_t c = 0; c < n; ++c) {
            uint64_t v = 0;
            for (uint32_t r = 0; r < n; ++r) v |= ((id[r] >> c) & 1ull) << r;
            inv[(m - 1) * kSobolCols + c] = v;
        }
    }
    return table;
}

std::mutex g_mutex;
std::map<std::pair<int, uint32_t>, uint32_t *> g_device_tables;  // (device, kind), kept until the process ends

int64_t inverse_mod(int64_t a, int64_t n) {
    int64_t r0 = n, r1 = ((a % n) + n) % n, t0 = 0, t1 = 1;
    while (r1 != 0) {
        const int64_t q = r0 / r1, r2 = r0 - q * r1, t2 = t0 - q * t1;
        r0 = r1;
        r1 = r2;
        t0 = t1;
        t1 = t2;
    }
    return ((t0 % n) + n) % n;
}

const uint32_t *device_table(uint32_t kind) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_mutex);
    auto it = g_device_tables.find({dev, kind});
    if (it != g_device_tables.end()) return it->second;
    static const std::vector<uint32_t> halton = make_halton_table();
    static const std::vector<uint32_t> sobol = make_sobol_table();
    const std::vector<uint32_t> &t = kind == kSamplerHalton ? halton : sobol;
    if (t.empty()) return nullptr;
    uint32_t *d = nullptr;
    if (hipMalloc((void **)&d, t.size() * sizeof(uint32_t)) != hipSuccess) return nullptr;
    if (hipMemcpy(d, t.data(), t.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    g_device_tables[{dev, kind}] = d;
    return d;
}

}  // namespace

bool sampler_launch(uint32_t kind, int x0, int y0, int w, int h, SamplerLaunch *out) {
    std::memset(out, 0, sizeof(*out));
    if (kind != kSamplerHalton && kind != kSamplerSobol) return false;
    out->table = device_table(kind);
    if (!out->table) return false;
    out->kind = kind;
    if (kind == kSamplerHalton) {
        const int res[2] = {w, h};
        int64_t scale[2];
        int32_t exps[2];
        for (int i = 0; i < 2; ++i) {
            const int base = i == 0 ? 2 : 3;
            int64_t sc = 1;
            int32_t e = 0;
            while (sc < std::min(res[i], kHaltonMaxRes)) {
                sc *= base;
                ++e;
            }
            scale[i] = sc;
            exps[i] = e;
        }
        out->exp0 = exps[0];
        out->exp1 = exps[1];
        out->scale0 = scale[0];
        out->scale1 = scale[1];
        out->stride = scale[0] * scale[1];
        out->inv0 = scale[0] > 1 ? inverse_mod(scale[1], scale[0]) : 0;
        out->inv1 = scale[1] > 1 ? inverse_mod(scale[0], scale[1]) : 0;
    } else {
        uint32_t m = 0;
        int64_t r = 1;
        while (r < std::max(w, h)) {
            r <<= 1;
            ++m;
        }
        if (m > kSobolMaxM) return false;
        out->x0 = x0;
        out->y0 = y0;
        out->log2_res = m;
        out->res = (float)r;
    }
    return true;
}

}  // namespace illume
