// Host side of the Halton sampler: the prime bases, their digit permutations
// (compute_radical_inverse_permutations, core/low_discrepancy.rs:1505-1529,
// with the default-seeded PCG32 of core/rng.rs and pbrt-v3's unbiased bounded
// draw) and the per-resolution pixel map (pbrt-v3 HaltonSampler::new).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>
#include "halton.h"
#include "kernels.h"

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

// [prime, prefix sum] x 1000 as uint32, then the permutations as uint16.
std::vector<uint32_t> make_table() {
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
    std::vector<uint32_t> table(2 * kHaltonDims + (total + 1) / 2, 0u);
    uint16_t *perms = reinterpret_cast<uint16_t *>(table.data() + 2 * kHaltonDims);
    HostPcg rng;
    uint32_t off = 0;
    for (uint32_t i = 0; i < kHaltonDims; ++i) {
        const uint32_t b = primes[i];
        table[2 * i] = b;
        table[2 * i + 1] = off;
        uint16_t *p = perms + off;
        for (uint32_t j = 0; j < b; ++j) p[j] = (uint16_t)j;
        for (uint32_t j = 0; j < b; ++j) std::swap(p[j], p[j + rng.below(b - j)]);
        off += b;
    }
    return table;
}

const std::vector<uint32_t> &host_table() {
    static const std::vector<uint32_t> t = make_table();
    return t;
}

std::mutex g_mutex;
std::map<int, uint32_t *> g_device_tables;  // per device, kept until the process ends

int64_t inverse_mod(int64_t a, int64_t n) {
    // extended Euclid on (a mod n, n)
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

}  // namespace

bool halton_launch(int res_x, int res_y, HaltonLaunch *out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    uint32_t *d = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        auto it = g_device_tables.find(dev);
        if (it != g_device_tables.end()) {
            d = it->second;
        } else {
            const std::vector<uint32_t> &t = host_table();
            if (hipMalloc((void **)&d, t.size() * sizeof(uint32_t)) != hipSuccess) return false;
            if (hipMemcpy(d, t.data(), t.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) {
                (void)hipFree(d);
                return false;
            }
            g_device_tables[dev] = d;
        }
    }
    out->table = d;
    const int res[2] = {res_x, res_y};
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
    return true;
}

}  // namespace illume
