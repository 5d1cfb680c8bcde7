// ORACLE — test infrastructure only (see pbrt.h). Halton sampler.
#include "halton.h"
#include <algorithm>
#include <cmath>
#include "pbrt.h"

namespace orc {

namespace {

// pbrt-v3 RNG::UniformUInt32(b): rejection below (2^32 - b) mod b. The
// reference's uniform_uint32_bounded (rng.rs:50-58) uses `& b` for the
// threshold; pbrt-v3's modulus is taken (it is the unbiased bound).
uint32_t bounded(Rng &rng, uint32_t b) {
    const uint32_t threshold = (~b + 1u) % b;
    while (true) {
        const uint32_t r = rng.uniform_uint32();
        if (r >= threshold) return r % b;
    }
}

HaltonTables build_tables() {
    HaltonTables t;
    int n = 0;
    for (uint32_t c = 2; n < kHaltonDims; ++c) {
        bool prime = true;
        for (uint32_t d = 2; d * d <= c; ++d)
            if (c % d == 0) {
                prime = false;
                break;
            }
        if (prime) t.primes[n++] = c;
    }
    uint32_t sum = 0;
    for (int i = 0; i < kHaltonDims; ++i) {
        t.prime_sums[i] = sum;
        sum += t.primes[i];
    }
    // compute_radical_inverse_permutations (low_discrepancy.rs:1505-1529) with
    // shuffle (sampling.rs:146-153): identity, then Fisher-Yates per base
    t.perms.resize(sum);
    Rng rng;  // default state and stream
    for (int i = 0; i < kHaltonDims; ++i) {
        uint16_t *p = t.perms.data() + t.prime_sums[i];
        const uint32_t b = t.primes[i];
        for (uint32_t j = 0; j < b; ++j) p[j] = (uint16_t)j;
        for (uint32_t j = 0; j < b; ++j) {
            const uint32_t other = j + bounded(rng, b - j);
            std::swap(p[j], p[other]);
        }
    }
    return t;
}

}  // namespace

const HaltonTables &halton_tables() {
    static const HaltonTables t = build_tables();
    return t;
}

uint32_t reverse_bits_32(uint32_t n) {
    n = (n << 16) | (n >> 16);
    n = ((n & 0x00ff00ffu) << 8) | ((n & 0xff00ff00u) >> 8);
    n = ((n & 0x0f0f0f0fu) << 4) | ((n & 0xf0f0f0f0u) >> 4);
    n = ((n & 0x33333333u) << 2) | ((n & 0xccccccccu) >> 2);
    n = ((n & 0x55555555u) << 1) | ((n & 0xaaaaaaaau) >> 1);
    return n;
}

uint64_t reverse_bits_64(uint64_t n) {
    const uint64_t n0 = reverse_bits_32((uint32_t)n), n1 = reverse_bits_32((uint32_t)(n >> 32));
    return (n0 << 32) | n1;
}

uint64_t inverse_radical_inverse(uint32_t base, uint64_t inverse, int n_digits) {
    uint64_t index = 0;
    for (int i = 0; i < n_digits; ++i) {
        const uint64_t digit = inverse % base;
        inverse /= base;
        index = index * base + digit;
    }
    return index;
}

float radical_inverse(int base_index, uint64_t a) {
    if (base_index == 0) return (float)reverse_bits_64(a) * 0x1p-64f;  // low_discrepancy.rs:468-474
    const uint32_t base = halton_tables().primes[base_index];
    const float inv_base = 1.f / (float)base;
    uint64_t reversed = 0;
    float inv_base_n = 1.f;
    while (a != 0) {
        const uint64_t next = a / base;
        const uint64_t digit = a - next * base;
        reversed = reversed * base + digit;
        inv_base_n *= inv_base;
        a = next;
    }
    return std::min((float)reversed * inv_base_n, OneMinusEpsilon);
}

float scrambled_radical_inverse(int base_index, uint64_t a, const uint16_t *perm) {
    const uint32_t base = halton_tables().primes[base_index];
    const float inv_base = 1.f / (float)base;
    uint64_t reversed = 0;
    float inv_base_n = 1.f;
    while (a != 0) {
        const uint64_t next = a / base;
        const uint64_t digit = a - next * base;
        reversed = reversed * base + perm[digit];
        inv_base_n *= inv_base;
        a = next;
    }
    return std::min(inv_base_n * ((float)reversed + inv_base * (float)perm[0] / (1.f - inv_base)), OneMinusEpsilon);
}

namespace {
// extended Euclid: a x + b y = gcd(a, b)
void extended_gcd(int64_t a, int64_t b, int64_t *x, int64_t *y) {
    if (b == 0) {
        *x = 1;
        *y = 0;
        return;
    }
    const int64_t d = a / b;
    int64_t xp, yp;
    extended_gcd(b, a % b, &xp, &yp);
    *x = yp;
    *y = xp - d * yp;
}
int64_t mod_pos(int64_t a, int64_t b) {
    const int64_t r = a - (a / b) * b;
    return r < 0 ? r + b : r;
}
int64_t multiplicative_inverse(int64_t a, int64_t n) {
    int64_t x, y;
    extended_gcd(a, n, &x, &y);
    return mod_pos(x, n);
}
constexpr int kMaxResolution = 128;
}  // namespace

void HaltonPixelMap::init(int res_x, int res_y) {
    const int res[2] = {res_x, res_y};
    for (int i = 0; i < 2; ++i) {
        const int base = i == 0 ? 2 : 3;
        int64_t s = 1;
        int e = 0;
        while (s < std::min(res[i], kMaxResolution)) {
            s *= base;
            ++e;
        }
        scale[i] = s;
        exp[i] = e;
    }
    stride = scale[0] * scale[1];
    mult_inv[0] = multiplicative_inverse(scale[1], scale[0]);
    mult_inv[1] = multiplicative_inverse(scale[0], scale[1]);
}

int64_t HaltonPixelMap::index_for_sample(int px, int py, int64_t sample_num) const {
    int64_t offset = 0;
    if (stride > 1) {
        const int pm[2] = {(int)mod_pos(px, kMaxResolution), (int)mod_pos(py, kMaxResolution)};
        for (int i = 0; i < 2; ++i) {
            const uint64_t dim_offset = inverse_radical_inverse(i == 0 ? 2 : 3, (uint64_t)pm[i], exp[i]);
            offset += (int64_t)dim_offset * (stride / scale[i]) * mult_inv[i];
        }
        offset %= stride;
    }
    return offset + sample_num * stride;
}

float HaltonPixelMap::sample_dimension(int64_t index, int dim) const {
    if (dim == 0) return radical_inverse(0, (uint64_t)index >> exp[0]);
    if (dim == 1) return radical_inverse(1, (uint64_t)(index / scale[1]));
    const HaltonTables &t = halton_tables();
    return scrambled_radical_inverse(dim, (uint64_t)index, t.perms.data() + t.prime_sums[dim]);
}

}  // namespace orc
