// ORACLE — test infrastructure only (see pbrt.h). The Halton sampler: the
// reference's radical-inverse primitives (core/low_discrepancy.rs) and pbrt-v3's
// HaltonSampler (absent from the reference; PBR book 3e 7.4) built on them.
#pragma once
#include <cstdint>
#include <vector>

namespace orc {

constexpr int kHaltonDims = 1000;  // PRIME_TABLE_SIZE (low_discrepancy.rs:18)

struct HaltonTables {
    uint32_t primes[kHaltonDims];      // the first 1000 primes (low_discrepancy.rs:20)
    uint32_t prime_sums[kHaltonDims];  // exclusive prefix sums (low_discrepancy.rs:84)
    std::vector<uint16_t> perms;       // compute_radical_inverse_permutations with the default Rng
};
const HaltonTables &halton_tables();

uint32_t reverse_bits_32(uint32_t n);                                           // low_discrepancy.rs:178-185
uint64_t reverse_bits_64(uint64_t n);                                           // low_discrepancy.rs:189-193
uint64_t inverse_radical_inverse(uint32_t base, uint64_t inverse, int n_digits);  // low_discrepancy.rs:196-205
float radical_inverse(int base_index, uint64_t a);                              // low_discrepancy.rs:422-439, 466-
float scrambled_radical_inverse(int base_index, uint64_t a, const uint16_t *perm);  // low_discrepancy.rs:441-463

// pbrt-v3 HaltonSampler for one sample-bounds resolution: base scales and
// exponents covering min(res, 128) pixels, the sample stride and the
// multiplicative inverses that map a pixel to its first sample index.
struct HaltonPixelMap {
    int exp[2];
    int64_t scale[2];
    int64_t stride;
    int64_t mult_inv[2];
    void init(int res_x, int res_y);
    int64_t index_for_sample(int px, int py, int64_t sample_num) const;  // GetIndexForSample
    float sample_dimension(int64_t index, int dim) const;               // SampleDimension
};

}  // namespace orc
