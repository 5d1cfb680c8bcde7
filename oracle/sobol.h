// ORACLE — test infrastructure only (see pbrt.h). The Sobol' sampler: the
// reference's sobol_sample_float and sobol_interval_to_index
// (core/low_discrepancy.rs:354-420) over generator matrices built from the
// Joe-Kuo direction numbers, and pbrt-v3's SobolSampler (absent from the
// reference; PBR book 3e 7.7) on top.
#pragma once
#include <cstdint>

namespace orc {

constexpr int kSobolDims = 1024;      // NUM_SOBOL_DIMENSIONS (sobolmatrices.rs:5)
constexpr int kSobolMatrixSize = 52;  // SOBOL_MATRIX_SIZE (sobolmatrices.rs:6)

struct SobolTables {
    uint32_t c[kSobolDims * kSobolMatrixSize];  // SOBOL_MATRICES_32: top 32 bits of the 64-bit columns
    uint64_t vdc[25][kSobolMatrixSize];         // VD_C_SOBOL_MATRICES[m-1][c], c < 52 - 2m
    uint64_t vdc_inv[26][kSobolMatrixSize];     // VD_C_SOBOL_MATRICES_INV[m-1][c], c < 2m
};
const SobolTables &sobol_tables();

float sobol_sample(int64_t index, int dim, uint32_t scramble);                  // low_discrepancy.rs:386-420
uint64_t sobol_interval_to_index(uint32_t m, uint64_t frame, int px, int py);  // low_discrepancy.rs:354-384

// pbrt-v3 SobolSampler over a sample-bounds rectangle [x0, x0 + w) x [y0, y0 + h).
struct SobolPixelMap {
    int x0, y0;
    int64_t resolution;  // smallest power of two >= max(w, h)
    uint32_t log2_res;
    void init(int x0_, int y0_, int w, int h);
    int64_t index_for_sample(int px, int py, int64_t sample_num) const;          // GetIndexForSample
    float sample_dimension(int64_t index, int dim, int px, int py) const;       // SampleDimension
};

}  // namespace orc
