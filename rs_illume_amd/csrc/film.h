// Filtered film: the sample splat of FilmTile::add_sample (core/film.rs:288-333)
// with pbrt-v3's bounds, as a device function over a float RGBW buffer.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace illume {

constexpr int kFilterTable = 16;  // FILTER_TABLE_WIDTH (film.rs:15)

// What the render kernel needs of a filter: its 16 x 16 weight table (device)
// and radius, and the sample bounds of the image (Film::get_sample_bounds).
struct FilmSplat {
    const float *table;
    float rx, ry, irx, iry;
    int32_t x0, y0;      // first sample pixel (may be negative)
    uint32_t wext, hext;  // sample bounds size
};

// Adds w * L and w to every pixel whose filter footprint holds the film
// position (px, py): ceil(p - 0.5 - r) <= pixel < floor(p - 0.5 + r) + 1,
// clipped to the image; w = table[min(15, floor(|(y - (py - 0.5)) / ry * 16|))]
// [min(15, floor(|(x - (px - 0.5)) / rx * 16|))]. Float atomics.
__device__ __forceinline__ void film_splat(const FilmSplat &f, float *__restrict__ rgbw, int width, int height,
                                           float px, float py, float lx, float ly, float lz) {
    const float dx = px - 0.5f, dy = py - 0.5f;
    const int x0 = max((int)ceilf(dx - f.rx), 0), y0 = max((int)ceilf(dy - f.ry), 0);
    const int x1 = min((int)floorf(dx + f.rx) + 1, width), y1 = min((int)floorf(dy + f.ry) + 1, height);
    for (int y = y0; y < y1; ++y) {
        const float fy = fabsf(((float)y - dy) * f.iry * (float)kFilterTable);
        const int iy = min(kFilterTable - 1, (int)floorf(fy));
        for (int x = x0; x < x1; ++x) {
            const float fx = fabsf(((float)x - dx) * f.irx * (float)kFilterTable);
            const int ix = min(kFilterTable - 1, (int)floorf(fx));
            const float w = f.table[iy * kFilterTable + ix];
            float *p = rgbw + 4 * ((size_t)y * width + x);
            atomicAdd(p, lx * w);
            atomicAdd(p + 1, ly * w);
            atomicAdd(p + 2, lz * w);
            atomicAdd(p + 3, w);
        }
    }
}

}  // namespace illume
