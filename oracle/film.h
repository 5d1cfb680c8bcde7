// ORACLE — test infrastructure only (see pbrt.h). Reconstruction filters
// (filters/*.rs) and the film's filtered sample splat (core/film.rs).
#pragma once
#include <cstdint>

namespace orc {

enum FilterType { FILTER_BOX = 0, FILTER_TRIANGLE = 1, FILTER_GAUSSIAN = 2, FILTER_MITCHELL = 3, FILTER_LANCZOS = 4 };

// Same layout as illume_filter_desc.
struct FilterDesc {
    int32_t type;
    float radius[2];
    float alpha;
    float b, c;
    float tau;
};

struct Filter {
    FilterDesc d;
    float inv_rx = 0, inv_ry = 0, exp_x = 0, exp_y = 0;
    bool init(const FilterDesc &desc);   // false for an unknown type or a radius outside (0, 32]
    float eval(float x, float y) const;  // Filter::evaluate of the type
};

// Film::new's weight table (film.rs:69-81) and the radius it was made for.
struct FilmFilter {
    float table[16 * 16];
    float rx, ry, irx, iry;
    void init(const Filter &f);
};

// Film::get_sample_bounds (film.rs:97-103) of a width x height image: pixels
// [x0, x1) x [y0, y1) receive camera samples.
void film_sample_bounds(const FilmFilter &ff, int width, int height, int *x0, int *y0, int *x1, int *y1);

// FilmTile::add_sample (film.rs:288-333) with pbrt-v3's bounds: every pixel
// (x, y) with ceil(p - 0.5 - r) <= (x, y) < floor(p - 0.5 + r) + 1 inside the
// image gets w * L and w (rgbw: 4 floats per pixel), w the table entry at
// min(15, floor(|(x - (p - 0.5)) / r * 16|)).
void film_add_sample(const FilmFilter &ff, float *rgbw, int width, int height, float px, float py, const float L[3]);

}  // namespace orc
