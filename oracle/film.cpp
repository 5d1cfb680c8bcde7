// ORACLE — test infrastructure only (see pbrt.h). Reconstruction filters and
// the filtered film splat, in float32 in the reference's operation order.
#include "film.h"
#include <algorithm>
#include <cmath>

namespace orc {

namespace {
constexpr float kPi = 3.14159265358979323846f;
constexpr int kTable = 16;  // FILTER_TABLE_WIDTH, film.rs:15

// GaussianFilter::gaussian (filters/gaussian.rs:29-31)
float gaussian(float alpha, float d, float expv) { return std::max(std::exp(-alpha * d * d) - expv, 0.f); }

// MitchellFilter::mitchell1d (filters/mitchell.rs:25-36)
float mitchell1d(float b, float c, float x) {
    x = std::fabs(2.f * x);
    if (x > 1.f)
        return ((-b - 6.f * c) * x * x * x + (6.f * b + 30.f * c) * x * x + (-12.f * b - 48.f * c) * x +
                (8.f * b + 24.f * c)) *
               (1.f / 6.f);
    return ((12.f - 9.f * b - 6.f * c) * x * x * x + (-18.f + 12.f * b + 6.f * c) * x * x + (6.f - 2.f * b)) *
           (1.f / 6.f);
}

// LanczosSincFilter::sinc / windowed_sinc (filters/sinc.rs:23-40)
float sinc(float x) {
    x = std::fabs(x);
    if (x < 1e-5f) return 1.f;
    return std::sin(kPi * x) / (kPi * x);
}
float windowed_sinc(float tau, float x, float radius) {
    x = std::fabs(x);
    if (x > radius) return 0.f;
    const float lanczos = sinc(x / tau);
    return sinc(x) * lanczos;
}
}  // namespace

bool Filter::init(const FilterDesc &desc) {
    d = desc;
    if (d.type < FILTER_BOX || d.type > FILTER_LANCZOS) return false;
    if (!(d.radius[0] > 0.f && d.radius[0] <= 32.f && d.radius[1] > 0.f && d.radius[1] <= 32.f)) return false;
    inv_rx = 1.f / d.radius[0];
    inv_ry = 1.f / d.radius[1];
    // GaussianFilter::new (gaussian.rs:17-25)
    exp_x = std::exp(-d.alpha * d.radius[0] * d.radius[0]);
    exp_y = std::exp(-d.alpha * d.radius[1] * d.radius[1]);
    return true;
}

float Filter::eval(float x, float y) const {
    switch (d.type) {
        case FILTER_BOX:  // box_filter.rs:30-32
            return 1.f;
        case FILTER_TRIANGLE:  // triangle.rs:30-32
            return std::max(d.radius[0] - std::fabs(x), 0.f) * std::max(d.radius[1] - std::fabs(y), 0.f);
        case FILTER_GAUSSIAN:  // gaussian.rs:43-45
            return gaussian(d.alpha, x, exp_x) * gaussian(d.alpha, y, exp_y);
        case FILTER_MITCHELL:  // mitchell.rs:49-51
            return mitchell1d(d.b, d.c, x * inv_rx) * mitchell1d(d.b, d.c, y * inv_ry);
        default:  // sinc.rs:52-54
            return windowed_sinc(d.tau, x, d.radius[0]) * windowed_sinc(d.tau, y, d.radius[1]);
    }
}

void FilmFilter::init(const Filter &f) {
    rx = f.d.radius[0];
    ry = f.d.radius[1];
    irx = f.inv_rx;
    iry = f.inv_ry;
    int offset = 0;
    for (int y = 0; y < kTable; ++y)
        for (int x = 0; x < kTable; ++x)
            table[offset++] = f.eval(((float)x + 0.5f) * rx / (float)kTable, ((float)y + 0.5f) * ry / (float)kTable);
}

void film_sample_bounds(const FilmFilter &ff, int width, int height, int *x0, int *y0, int *x1, int *y1) {
    *x0 = (int)std::floor(0.f + 0.5f - ff.rx);
    *y0 = (int)std::floor(0.f + 0.5f - ff.ry);
    *x1 = (int)std::ceil((float)width - 0.5f + ff.rx);
    *y1 = (int)std::ceil((float)height - 0.5f + ff.ry);
}

void film_add_sample(const FilmFilter &ff, float *rgbw, int width, int height, float px, float py, const float L[3]) {
    const float dx = px - 0.5f, dy = py - 0.5f;
    int x0 = (int)std::ceil(dx - ff.rx), y0 = (int)std::ceil(dy - ff.ry);
    int x1 = (int)std::floor(dx + ff.rx) + 1, y1 = (int)std::floor(dy + ff.ry) + 1;
    x0 = std::max(x0, 0);
    y0 = std::max(y0, 0);
    x1 = std::min(x1, width);
    y1 = std::min(y1, height);
    for (int y = y0; y < y1; ++y) {
        const float fy = std::fabs(((float)y - dy) * ff.iry * (float)kTable);
        const int iy = std::min(kTable - 1, (int)std::floor(fy));
        for (int x = x0; x < x1; ++x) {
            const float fx = std::fabs(((float)x - dx) * ff.irx * (float)kTable);
            const int ix = std::min(kTable - 1, (int)std::floor(fx));
            const float w = ff.table[iy * kTable + ix];
            float *p = rgbw + 4 * ((size_t)y * width + x);
            p[0] += L[0] * w;
            p[1] += L[1] * w;
            p[2] += L[2] * w;
            p[3] += w;
        }
    }
}

}  // namespace orc
