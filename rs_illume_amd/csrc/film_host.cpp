// Reconstruction filters (filters/*.rs) tabulated for the device film, and
// the illume_filter_* / illume_film_resolve entry points.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include "../../include/illume_hip.h"
#include "film.h"
#include "kernels.h"

namespace illume {
int abi_fail(int status, const std::string &msg);
int abi_hip_fail(hipError_t e, const char *what);
}  // namespace illume

using namespace illume;

struct illume_filter {
    illume_filter_desc desc;
    float table[kFilterTable * kFilterTable];
    float *d_table = nullptr;
};

namespace {

constexpr float kPiF = 3.14159265358979323846f;

// Filter::evaluate of each type, float32 in the reference's operation order.
struct FilterEval {
    illume_filter_desc d;
    float inv_rx, inv_ry, exp_x, exp_y;

    explicit FilterEval(const illume_filter_desc &desc) : d(desc) {
        inv_rx = 1.f / d.radius[0];
        inv_ry = 1.f / d.radius[1];
        exp_x = std::exp(-d.alpha * d.radius[0] * d.radius[0]);  // gaussian.rs:17-25
        exp_y = std::exp(-d.alpha * d.radius[1] * d.radius[1]);
    }
    float gauss(float v, float e) const { return std::max(std::exp(-d.alpha * v * v) - e, 0.f); }  // gaussian.rs:29-31
    float mitchell(float v) const {  // mitchell.rs:25-36
        const float B = d.b, C = d.c;
        v = std::fabs(2.f * v);
        if (v > 1.f)
            return ((-B - 6.f * C) * v * v * v + (6.f * B + 30.f * C) * v * v + (-12.f * B - 48.f * C) * v +
                    (8.f * B + 24.f * C)) *
                   (1.f / 6.f);
        return ((12.f - 9.f * B - 6.f * C) * v * v * v + (-18.f + 12.f * B + 6.f * C) * v * v + (6.f - 2.f * B)) *
               (1.f / 6.f);
    }
    static float sinc(float v) {  // sinc.rs:23-30
        v = std::fabs(v);
        return v < 1e-5f ? 1.f : std::sin(kPiF * v) / (kPiF * v);
    }
    float lanczos(float v, float r) const {  // sinc.rs:32-40
        v = std::fabs(v);
        if (v > r) return 0.f;
        const float window = sinc(v / d.tau);
        return sinc(v) * window;
    }
    float operator()(float x, float y) const {
        switch (d.type) {
            case ILLUME_FILTER_BOX: return 1.f;                                          // box_filter.rs:30-32
            case ILLUME_FILTER_TRIANGLE:                                                  // triangle.rs:30-32
                return std::max(d.radius[0] - std::fabs(x), 0.f) * std::max(d.radius[1] - std::fabs(y), 0.f);
            case ILLUME_FILTER_GAUSSIAN: return gauss(x, exp_x) * gauss(y, exp_y);        // gaussian.rs:43-45
            case ILLUME_FILTER_MITCHELL: return mitchell(x * inv_rx) * mitchell(y * inv_ry);  // mitchell.rs:49-51
            default: return lanczos(x, d.radius[0]) * lanczos(y, d.radius[1]);           // sinc.rs:52-54
        }
    }
};

bool valid_desc(const illume_filter_desc &d) {
    if (d.type < ILLUME_FILTER_BOX || d.type > ILLUME_FILTER_LANCZOS) return false;
    for (int i = 0; i < 2; ++i)
        if (!(d.radius[i] > 0.f && d.radius[i] <= 32.f)) return false;
    if (!std::isfinite(d.alpha) || !std::isfinite(d.b) || !std::isfinite(d.c) || !std::isfinite(d.tau)) return false;
    if (d.type == ILLUME_FILTER_LANCZOS && !(d.tau > 0.f)) return false;
    return true;
}

}  // namespace

namespace illume {
// The splat parameters of a filter for a width x height image: its table,
// radius and Film::get_sample_bounds (film.rs:97-103).
void filter_splat(const illume_filter *f, uint32_t width, uint32_t height, FilmSplat *out) {
    out->table = f->d_table;
    out->rx = f->desc.radius[0];
    out->ry = f->desc.radius[1];
    out->irx = 1.f / out->rx;
    out->iry = 1.f / out->ry;
    out->x0 = (int32_t)std::floor(0.f + 0.5f - out->rx);
    out->y0 = (int32_t)std::floor(0.f + 0.5f - out->ry);
    const int32_t x1 = (int32_t)std::ceil((float)width - 0.5f + out->rx);
    const int32_t y1 = (int32_t)std::ceil((float)height - 0.5f + out->ry);
    out->wext = (uint32_t)(x1 - out->x0);
    out->hext = (uint32_t)(y1 - out->y0);
}
}  // namespace illume

extern "C" {

int illume_filter_create(const illume_filter_desc *desc, illume_filter **out) {
    if (!desc || !out) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (!valid_desc(*desc))
        return abi_fail(ILLUME_ERR_INVALID_ARG, "invalid filter (type 0..4, radius in (0, 32], finite parameters, "
                                                "Lanczos tau > 0)");
    illume_filter *f = new (std::nothrow) illume_filter;
    if (!f) return abi_fail(ILLUME_ERR_OUT_OF_MEMORY, "filter");
    f->desc = *desc;
    // Film::new's table (film.rs:69-81)
    const FilterEval eval(*desc);
    for (int y = 0, k = 0; y < kFilterTable; ++y)
        for (int x = 0; x < kFilterTable; ++x, ++k)
            f->table[k] = eval(((float)x + 0.5f) * desc->radius[0] / (float)kFilterTable,
                               ((float)y + 0.5f) * desc->radius[1] / (float)kFilterTable);
    hipError_t e = hipMalloc((void **)&f->d_table, sizeof(f->table));
    if (e == hipSuccess) e = hipMemcpy(f->d_table, f->table, sizeof(f->table), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (f->d_table) (void)hipFree(f->d_table);
        delete f;
        return abi_hip_fail(e, "filter table");
    }
    *out = f;
    return ILLUME_OK;
}

int illume_filter_destroy(illume_filter *filter) {
    if (!filter) return ILLUME_OK;
    if (filter->d_table) (void)hipFree(filter->d_table);
    delete filter;
    return ILLUME_OK;
}

int illume_filter_table(const illume_filter *filter, float *host_table256) {
    if (!filter || !host_table256) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    for (int i = 0; i < kFilterTable * kFilterTable; ++i) host_table256[i] = filter->table[i];
    return ILLUME_OK;
}

int illume_film_resolve(const float *rgbw_sum, uint64_t n_pixels, float scale, float *rgb_out, void *stream) {
    if (n_pixels && (!rgbw_sum || !rgb_out)) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    const hipError_t e = launch_resolve(rgbw_sum, n_pixels, scale, rgb_out, (hipStream_t)stream);
    if (e != hipSuccess) return abi_hip_fail(e, "film resolve launch");
    return ILLUME_OK;
}

}  // extern "C"
