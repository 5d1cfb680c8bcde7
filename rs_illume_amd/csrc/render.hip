// Megakernel path tracer: pbrt-v3's PathIntegrator (the reference has the
// SamplerIntegrator trait only, core/integrator.rs:31-35) with the
// reference's estimate_direct / uniform_sample_one_light MIS
// (integrator.rs:321-494), camera sampling (sampler.rs:78-84,
// perspective.rs:103-126) and the radiance guard of
// SamplerIntegratorBase::render (integrator.rs:113-122).
//
// One thread traces one camera path through all of its bounces (closest-hit
// extension rays, any-hit shadow rays and MIS rays share the traversal of
// trace.h). Lanes of a wave hold neighbouring pixels of the same sample
// index, so primary rays are coherent. Each finished path adds its radiance
// to the caller's float RGB sum buffer with float atomics; ray counts are
// reduced per block in LDS before one 64-bit atomic per block.
#include <hip/hip_runtime.h>
#include "film.h"
#include "kernels.h"
#include "path.h"

#pragma clang fp contract(off)

namespace illume {

namespace {

constexpr int kRenderBlock = 128;

// FILTERED: camera samples over the sample bounds of `fs` (stream = index in
// those bounds) splatted through the filter into 4-float pixels; otherwise
// one sample pixel per image pixel, added to 3-float pixels.
template <int BLOCK, bool FILTERED>
__global__ __launch_bounds__(BLOCK) void render_kernel(DevScene sc, DevCamera cam, uint32_t spp, uint32_t max_depth,
                                                        uint64_t seed, uint32_t sample_offset, float *__restrict__ film,
                                                        unsigned long long *__restrict__ counters, FilmSplat fs,
                                                        SamplerLaunch hl) {
    __shared__ int lds_stack[kStackLds * BLOCK];
    __shared__ uint32_t blk_counts[2];
    if (threadIdx.x < 2) blk_counts[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t n_pix = FILTERED ? (uint64_t)fs.wext * fs.hext : (uint64_t)cam.width * cam.height;
    const uint64_t gid = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    PathCounters pc{0, 0};
    if (gid < n_pix * spp) {
        const uint64_t pix = gid % n_pix;
        const uint32_t s = (uint32_t)(gid / n_pix) + sample_offset;
        int spill[kSpill];
        TraceStack<BLOCK> st;
        st.reset(lds_stack + threadIdx.x, spill);
        PathState ps;
        float fx, fy;
        if (FILTERED)
            path_start_at(cam, seed, pix, fs.x0 + (int)(pix % fs.wext), fs.y0 + (int)(pix / fs.wext), s, ps, fx, fy,
                          hl);
        else
            path_start_at(cam, seed, pix, (int)(pix % cam.width), (int)(pix / cam.width), s, ps, fx, fy, hl);
        GuideCfg gc{};
        uint32_t n_recs = 0;
        while (ps.active) {
            TraceHit h;
            const bool found = scene_closest<BLOCK>(sc, ps.o, ps.d, ps.t_max, h, st, pc);
            path_shade<BLOCK>(sc, max_depth, ps, found, h, nullptr, gc, nullptr, n_recs, st, pc);
        }
        const f3 L = guard_radiance(ps.L);
        if (FILTERED) {
            film_splat(fs, film, (int)cam.width, (int)cam.height, fx, fy, L.x, L.y, L.z);
        } else {
            atomicAdd(film + 3 * pix, L.x);
            atomicAdd(film + 3 * pix + 1, L.y);
            atomicAdd(film + 3 * pix + 2, L.z);
        }
    }
    if (counters) {
        atomicAdd(&blk_counts[0], pc.closest);
        atomicAdd(&blk_counts[1], pc.any);
        __syncthreads();
        if (threadIdx.x == 0) {
            atomicAdd(counters, (unsigned long long)blk_counts[0]);
            atomicAdd(counters + 1, (unsigned long long)blk_counts[1]);
        }
    }
}

// Film::write_image's per-pixel normalisation (film.rs:187-226).
__global__ __launch_bounds__(256) void resolve_kernel(const float *__restrict__ rgbw, uint64_t n, float scale,
                                                      float *__restrict__ rgb) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float r = rgbw[4 * i], g = rgbw[4 * i + 1], b = rgbw[4 * i + 2];
    const float w = rgbw[4 * i + 3];
    if (w != 0.f) {
        const float inv = 1.f / w;
        r = fmaxf(r * inv, 0.f);
        g = fmaxf(g * inv, 0.f);
        b = fmaxf(b * inv, 0.f);
    }
    rgb[3 * i] = r * scale;
    rgb[3 * i + 1] = g * scale;
    rgb[3 * i + 2] = b * scale;
}

}  // namespace

hipError_t launch_render(const DevScene &sc, const DevCamera &cam, uint32_t spp, uint32_t max_depth, uint64_t seed,
                         uint32_t sample_offset, float *rgb, unsigned long long *counters, const SamplerLaunch &hl,
                         hipStream_t stream) {
    const uint64_t n = (uint64_t)cam.width * cam.height * spp;
    if (n == 0) return hipSuccess;
    const uint64_t grid = (n + kRenderBlock - 1) / kRenderBlock;
    if (grid > 0x7fffffffULL) return hipErrorInvalidValue;
    render_kernel<kRenderBlock, false><<<(uint32_t)grid, kRenderBlock, 0, stream>>>(
        sc, cam, spp, max_depth, seed, sample_offset, rgb, counters, FilmSplat{}, hl);
    return hipGetLastError();
}

hipError_t launch_render_filtered(const DevScene &sc, const DevCamera &cam, const FilmSplat &fs, uint32_t spp,
                                  uint32_t max_depth, uint64_t seed, uint32_t sample_offset, float *rgbw,
                                  unsigned long long *counters, const SamplerLaunch &hl, hipStream_t stream) {
    const uint64_t n = (uint64_t)fs.wext * fs.hext * spp;
    if (n == 0) return hipSuccess;
    const uint64_t grid = (n + kRenderBlock - 1) / kRenderBlock;
    if (grid > 0x7fffffffULL) return hipErrorInvalidValue;
    render_kernel<kRenderBlock, true><<<(uint32_t)grid, kRenderBlock, 0, stream>>>(
        sc, cam, spp, max_depth, seed, sample_offset, rgbw, counters, fs, hl);
    return hipGetLastError();
}

hipError_t launch_resolve(const float *rgbw, uint64_t n, float scale, float *rgb, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint64_t grid = (n + 255) / 256;
    if (grid > 0x7fffffffULL) return hipErrorInvalidValue;
    resolve_kernel<<<(uint32_t)grid, 256, 0, stream>>>(rgbw, n, scale, rgb);
    return hipGetLastError();
}

}  // namespace illume
