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
// to the caller's float RGB sum buffer with float atomics, or, through a
// reconstruction filter, writes its sample for the gather pass that films it;
// ray counts are reduced per block in LDS before one 64-bit atomic per block.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "film.h"
#include "kernels.h"
#include "path.h"

#pragma clang fp contract(off)

namespace illume {

namespace {

constexpr int kRenderBlock = 128;

// FILTERED: camera samples over the sample bounds of `fs` (stream = index in
// those bounds), each written to the sample buffer `film` (radiance and film
// position at its path number) for gather_kernel; otherwise one sample pixel
// per image pixel, added to 3-float pixels.
// At least 3 waves per SIMD (168 VGPRs, one or two spilled): the traversals
// inside a bounce hide their node fetches better than at the 2 waves the
// natural 182 VGPRs give (Cornell box 256^2 x 16 spp 1.31 -> 1.04 ms, a 256^2
// terrain view 1.62 -> 1.45 ms); 4 waves spill 40+ VGPRs and lose on the terrain.
template <int BLOCK, bool FILTERED>
__global__ __launch_bounds__(BLOCK, 3) void render_kernel(DevScene sc, DevCamera cam, uint32_t spp, uint32_t max_depth,
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
            // the gather pass films it (film = the sample buffer: [gid] (L, fx), then [n + gid] fy)
            reinterpret_cast<float4 *>(film)[gid] = make_float4(L.x, L.y, L.z, fx);
            film[4 * n_pix * spp + gid] = fy;
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

// SamplerIntegratorBase::render with AOIntegrator::li (integrators/ao.rs:45-93):
// one thread per camera sample. At the closest hit the frame is n = the
// geometric normal facing the ray's origin, s = normalize(dpdu), t = n_g x s;
// each of n_ao hemisphere directions (cosine or uniform, two sampler
// dimensions each after the camera sample) adds (wi . n) / (pdf n_ao) when its
// shadow ray (an any-hit query, t_max infinite) is unoccluded. The grey value
// goes to all three channels of the pixel's sum.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK, 4) void render_ao_kernel(DevScene sc, DevCamera cam, uint32_t spp, uint32_t n_ao,
                                                           bool cos_sample, uint64_t seed, uint32_t sample_offset,
                                                           float *__restrict__ film,
                                                           unsigned long long *__restrict__ counters,
                                                           SamplerLaunch hl) {
    __shared__ int lds_stack[kStackLds * BLOCK];
    __shared__ uint32_t blk_counts[2];
    if (threadIdx.x < 2) blk_counts[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t n_pix = (uint64_t)cam.width * cam.height;
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
        path_start_at(cam, seed, pix, (int)(pix % cam.width), (int)(pix / cam.width), s, ps, fx, fy, hl);
        TraceHit h;
        float ao = 0.f;
        if (scene_closest<BLOCK>(sc, ps.o, ps.d, ps.t_max, h, st, pc)) {
            Isect it;
            hit_isect(sc, h, ps.o, ps.d, it);
            const f3 n = faceforward(it.n, neg(ps.d));
            const f3 sv = normalize(it.dpdu);
            const f3 tv = cross(it.n, sv);
            for (uint32_t i = 0; i < n_ao; ++i) {
                const float u0 = ps.rng.get_1d(), u1 = ps.rng.get_1d();
                f3 w;
                float pdf;
                if (cos_sample) {
                    w = cosine_sample_hemisphere(u0, u1);
                    pdf = fabsf(w.z) * kInvPi;
                } else {
                    w = uniform_sample_hemisphere(u0, u1);
                    pdf = kInv2Pi;
                }
                const f3 wi = mk(sv.x * w.x + tv.x * w.y + n.x * w.z, sv.y * w.x + tv.y * w.y + n.y * w.z,
                                 sv.z * w.x + tv.z * w.y + n.z * w.z);
                f3 so;
                spawn_ray(it, wi, so);
                if (!scene_any<BLOCK>(sc, so, wi, __builtin_inff(), st, pc)) ao += dot(wi, n) / (pdf * (float)n_ao);
            }
        }
        const f3 L = guard_radiance(mk(ao, ao, ao));
        atomicAdd(film + 3 * pix, L.x);
        atomicAdd(film + 3 * pix + 1, L.y);
        atomicAdd(film + 3 * pix + 2, L.z);
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

// SamplerIntegrator::li for caller-given rays (one thread per ray): ray i
// takes the sampler stream (seed, i, sample), skips the five dimensions of a
// camera sample and runs PathIntegrator::li from the ray, so the camera ray of
// pixel p for sample s gives illume_render's sample s of pixel p. L_out gets
// the guarded radiance (3 floats per ray).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK, 3) void li_batch_kernel(DevScene sc, const illume_ray *__restrict__ rays, uint32_t n,
                                                          uint32_t max_depth, uint64_t seed, uint32_t sample,
                                                          float *__restrict__ L_out,
                                                          unsigned long long *__restrict__ counters) {
    __shared__ int lds_stack[kStackLds * BLOCK];
    __shared__ uint32_t blk_counts[2];
    if (threadIdx.x < 2) blk_counts[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    PathCounters pc{0, 0};
    if (i < n) {
        int spill[kSpill];
        TraceStack<BLOCK> st;
        st.reset(lds_stack + threadIdx.x, spill);
        const float4 *rp = reinterpret_cast<const float4 *>(rays + i);
        const float4 r0 = rp[0], r1 = rp[1];
        PathState ps;
        ps.rng.seed(seed, i, sample);
        for (int k = 0; k < 5; ++k) ps.rng.get_1d();  // film 2D, time 1D, lens 2D
        ps.o = mk(r0.x, r0.y, r0.z);
        ps.d = mk(r0.w, r1.x, r1.y);
        ps.t_max = r1.z;
        ps.beta = mk(1.f, 1.f, 1.f);
        ps.L = mk(0.f, 0.f, 0.f);
        ps.eta_scale = 1.f;
        ps.bounces = 0;
        ps.specular = false;
        ps.active = true;
        GuideCfg gc{};
        uint32_t n_recs = 0;
        while (ps.active) {
            TraceHit h;
            const bool found = scene_closest<BLOCK>(sc, ps.o, ps.d, ps.t_max, h, st, pc);
            path_shade<BLOCK>(sc, max_depth, ps, found, h, nullptr, gc, nullptr, n_recs, st, pc);
        }
        const f3 L = guard_radiance(ps.L);
        L_out[3 * (size_t)i] = L.x;
        L_out[3 * (size_t)i + 1] = L.y;
        L_out[3 * (size_t)i + 2] = L.z;
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

// Batched megakernel: one thread per path of the whole batch; lanes find
// their job (scene, camera, film offset, sampler) by binary search over the
// jobs' first path numbers, so one launch renders images of any sizes.
// Held to at least 5 waves per SIMD (96 VGPRs; the spills stay in cache):
// against 3 waves, 32 procedural scenes render 18% and 16 terrain views 17%
// faster (6 waves equal, 8 slower).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK, 5) void render_batch_kernel(const RenderJob *__restrict__ jobs, uint32_t n_jobs,
                                                                 unsigned long long total, float *__restrict__ film,
                                                                 unsigned long long *__restrict__ counters) {
    __shared__ int lds_stack[kStackLds * BLOCK];
    __shared__ uint32_t blk_counts[2];
    if (threadIdx.x < 2) blk_counts[threadIdx.x] = 0;
    __syncthreads();
    const unsigned long long g = (unsigned long long)blockIdx.x * BLOCK + threadIdx.x;
    PathCounters pc{0, 0};
    if (g < total) {
        uint32_t lo = 0, hi = n_jobs;  // last job with path_begin <= g
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (jobs[mid].path_begin <= g) lo = mid;
            else hi = mid;
        }
        const RenderJob &job = jobs[lo];
        const DevScene &sc = *job.scene;
        const uint64_t n_pix = (uint64_t)job.cam.width * job.cam.height;
        const unsigned long long local = g - job.path_begin;
        const uint64_t pix = local % n_pix;
        int spill[kSpill];
        TraceStack<BLOCK> st;
        st.reset(lds_stack + threadIdx.x, spill);
        PathState ps;
        float fx, fy;
        path_start_at(job.cam, job.seed, pix, (int)(pix % job.cam.width), (int)(pix / job.cam.width),
                      job.sample_offset + (uint32_t)(local / n_pix), ps, fx, fy, job.sl);
        GuideCfg gc{};
        uint32_t n_recs = 0;
        while (ps.active) {
            TraceHit h;
            const bool found = scene_closest<BLOCK>(sc, ps.o, ps.d, ps.t_max, h, st, pc);
            path_shade<BLOCK>(sc, job.max_depth, ps, found, h, nullptr, gc, nullptr, n_recs, st, pc);
        }
        const f3 L = guard_radiance(ps.L);
        float *px = film + job.film_offset + 3 * pix;
        atomicAdd(px, L.x);
        atomicAdd(px + 1, L.y);
        atomicAdd(px + 2, L.z);
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

// The filtered film of a chunk of spp samples per sample pixel, by gathering:
// one lane per image pixel visits, in a fixed order (sample, then sample-pixel
// rows and columns), every sample whose sample pixel lies within the filter
// radius (plus one pixel of margin), applies film_splat's footprint test and
// weight to it with the same float operations, and adds the sums to its RGBW
// pixel once. No atomics: the film is the same bits on every run.
__global__ __launch_bounds__(256) void gather_kernel(FilmSplat fs, const float4 *__restrict__ smp_l,
                                                     const float *__restrict__ smp_y, uint32_t spp, uint32_t width,
                                                     uint32_t height, float *__restrict__ rgbw) {
    const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= width || y >= height) return;
    const uint64_t n_pix = (uint64_t)fs.wext * fs.hext;
    const int sx0 = max((int)floorf((float)x + 0.5f - fs.rx) - 1, fs.x0);
    const int sx1 = min((int)floorf((float)x + 0.5f + fs.rx) + 1, fs.x0 + (int)fs.wext - 1);
    const int sy0 = max((int)floorf((float)y + 0.5f - fs.ry) - 1, fs.y0);
    const int sy1 = min((int)floorf((float)y + 0.5f + fs.ry) + 1, fs.y0 + (int)fs.hext - 1);
    float r = 0.f, g = 0.f, b = 0.f, wsum = 0.f;
    for (uint32_t s = 0; s < spp; ++s) {
        for (int sy = sy0; sy <= sy1; ++sy) {
            const uint64_t row = (uint64_t)s * n_pix + (uint64_t)(sy - fs.y0) * fs.wext;
            for (int sx = sx0; sx <= sx1; ++sx) {
                const uint64_t i = row + (uint64_t)(sx - fs.x0);
                const float4 a = smp_l[i];
                const float dx = a.w - 0.5f, dy = smp_y[i] - 0.5f;
                const int x0 = max((int)ceilf(dx - fs.rx), 0), x1 = min((int)floorf(dx + fs.rx) + 1, (int)width);
                const int y0 = max((int)ceilf(dy - fs.ry), 0), y1 = min((int)floorf(dy + fs.ry) + 1, (int)height);
                if ((int)x < x0 || (int)x >= x1 || (int)y < y0 || (int)y >= y1) continue;
                const float fy = fabsf(((float)y - dy) * fs.iry * (float)kFilterTable);
                const int iy = min(kFilterTable - 1, (int)floorf(fy));
                const float fx = fabsf(((float)x - dx) * fs.irx * (float)kFilterTable);
                const int ix = min(kFilterTable - 1, (int)floorf(fx));
                const float w = fs.table[iy * kFilterTable + ix];
                r += a.x * w;
                g += a.y * w;
                b += a.z * w;
                wsum += w;
            }
        }
    }
    float *p = rgbw + 4 * ((size_t)y * width + x);
    p[0] += r;
    p[1] += g;
    p[2] += b;
    p[3] += wsum;
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

hipError_t launch_render_ao(const DevScene &sc, const DevCamera &cam, uint32_t spp, uint32_t n_ao, bool cos_sample,
                            uint64_t seed, uint32_t sample_offset, float *rgb, unsigned long long *counters,
                            const SamplerLaunch &hl, hipStream_t stream) {
    const uint64_t n = (uint64_t)cam.width * cam.height * spp;
    if (n == 0) return hipSuccess;
    const uint64_t grid = (n + kRenderBlock - 1) / kRenderBlock;
    if (grid > 0x7fffffffULL) return hipErrorInvalidValue;
    render_ao_kernel<kRenderBlock><<<(uint32_t)grid, kRenderBlock, 0, stream>>>(sc, cam, spp, n_ao, cos_sample, seed,
                                                                                sample_offset, rgb, counters, hl);
    return hipGetLastError();
}

__global__ void zero_floats_kernel(float *__restrict__ p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

hipError_t launch_li_batch(const DevScene &sc, const illume_ray *rays, uint32_t n, uint32_t max_depth, uint64_t seed,
                           uint32_t sample, float *L_out, unsigned long long *counters, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    if (!sc.nodes) {  // nothing to hit, no lights: black
        const uint64_t m = 3ull * n;
        zero_floats_kernel<<<(uint32_t)((m + 255) / 256), 256, 0, stream>>>(L_out, m);
        return hipGetLastError();
    }
    li_batch_kernel<kRenderBlock><<<(n + kRenderBlock - 1) / kRenderBlock, kRenderBlock, 0, stream>>>(
        sc, rays, n, max_depth, seed, sample, L_out, counters);
    return hipGetLastError();
}

hipError_t launch_render_filtered(const DevScene &sc, const DevCamera &cam, const FilmSplat &fs, uint32_t spp,
                                  uint32_t max_depth, uint64_t seed, uint32_t sample_offset, float *rgbw,
                                  unsigned long long *counters, const SamplerLaunch &hl, hipStream_t stream) {
    const uint64_t n_pix = (uint64_t)fs.wext * fs.hext;
    if (n_pix == 0 || spp == 0) return hipSuccess;
    // chunks of at most 2^24 samples (20 bytes each) unless one sample per pixel is more
    constexpr uint64_t kChunkSamples = 1ull << 24;
    const uint32_t chunk = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(spp, kChunkSamples / n_pix));
    const uint64_t cap = n_pix * chunk;
    if ((cap + kRenderBlock - 1) / kRenderBlock > 0x7fffffffULL) return hipErrorInvalidValue;
    float *buf = nullptr;
    hipError_t e = hipMallocAsync((void **)&buf, cap * 5 * sizeof(float), stream);
    if (e != hipSuccess) return e;
    const dim3 ggrid((cam.width + 15) / 16, (cam.height + 15) / 16);
    for (uint32_t s0 = 0; s0 < spp && e == hipSuccess; s0 += chunk) {
        const uint32_t k = std::min(chunk, spp - s0);
        const uint64_t n = n_pix * k;
        render_kernel<kRenderBlock, true><<<(uint32_t)((n + kRenderBlock - 1) / kRenderBlock), kRenderBlock, 0,
                                           stream>>>(sc, cam, k, max_depth, seed, sample_offset + s0, buf, counters,
                                                     fs, hl);
        gather_kernel<<<ggrid, 256, 0, stream>>>(fs, reinterpret_cast<const float4 *>(buf), buf + 4 * n, k,
                                                 cam.width, cam.height, rgbw);
        e = hipGetLastError();
    }
    const hipError_t ef = hipFreeAsync(buf, stream);
    return e != hipSuccess ? e : ef;
}

hipError_t launch_render_batch(const RenderJob *jobs, uint32_t n_jobs, unsigned long long total, float *film,
                               unsigned long long *counters, hipStream_t stream) {
    if (total == 0 || n_jobs == 0) return hipSuccess;
    const unsigned long long grid = (total + kRenderBlock - 1) / kRenderBlock;
    if (grid > 0x7fffffffULL) return hipErrorInvalidValue;
    render_batch_kernel<kRenderBlock><<<(uint32_t)grid, kRenderBlock, 0, stream>>>(jobs, n_jobs, total, film,
                                                                                   counters);
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
