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
#include "kernels.h"
#include "path.h"

#pragma clang fp contract(off)

namespace illume {

namespace {

constexpr int kRenderBlock = 128;

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void render_kernel(DevScene sc, DevCamera cam, uint32_t spp, uint32_t max_depth,
                                                        uint64_t seed, uint32_t sample_offset, float *__restrict__ rgb,
                                                        unsigned long long *__restrict__ counters) {
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
        TraceStack<BLOCK> st;
        st.reset(lds_stack + threadIdx.x);
        PathState ps;
        path_start(cam, seed, pix, s, ps);
        GuideCfg gc{};
        uint32_t n_recs = 0;
        while (ps.active) {
            TraceHit h;
            const bool found = scene_closest<BLOCK>(sc, ps.o, ps.d, ps.t_max, h, st, pc);
            path_shade<BLOCK>(sc, max_depth, ps, found, h, nullptr, gc, nullptr, n_recs, st, pc);
        }
        const f3 L = guard_radiance(ps.L);
        atomicAdd(rgb + 3 * pix, L.x);
        atomicAdd(rgb + 3 * pix + 1, L.y);
        atomicAdd(rgb + 3 * pix + 2, L.z);
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

}  // namespace

hipError_t launch_render(const DevScene &sc, const DevCamera &cam, uint32_t spp, uint32_t max_depth, uint64_t seed,
                         uint32_t sample_offset, float *rgb, unsigned long long *counters, hipStream_t stream) {
    const uint64_t n = (uint64_t)cam.width * cam.height * spp;
    if (n == 0) return hipSuccess;
    const uint64_t grid = (n + kRenderBlock - 1) / kRenderBlock;
    if (grid > 0x7fffffffULL) return hipErrorInvalidValue;
    render_kernel<kRenderBlock><<<(uint32_t)grid, kRenderBlock, 0, stream>>>(sc, cam, spp, max_depth, seed,
                                                                             sample_offset, rgb, counters);
    return hipGetLastError();
}

}  // namespace illume
