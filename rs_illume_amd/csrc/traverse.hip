// Batched ray queries: closest hit, any hit and the instrumented (counting)
// traversal, one ray per lane over the traversal in trace.h.
#include <hip/hip_runtime.h>
#include "kernels.h"
#include "trace.h"

#pragma clang fp contract(off)

namespace illume {

namespace {

// Launch shape per node width (measured on the 2^20-triangle / 2^24-ray
// workload): two-child nodes run best in 128-lane blocks at the natural
// register count; four-child nodes in 64-lane blocks limited to 64 VGPRs
// (8 waves per SIMD), which keeps more node fetches in flight.
template <int WIDTH>
struct TraverseShape {
    static constexpr int kBlock = 128, kWaves = 1;
};
template <>
struct TraverseShape<4> {
    static constexpr int kBlock = 64, kWaves = 8;
};

template <bool ANY, bool COUNT, int WIDTH, int kBlock = TraverseShape<WIDTH>::kBlock>
__global__ __launch_bounds__(kBlock, TraverseShape<WIDTH>::kWaves) void traverse_kernel(
    const void *__restrict__ nodes, const PrimRecord *__restrict__ prims, const illume_ray *__restrict__ rays,
    uint32_t n, illume_hit *__restrict__ hits, uint8_t *__restrict__ occluded, uint32_t *__restrict__ counts) {
    __shared__ int lds_stack[kStackLds * kBlock];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 *rp = reinterpret_cast<const float4 *>(rays + i);
    const float4 r0 = rp[0], r1 = rp[1];
    int spill[kSpill];
    TraceStack<kBlock> st;
    st.reset(lds_stack + threadIdx.x, spill);
    TraceHit h;
    uint32_t n_box = 0, n_prim = 0;
    const f3 o = mk(r0.x, r0.y, r0.z), d = mk(r0.w, r1.x, r1.y);
    const bool found =
        WIDTH == 4
            ? trace_ray4<ANY, COUNT, kBlock>((const GpuNode4Q *)nodes, prims, o, d, r1.z, h, st, n_box, n_prim)
            : trace_ray<ANY, COUNT, kBlock>((const GpuNode *)nodes, prims, o, d, r1.z, h, st, n_box, n_prim);
    if (COUNT) {
        counts[2 * (size_t)i] = n_box;
        counts[2 * (size_t)i + 1] = n_prim;
    } else if (ANY) {
        occluded[i] = found ? 1 : 0;
    } else {
        reinterpret_cast<float4 *>(hits)[i] = make_float4(__uint_as_float(h.prim), h.t, h.b1, h.b2);
    }
}

__global__ void fill_miss_kernel(uint32_t n, illume_hit *hits, uint8_t *occluded, uint32_t *counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (hits) reinterpret_cast<float4 *>(hits)[i] = make_float4(__uint_as_float(ILLUME_MISS), 0.f, 0.f, 0.f);
    if (occluded) occluded[i] = 0;
    if (counts) {
        counts[2 * (size_t)i] = 0;
        counts[2 * (size_t)i + 1] = 0;
    }
}

}  // namespace

int max_supported_depth() { return kMaxTraceDepth; }

template <int W>
void launch_w(const void *nodes, const PrimRecord *prims, const illume_ray *rays, uint32_t n, illume_hit *hits,
              uint8_t *occluded, uint32_t *counts, bool any, hipStream_t stream) {
    constexpr int kBlock = TraverseShape<W>::kBlock;
    const uint32_t grid = (n + kBlock - 1) / kBlock;
    if (counts) {
        if (any)
            traverse_kernel<true, true, W><<<grid, kBlock, 0, stream>>>(nodes, prims, rays, n, nullptr, nullptr, counts);
        else
            traverse_kernel<false, true, W><<<grid, kBlock, 0, stream>>>(nodes, prims, rays, n, nullptr, nullptr, counts);
    } else if (any) {
        traverse_kernel<true, false, W><<<grid, kBlock, 0, stream>>>(nodes, prims, rays, n, nullptr, occluded, nullptr);
    } else {
        traverse_kernel<false, false, W><<<grid, kBlock, 0, stream>>>(nodes, prims, rays, n, hits, nullptr, nullptr);
    }
}

hipError_t launch_traverse(const void *nodes, int width, const PrimRecord *prims, bool empty, const illume_ray *rays,
                           uint32_t n, illume_hit *hits, uint8_t *occluded, uint32_t *counts, bool any,
                           hipStream_t stream) {
    if (n == 0) return hipSuccess;
    if (empty) {
        fill_miss_kernel<<<(n + 255) / 256, 256, 0, stream>>>(n, counts ? nullptr : (any ? nullptr : hits),
                                                              counts ? nullptr : (any ? occluded : nullptr), counts);
        return hipGetLastError();
    }
    if (width == 4)
        launch_w<4>(nodes, prims, rays, n, hits, occluded, counts, any, stream);
    else
        launch_w<2>(nodes, prims, rays, n, hits, occluded, counts, any, stream);
    return hipGetLastError();
}

}  // namespace illume
