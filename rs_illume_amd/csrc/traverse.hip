// Batched ray queries: closest hit, any hit and the instrumented (counting)
// traversal, one ray per lane over the traversal in trace.h.
#include <hip/hip_runtime.h>
#include "kernels.h"
#include "trace.h"

#pragma clang fp contract(off)

namespace illume {

namespace {

constexpr int kBlock = 256;

template <bool ANY, bool COUNT>
__global__ __launch_bounds__(kBlock) void traverse_kernel(const GpuNode *__restrict__ nodes,
                                                          const PrimRecord *__restrict__ prims,
                                                          const illume_ray *__restrict__ rays, uint32_t n,
                                                          illume_hit *__restrict__ hits, uint8_t *__restrict__ occluded,
                                                          uint32_t *__restrict__ counts) {
    __shared__ int lds_stack[kStackLds * kBlock];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 *rp = reinterpret_cast<const float4 *>(rays + i);
    const float4 r0 = rp[0], r1 = rp[1];
    TraceStack<kBlock> st;
    st.reset(lds_stack + threadIdx.x);
    TraceHit h;
    uint32_t n_box = 0, n_prim = 0;
    const bool found =
        trace_ray<ANY, COUNT, kBlock>(nodes, prims, mk(r0.x, r0.y, r0.z), mk(r0.w, r1.x, r1.y), r1.z, h, st, n_box, n_prim);
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

hipError_t launch_traverse(const GpuNode *nodes, const PrimRecord *prims, bool empty, const illume_ray *rays,
                           uint32_t n, illume_hit *hits, uint8_t *occluded, uint32_t *counts, bool any,
                           hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t grid = (n + kBlock - 1) / kBlock;
    if (empty) {
        fill_miss_kernel<<<grid, kBlock, 0, stream>>>(n, counts ? nullptr : (any ? nullptr : hits),
                                                      counts ? nullptr : (any ? occluded : nullptr), counts);
        return hipGetLastError();
    }
    if (counts) {
        if (any)
            traverse_kernel<true, true><<<grid, kBlock, 0, stream>>>(nodes, prims, rays, n, nullptr, nullptr, counts);
        else
            traverse_kernel<false, true><<<grid, kBlock, 0, stream>>>(nodes, prims, rays, n, nullptr, nullptr, counts);
    } else if (any) {
        traverse_kernel<true, false><<<grid, kBlock, 0, stream>>>(nodes, prims, rays, n, nullptr, occluded, nullptr);
    } else {
        traverse_kernel<false, false><<<grid, kBlock, 0, stream>>>(nodes, prims, rays, n, hits, nullptr, nullptr);
    }
    return hipGetLastError();
}

}  // namespace illume
