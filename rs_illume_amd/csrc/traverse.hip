// BVH traversal kernels: closest hit (BVHAccel::intersect, accelerators/
// bvh.rs:563-612) and any hit (BVHAccel::intersect_p, bvh.rs:614-662) over the
// 64-byte two-child layout.
//
// One ray per lane. Each visit of an inner node reads both child boxes (one
// 64-byte record), tests them with the conservative slab test, descends into
// the nearer hit child and pushes the farther. The traversal stack lives in
// LDS (kStackLds entries per lane, a ring laid out [slot][lane] so a wave's
// pushes and pops hit 64 distinct banks); entries deeper than that spill to a
// private array that is touched only by rays whose stack outgrows the ring.
// Primitive tests are the reference's FP32 arithmetic (geom.h), so hit ids and
// distances match it exactly.
#include <hip/hip_runtime.h>
#include "geom.h"
#include "kernels.h"
#include "layout.h"

#pragma clang fp contract(off)

namespace illume {

namespace {

constexpr int kBlock = 256;
constexpr int kStackLds = 16;  // power of two
constexpr int kSpill = 64;

struct Stack {
    int *lds;  // this lane's column: lds[slot * kBlock]
    int sp;
    int spill[kSpill];
    __device__ __forceinline__ void push(int v) {
        int slot = sp & (kStackLds - 1);
        if (sp >= kStackLds) spill[sp - kStackLds] = lds[slot * kBlock];
        lds[slot * kBlock] = v;
        ++sp;
    }
    __device__ __forceinline__ int pop() {
        --sp;
        int slot = sp & (kStackLds - 1);
        int v = lds[slot * kBlock];
        if (sp >= kStackLds) lds[slot * kBlock] = spill[sp - kStackLds];
        return v;
    }
};

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
    const f3 o = mk(r0.x, r0.y, r0.z);
    const f3 d = mk(r0.w, r1.x, r1.y);
    const f3 inv = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
    float t_max = r1.z;

    uint32_t hit_id = ILLUME_MISS;
    float hit_t = 0.f, hit_b1 = 0.f, hit_b2 = 0.f;
    bool occl = false;
    uint32_t n_box = 0, n_prim = 0;

    Stack st;
    st.lds = lds_stack + threadIdx.x;
    st.sp = 0;
    int cur = 0;
    while (true) {
        if (cur >= 0) {
            const float4 *np = reinterpret_cast<const float4 *>(nodes + cur);
            const float4 a = np[0], b = np[1], c = np[2];
            const int4 l = reinterpret_cast<const int4 *>(np)[3];
            if (COUNT) n_box += 2;
            float t0, t1;
            const bool h0 = slab(a.x, a.y, a.z, a.w, c.x, c.y, o, inv, t_max, t0);
            const bool h1 = slab(b.x, b.y, b.z, b.w, c.z, c.w, o, inv, t_max, t1);
            if (h0 && h1) {
                int near = l.x, far = l.y;
                if (t1 < t0) {
                    near = l.y;
                    far = l.x;
                }
                st.push(far);
                cur = near;
                continue;
            }
            if (h0) {
                cur = l.x;
                continue;
            }
            if (h1) {
                cur = l.y;
                continue;
            }
        } else {
            const uint32_t first = leaf_first(cur), cnt = leaf_count(cur);
            for (uint32_t k = 0; k < cnt; ++k) {
                const float4 *pp = reinterpret_cast<const float4 *>(prims + first + k);
                const float4 pa = pp[0];
                const uint32_t tag = __float_as_uint(pa.w);
                if (COUNT) ++n_prim;
                if (tag & kSphereBit) {
                    const float4 pb = pp[1];
                    float t;
                    if (sphere_test(mk(pa.x, pa.y, pa.z), pb.x, o, d, t_max, t)) {
                        if (ANY) {
                            occl = true;
                            break;
                        }
                        t_max = t;
                        hit_id = tag & ~kSphereBit;
                        hit_t = t;
                        hit_b1 = 0.f;
                        hit_b2 = 0.f;
                    }
                } else {
                    const float4 pb = pp[1], pc = pp[2];
                    const f3 p0 = mk(pa.x, pa.y, pa.z), p1 = mk(pb.x, pb.y, pb.z), p2 = mk(pc.x, pc.y, pc.z);
                    float t, b0, b1, b2;
                    if (triangle_test(p0, p1, p2, o, d, t_max, t, b0, b1, b2)) {
                        if (ANY) {
                            occl = true;
                            break;
                        }
                        if (!triangle_degenerate(p0, p1, p2)) {
                            t_max = t;
                            hit_id = tag;
                            hit_t = t;
                            hit_b1 = b1;
                            hit_b2 = b2;
                        }
                    }
                }
            }
            if (ANY && occl) break;
        }
        if (st.sp == 0) break;
        cur = st.pop();
    }
    if (COUNT) {
        counts[2 * (size_t)i] = n_box;
        counts[2 * (size_t)i + 1] = n_prim;
    } else if (ANY) {
        occluded[i] = occl ? 1 : 0;
    } else {
        reinterpret_cast<float4 *>(hits)[i] = make_float4(__uint_as_float(hit_id), hit_t, hit_b1, hit_b2);
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

int max_supported_depth() { return kStackLds + kSpill; }

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
