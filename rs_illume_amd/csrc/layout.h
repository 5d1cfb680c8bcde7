// Device memory layout of a built BVH, shared by the host builder and the
// HIP kernels.
//
// Inner nodes hold the boxes of BOTH children (64 bytes, one cache-line-sized
// record per visit), so a visit tests two boxes with one 64-byte read and
// never reads the box of a node it does not descend into. Leaves are not
// nodes: a child reference either names an inner node (>= 0) or a contiguous
// run of 1..15 primitive records (< 0).
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace illume {

// 64-byte two-child node. Field order groups the x/y slabs of both children,
// then both z slabs, then the links, so one float4 per slab pair is loaded.
struct alignas(64) GpuNode {
    float c0_lo_x, c0_hi_x, c0_lo_y, c0_hi_y;
    float c1_lo_x, c1_hi_x, c1_lo_y, c1_hi_y;
    float c0_lo_z, c0_hi_z, c1_lo_z, c1_hi_z;
    int32_t child0, child1, pad0, pad1;
};
static_assert(sizeof(GpuNode) == 64, "GpuNode must be 64 bytes");

// 128-byte four-child node, the builder's form (the device gets GpuNode4Q):
// the six slabs of the four children as float4 rows (lo_x, hi_x, lo_y, hi_y,
// lo_z, hi_z), then the four child references. Unused child slots have all
// planes at +inf and an empty leaf reference. Half the depth of the binary
// tree: half the dependent node fetches per ray.
struct alignas(128) GpuNode4 {
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    int32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(GpuNode4) == 128, "GpuNode4 must be 128 bytes");

// 64-byte four-child node with quantized child boxes (one cache line per
// visit instead of two). A child plane is fma(q, 2^(e - 127), origin) for a
// byte q, evaluated in float32 (one rounding, the product being exact);
// the builder picks q so that every plane evaluated this way lies on or
// outside the child's float box, so the dequantized box contains it and the
// slab test on it enters every box the unquantized test enters. Unused slots
// hold kEmptyChild and are never entered.
struct alignas(64) GpuNode4Q {
    float ox, oy, oz;
    uint8_t ex, ey, ez, pad0;  // biased exponents of the three quantization steps
    uint8_t lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    int32_t child[4];
    uint32_t pad1[2];
};
static_assert(sizeof(GpuNode4Q) == 64, "GpuNode4Q must be 64 bytes");
constexpr int32_t kEmptyChild = -1;  // ~(0 << 4 | 0): a leaf of no primitives

// Leaf reference: ~((first_slot << 4) | count), count in 1..15.
constexpr int kLeafCountBits = 4;
constexpr int kMaxLeafPrims = (1 << kLeafCountBits) - 1;
constexpr uint32_t kMaxPrimSlots = 1u << 27;
__host__ __device__ inline int32_t make_leaf_ref(uint32_t first_slot, uint32_t count) {
    return (int32_t)~((first_slot << kLeafCountBits) | count);
}
__host__ __device__ inline uint32_t leaf_first(int32_t ref) { return ((uint32_t)~ref) >> kLeafCountBits; }
__host__ __device__ inline uint32_t leaf_count(int32_t ref) { return ((uint32_t)~ref) & kMaxLeafPrims; }

// Primitive record: 48 bytes = 3 x float4.
//   triangle: (p0.xyz, id) (p1.xyz, 0) (p2.xyz, 0)   vertices in mesh index order
//   sphere:   (center.xyz, id | kSphereBit) (radius, 0, 0, 0) (0, 0, 0, 0)
// id is the caller's primitive id stored as raw bits.
constexpr uint32_t kSphereBit = 0x80000000u;
struct PrimRecord {
    float a[4], b[4], c[4];
};
static_assert(sizeof(PrimRecord) == 48, "PrimRecord must be 48 bytes");

}  // namespace illume
