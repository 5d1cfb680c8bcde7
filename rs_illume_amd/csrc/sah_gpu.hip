// Binned SAH BVH built on the device: BVHAccel::new with SplitMethod::SAH
// (accelerators/bvh.rs:32-247, recursive_build 95-247) under the host
// builder's split rule (bvh_builder.h: kSahBins bins on each axis, split cost
// kSahTraversalCost + (n0 A0 + n1 A1) / A against leaf cost n, count splits
// for coincident centroids, object medians from depth kSahMedianDepth on,
// stable partitions). Bins hold exact min/max boxes and integer counts, so
// every split decision is the host's, bit for bit, and the tree is the host
// tree: the same inner nodes with the same boxes, the same leaves with their
// primitives in the same order. The one exception is an object median of a
// node of more than kSmall primitives below depth kSahMedianDepth, split here
// by count in the current order (the host selects by centroid).
//
//   1. primitive boxes (one thread per primitive), the root's box and
//      centroid bounds;
//   2. top phase, one pass per level over the nodes of more than kSmall
//      primitives. Every node's range is cut into chunks of at least kChunk
//      primitives (at most kMaxChunks per node), one block per chunk: binning
//      into LDS bins (each bin: count, box, centroid bounds), flushed once per
//      block with global atomics; one wave per node evaluates the 3 x 15
//      candidate splits and takes both children's boxes and centroid bounds
//      from the bins; a flag per primitive, a device-wide exclusive scan, and a
//      stable scatter of each node's range into the other buffer;
//   3. bottom phase: one wave per node of at most kSmall primitives builds
//      its whole subtree with the primitives in LDS (binning with LDS
//      atomics, partition by ballot), depth first, children's boxes from the
//      bins as above;
//   4. leaf-ordered primitive records; for width 4, the collapse (open the
//      inner child of largest area until four) and quantization of each
//      four-child node on the device, one launch per four-child level.
// A node is written by its own step into its parent's two-child node (its box
// and its reference), so no pass waits for a whole level of boxes.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_scan.hpp>
#include <algorithm>
#include <cstring>
#include <vector>
#include "bvh_builder.h"
#include "layout.h"
#include "lbvh.h"
#include "quantize.h"

namespace illume {
namespace {

constexpr int kSmall = 128;        // nodes of at most this many primitives: one wave builds the subtree
constexpr int kStack = 48;         // pending nodes of a bottom-phase subtree (depth bound: 40 + 7 + 1)
constexpr uint32_t kChunk = 2048;  // primitives per block of a top-phase pass, at least
constexpr uint32_t kMaxChunks = 256;  // blocks per node, at most (bounds the atomics on one bin)
constexpr int kBinSlots = 3 * kSahBins;
constexpr int kBinWords = 13;   // count, box lo xyz, box hi xyz, centroid lo xyz, centroid hi xyz (encoded)
constexpr int kGBinWords = 16;  // the same, padded, in global memory
constexpr int kMaxTopLevels = 256;

// Order-preserving float <-> uint32 encoding (min/max on the encodings are
// min/max on the floats).
__device__ __forceinline__ uint32_t enc(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float dec(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}
constexpr uint32_t kEncPosInf = 0xff800000u;  // enc(+inf)
constexpr uint32_t kEncNegInf = 0x007fffffu;  // enc(-inf)

struct PBox {
    float4 lo, hi;  // lo.w: the primitive's id (raw bits)
};

// A node of the top phase or the root of a bottom-phase subtree.
struct BNode {
    uint32_t box[6];     // encoded lo xyz, hi xyz
    uint32_t cbox[6];    // encoded centroid bounds
    uint32_t start, count;
    int32_t parent;      // two-child node slot of the parent, -1 for the root
    uint32_t depth_side; // depth << 2 | buffer holding the range << 1 | which child of the parent
};
static_assert(sizeof(BNode) == 64, "BNode is 64 bytes");

struct Split {          // decision for a top-phase node
    uint32_t mode;      // 0: by bin along axis, 1: first lcnt in the current order
    uint32_t axis, bin;
    float k, lo;        // bin = clamp((int)((c - lo) * k), 0, kSahBins - 1)
    uint32_t start, lcnt;
    uint32_t child[2];  // BNode ids
    uint32_t pad[3];
};

// Device counters. Two-child node slots are handed out in blocks (`alloc`):
// one per top-phase node, m - 1 per bottom-phase subtree of m primitives (its
// inner nodes at most), n - 1 in all; `inner` counts the slots used, `used`
// marks them.
struct Ctr {
    uint32_t inner, bnodes, next_front, small, leaves, depth, overflow, n4, alloc, pad[3];
};

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ uint32_t lanes_below(uint64_t m, int lane) {
    return (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}
__device__ __forceinline__ float pick3(int a, float x, float y, float z) { return a == 0 ? x : (a == 1 ? y : z); }
__device__ __forceinline__ float bcast(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ float centroid(const PBox &p, int a) {
    return pick3(a, p.lo.x, p.lo.y, p.lo.z) * 0.5f + pick3(a, p.hi.x, p.hi.y, p.hi.z) * 0.5f;
}
__device__ __forceinline__ int bin_of(float c, float lo, float k) {
    const int b = (int)((c - lo) * k);
    return min(max(b, 0), kSahBins - 1);
}
// Number of blocks a top-phase pass gives a node of `count` primitives.
__host__ __device__ __forceinline__ uint32_t chunks_of(uint32_t count) {
    const uint32_t c = (count + kChunk - 1) / kChunk;
    return c < 1 ? 1 : (c > kMaxChunks ? kMaxChunks : c);
}

// Adds primitive p to the bins of axis a (bin b) at `bins` (kBinWords
// encoded words per bin, `stride` apart).
__device__ __forceinline__ void bin_add(uint32_t *dst, const uint32_t *e) {
    atomicAdd(dst, 1u);
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        atomicMin(dst + 1 + x, e[x]);
        atomicMax(dst + 4 + x, e[3 + x]);
        atomicMin(dst + 7 + x, e[6 + x]);
        atomicMax(dst + 10 + x, e[9 + x]);
    }
}
// The 12 encoded bounds of a primitive: box lo, box hi, centroid (twice).
__device__ __forceinline__ void prim_words(const PBox &p, uint32_t *e) {
    const float lo[3] = {p.lo.x, p.lo.y, p.lo.z}, hi[3] = {p.hi.x, p.hi.y, p.hi.z};
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        const uint32_t c = enc(lo[x] * 0.5f + hi[x] * 0.5f);
        e[x] = enc(lo[x]);
        e[3 + x] = enc(hi[x]);
        e[6 + x] = c;
        e[9 + x] = c;
    }
}
__device__ __forceinline__ void clear_bin(uint32_t *e) {
    e[0] = 0;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        e[1 + x] = e[7 + x] = kEncPosInf;
        e[4 + x] = e[10 + x] = kEncNegInf;
    }
}

// Best split over bins (kBinSlots bins of `stride` words): lane 15 a + i
// evaluates the split after bin i on axis a, as the host loop does, and keeps
// its two sides' unions (box lo hi, centroid lo hi: 12 floats each); the
// wave's minimum cost, ties to the lowest lane (the host keeps the first
// strictly smaller cost in axis-then-bin order). cost == INFINITY: no
// candidate.
struct Best {
    float cost;
    int lane;
    uint32_t lcnt;
};
__device__ __forceinline__ Best sah_best(const uint32_t *bins, int stride, const float *ext3, float parent_area,
                                         int lane, float *lb, float *rb) {
    float cost = INFINITY;
    uint32_t lcnt = 0;
#pragma unroll
    for (int x = 0; x < 12; ++x) {
        const bool lo = x < 3 || (x >= 6 && x < 9);
        lb[x] = rb[x] = lo ? INFINITY : -INFINITY;
    }
    if (lane < 3 * (kSahBins - 1)) {
        const int a = lane / (kSahBins - 1), i = lane % (kSahBins - 1);
        const float ext = pick3(a, ext3[0], ext3[1], ext3[2]);
        if (ext > 0) {
            uint32_t rcnt = 0;
            for (int b = 0; b < kSahBins; ++b) {
                const uint32_t *e = bins + (a * kSahBins + b) * stride;
                const uint32_t cnt = e[0];
                if (cnt == 0) continue;
                float v[12];
#pragma unroll
                for (int x = 0; x < 12; ++x) v[x] = dec(e[1 + x]);
                const bool left = b <= i;
                if (left) lcnt += cnt;
                else rcnt += cnt;
#pragma unroll
                for (int x = 0; x < 12; ++x) {
                    const bool lo = x < 3 || (x >= 6 && x < 9);
                    const float l2 = lo ? fminf(lb[x], v[x]) : fmaxf(lb[x], v[x]);
                    const float r2 = lo ? fminf(rb[x], v[x]) : fmaxf(rb[x], v[x]);
                    lb[x] = left ? l2 : lb[x];
                    rb[x] = left ? rb[x] : r2;
                }
            }
            if (lcnt != 0 && rcnt != 0)
                cost = kSahTraversalCost + ((float)lcnt * sah_area(lb[3] - lb[0], lb[4] - lb[1], lb[5] - lb[2]) +
                                            (float)rcnt * sah_area(rb[3] - rb[0], rb[4] - rb[1], rb[5] - rb[2])) /
                                               fmaxf(parent_area, 1e-30f);
        }
    }
    Best b{cost, lane, lcnt};
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float c2 = __shfl_xor(b.cost, o);
        const int l2 = __shfl_xor(b.lane, o);
        const uint32_t n2 = __shfl_xor(b.lcnt, o);
        if (c2 < b.cost || (c2 == b.cost && l2 < b.lane)) {
            b.cost = c2;
            b.lane = l2;
            b.lcnt = n2;
        }
    }
    b.lane = __builtin_amdgcn_readfirstlane(b.lane);
    b.lcnt = __builtin_amdgcn_readfirstlane(b.lcnt);
    return b;
}

// Child `side` of two-child node `slot`: its box and reference.
__device__ __forceinline__ void set_child(GpuNode *nodes, int32_t slot, uint32_t side, const float *b, int32_t ref) {
    GpuNode &g = nodes[slot];
    if (side == 0) {
        g.c0_lo_x = b[0]; g.c0_hi_x = b[3]; g.c0_lo_y = b[1]; g.c0_hi_y = b[4];
        g.c0_lo_z = b[2]; g.c0_hi_z = b[5];
        g.child0 = ref;
    } else {
        g.c1_lo_x = b[0]; g.c1_hi_x = b[3]; g.c1_lo_y = b[1]; g.c1_hi_y = b[4];
        g.c1_lo_z = b[2]; g.c1_hi_z = b[5];
        g.child1 = ref;
    }
}

// ---- 1. primitive boxes, the root's box and centroid bounds
__global__ __launch_bounds__(256) void prim_kernel(const float *__restrict__ verts, const uint32_t *__restrict__ idx,
                                                   uint32_t n_tris, const float *__restrict__ spheres, uint32_t n,
                                                   PBox *__restrict__ pb, BNode *__restrict__ root) {
    __shared__ float red[4][12];
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    float v[12];  // lo xyz, clo xyz (minimised); hi xyz, chi xyz (maximised)
    for (int a = 0; a < 6; ++a) v[a] = INFINITY, v[6 + a] = -INFINITY;
    if (p < n) {
        float lo[3], hi[3];
        if (p < n_tris) {
            for (int a = 0; a < 3; ++a) lo[a] = hi[a] = verts[3 * (size_t)idx[3 * (size_t)p] + a];
            for (int k = 1; k < 3; ++k)
                for (int a = 0; a < 3; ++a) {
                    const float x = verts[3 * (size_t)idx[3 * (size_t)p + k] + a];
                    lo[a] = fminf(lo[a], x);
                    hi[a] = fmaxf(hi[a], x);
                }
        } else {
            // one ulp outward, as the host builder does
            const float *s = spheres + 4 * (size_t)(p - n_tris);
            for (int a = 0; a < 3; ++a) {
                lo[a] = nextafterf(s[a] + (-s[3]), -INFINITY);
                hi[a] = nextafterf(s[a] + s[3], INFINITY);
            }
        }
        PBox b;
        b.lo = make_float4(lo[0], lo[1], lo[2], __uint_as_float(p));
        b.hi = make_float4(hi[0], hi[1], hi[2], 0.f);
        pb[p] = b;
        for (int a = 0; a < 3; ++a) {
            const float c = lo[a] * 0.5f + hi[a] * 0.5f;
            v[a] = lo[a];
            v[3 + a] = c;
            v[6 + a] = hi[a];
            v[9 + a] = c;
        }
    }
    for (int a = 0; a < 6; ++a) {
        v[a] = wave_min(v[a]);
        v[6 + a] = wave_max(v[6 + a]);
    }
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0)
        for (int a = 0; a < 12; ++a) red[wave][a] = v[a];
    __syncthreads();
    if (threadIdx.x < 12) {
        const int a = threadIdx.x;
        float r = red[0][a];
        for (int w = 1; w < 4; ++w) r = a < 6 ? fminf(r, red[w][a]) : fmaxf(r, red[w][a]);
        // box lo (0..2), cbox lo (3..5), box hi (6..8), cbox hi (9..11)
        uint32_t *dst = a < 3 ? &root->box[a] : a < 6 ? &root->cbox[a - 3] : a < 9 ? &root->box[a - 3] : &root->cbox[a - 6];
        if (a < 6) atomicMin(dst, enc(r));
        else atomicMax(dst, enc(r));
    }
}

__global__ void init_root_kernel(BNode *root, uint32_t n, uint32_t *front, uint32_t *chunk_end, uint32_t *small_list,
                                 Ctr *ctr) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int a = 0; a < 3; ++a) {
        root->box[a] = root->cbox[a] = kEncPosInf;
        root->box[3 + a] = root->cbox[3 + a] = kEncNegInf;
    }
    root->start = 0;
    root->count = n;
    root->parent = -1;
    root->depth_side = 0;
    Ctr c{};
    c.bnodes = 1;
    if (n > (uint32_t)kSmall) {
        front[0] = 0;
        chunk_end[0] = chunks_of(n);
        c.next_front = 1;
    } else {
        small_list[0] = 0;
        c.small = 1;
    }
    *ctr = c;
}

__global__ void clear_bins_kernel(uint32_t *__restrict__ bins, uint32_t n_entries) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_entries) return;
    clear_bin(bins + (size_t)i * kGBinWords);
}

// ---- 2. top phase
// The chunk of block b: frontier node f (chunk_end is the inclusive scan of
// the nodes' chunk counts) and its position range [p0, p1). False past the
// last chunk.
struct Chunk {
    uint32_t f, p0, p1;
};
__device__ __forceinline__ bool find_chunk(const uint32_t *__restrict__ chunk_end, uint32_t nf,
                                           const uint32_t *__restrict__ front, const BNode *__restrict__ bnodes,
                                           uint32_t b, Chunk *c) {
    if (b >= chunk_end[nf - 1]) return false;
    uint32_t lo = 0, hi = nf - 1;  // first f with chunk_end[f] > b
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (chunk_end[mid] > b) hi = mid;
        else lo = mid + 1;
    }
    const uint32_t first = lo ? chunk_end[lo - 1] : 0u, nch = chunk_end[lo] - first;
    const BNode &B = bnodes[front[lo]];
    const uint32_t per = (B.count + nch - 1) / nch;
    c->f = lo;
    c->p0 = B.start + min(B.count, (b - first) * per);
    c->p1 = B.start + min(B.count, (b - first + 1) * per);
    return true;
}

__global__ __launch_bounds__(256) void bin_kernel(const PBox *__restrict__ pb, const uint32_t *__restrict__ chunk_end,
                                                  uint32_t nf, const uint32_t *__restrict__ front,
                                                  const BNode *__restrict__ bnodes, uint32_t *__restrict__ bins) {
    __shared__ uint32_t lb[kBinSlots * kBinWords];
    Chunk c;
    if (!find_chunk(chunk_end, nf, front, bnodes, blockIdx.x, &c)) return;
    for (int i = threadIdx.x; i < kBinSlots; i += 256) clear_bin(lb + i * kBinWords);
    const BNode &B = bnodes[front[c.f]];
    float lo[3], k[3];
    bool ok[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = dec(B.cbox[a]);
        const float ext = dec(B.cbox[3 + a]) - lo[a];
        ok[a] = ext > 0;
        k[a] = ok[a] ? (float)kSahBins / ext : 0.f;
    }
    __syncthreads();
    for (uint32_t pos = c.p0 + threadIdx.x; pos < c.p1; pos += 256) {
        const PBox p = pb[pos];
        uint32_t e[12];
        prim_words(p, e);
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (ok[a]) bin_add(lb + (a * kSahBins + bin_of(centroid(p, a), lo[a], k[a])) * kBinWords, e);
    }
    __syncthreads();
    if (threadIdx.x < kBinSlots) {
        const uint32_t *src = lb + threadIdx.x * kBinWords;
        if (src[0]) {
            uint32_t *dst = bins + ((size_t)c.f * kBinSlots + threadIdx.x) * kGBinWords;
            atomicAdd(dst, src[0]);
            for (int x = 0; x < 3; ++x) {
                atomicMin(dst + 1 + x, src[1 + x]);
                atomicMax(dst + 4 + x, src[4 + x]);
                atomicMin(dst + 7 + x, src[7 + x]);
                atomicMax(dst + 10 + x, src[10 + x]);
            }
        }
    }
}

__device__ __forceinline__ void enc_bounds(uint32_t *box, uint32_t *cbox, const float *v) {
    for (int x = 0; x < 6; ++x) {
        box[x] = enc(v[x]);
        cbox[x] = enc(v[6 + x]);
    }
}

// One wave per top-phase node: the split, the node's slot and its entry in
// the parent, the two children's records (boxes and centroid bounds from the
// bins; a count split leaves them to the scatter), the next frontier with each
// large child's chunk count.
__global__ __launch_bounds__(256) void split_kernel(const uint32_t *__restrict__ front, uint32_t nf,
                                                    BNode *__restrict__ bnodes, const uint32_t *__restrict__ bins,
                                                    Split *__restrict__ split, uint32_t *__restrict__ next_front,
                                                    uint32_t *__restrict__ next_chunks,
                                                    uint32_t *__restrict__ small_list, GpuNode *__restrict__ nodes,
                                                    uint8_t *__restrict__ used, Ctr *__restrict__ ctr,
                                                    uint32_t next_buf) {
    const uint32_t f = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (f >= nf) return;
    const uint32_t id = front[f];
    const BNode B = bnodes[id];
    float box[6], cb_lo[3], ext3[3];
    for (int a = 0; a < 3; ++a) {
        box[a] = dec(B.box[a]);
        box[3 + a] = dec(B.box[3 + a]);
        cb_lo[a] = dec(B.cbox[a]);
        ext3[a] = dec(B.cbox[3 + a]) - cb_lo[a];
    }
    const uint32_t n = B.count, depth = B.depth_side >> 2;
    const float parent_area = sah_area(box[3] - box[0], box[4] - box[1], box[5] - box[2]);
    float lb[12], rb[12];
    const Best best = sah_best(bins + (size_t)f * kBinSlots * kGBinWords, kGBinWords, ext3, parent_area, lane, lb, rb);
    float cv[2][12];
#pragma unroll
    for (int x = 0; x < 12; ++x) {
        cv[0][x] = bcast(lb[x], best.lane);
        cv[1][x] = bcast(rb[x], best.lane);
    }
    if (lane != 0) return;
    Split s;
    s.start = B.start;
    s.pad[0] = s.pad[1] = s.pad[2] = 0;
    if (depth >= (uint32_t)kSahMedianDepth || best.cost == INFINITY) {
        // coincident centroids (the host's count split) or, below the median
        // depth, a count split standing in for the host's object median
        s.mode = 1;
        s.axis = s.bin = 0;
        s.k = s.lo = 0.f;
        s.lcnt = n / 2;
    } else {
        s.mode = 0;
        s.axis = best.lane / (kSahBins - 1);
        s.bin = best.lane % (kSahBins - 1);
        s.lo = pick3((int)s.axis, cb_lo[0], cb_lo[1], cb_lo[2]);
        s.k = (float)kSahBins / pick3((int)s.axis, ext3[0], ext3[1], ext3[2]);
        s.lcnt = best.lcnt;
    }
    const int32_t slot = (int32_t)atomicAdd(&ctr->alloc, 1u);
    atomicAdd(&ctr->inner, 1u);
    used[slot] = 1;
    if (B.parent >= 0) set_child(nodes, B.parent, B.depth_side & 1u, box, slot);
    const uint32_t c0 = atomicAdd(&ctr->bnodes, 2u);
    for (uint32_t k = 0; k < 2; ++k) {
        BNode c;
        if (s.mode == 0) {
            enc_bounds(c.box, c.cbox, cv[k]);
        } else {
            for (int a = 0; a < 3; ++a) {
                c.box[a] = c.cbox[a] = kEncPosInf;
                c.box[3 + a] = c.cbox[3 + a] = kEncNegInf;
            }
        }
        c.start = k ? s.start + s.lcnt : s.start;
        c.count = k ? n - s.lcnt : s.lcnt;
        c.parent = slot;
        c.depth_side = (depth + 1) << 2 | next_buf << 1 | k;
        bnodes[c0 + k] = c;
        s.child[k] = c0 + k;
        if (c.count > (uint32_t)kSmall) {
            const uint32_t cf = atomicAdd(&ctr->next_front, 1u);
            next_front[cf] = c0 + k;
            next_chunks[cf] = chunks_of(c.count);
        } else {
            small_list[atomicAdd(&ctr->small, 1u)] = c0 + k;
        }
    }
    split[f] = s;
}

__device__ __forceinline__ uint32_t goes_left(const Split &s, const PBox &p, uint32_t pos) {
    if (s.mode == 0) return bin_of(centroid(p, (int)s.axis), s.lo, s.k) <= (int)s.bin ? 1u : 0u;
    return pos - s.start < s.lcnt ? 1u : 0u;
}

__global__ __launch_bounds__(256) void flag_kernel(const PBox *__restrict__ pb, const uint32_t *__restrict__ chunk_end,
                                                   uint32_t nf, const uint32_t *__restrict__ front,
                                                   const BNode *__restrict__ bnodes, const Split *__restrict__ split,
                                                   uint32_t *__restrict__ flags) {
    Chunk c;
    if (!find_chunk(chunk_end, nf, front, bnodes, blockIdx.x, &c)) return;
    const Split s = split[c.f];
    for (uint32_t pos = c.p0 + threadIdx.x; pos < c.p1; pos += 256) flags[pos] = goes_left(s, pb[pos], pos);
}

// Stable partition of every top-phase node's range by its flags into the
// other buffer. Children of count splits get their boxes and centroid bounds
// here (wave-reduced atomics).
__global__ __launch_bounds__(256) void scatter_kernel(const PBox *__restrict__ pb,
                                                      const uint32_t *__restrict__ chunk_end, uint32_t nf,
                                                      const uint32_t *__restrict__ front,
                                                      BNode *__restrict__ bnodes, const Split *__restrict__ split,
                                                      const uint32_t *__restrict__ flags,
                                                      const uint32_t *__restrict__ excl, PBox *__restrict__ pb2) {
    Chunk c;
    if (!find_chunk(chunk_end, nf, front, bnodes, blockIdx.x, &c)) return;
    const Split s = split[c.f];
    const int lane = threadIdx.x % 64;
    const uint32_t base = excl[s.start];
    for (uint32_t p0 = c.p0; p0 < c.p1; p0 += 256) {
        const uint32_t pos = p0 + threadIdx.x;
        uint32_t k = 2;  // no primitive
        PBox p;
        p.lo = p.hi = make_float4(0.f, 0.f, 0.f, 0.f);
        if (pos < c.p1) {
            p = pb[pos];
            const uint32_t before = excl[pos] - base;
            k = flags[pos] ? 0u : 1u;
            pb2[k == 0 ? s.start + before : s.start + s.lcnt + (pos - s.start - before)] = p;
        }
        if (s.mode == 0) continue;
        // count split: the children's bounds
#pragma unroll
        for (uint32_t side = 0; side < 2; ++side) {
            const bool in = k == side;
            if (!__ballot(in)) continue;
            float v[12];
            const float lo[3] = {p.lo.x, p.lo.y, p.lo.z}, hi[3] = {p.hi.x, p.hi.y, p.hi.z};
            for (int a = 0; a < 3; ++a) {
                const float ce = lo[a] * 0.5f + hi[a] * 0.5f;
                v[a] = in ? lo[a] : INFINITY;
                v[3 + a] = in ? ce : INFINITY;
                v[6 + a] = in ? hi[a] : -INFINITY;
                v[9 + a] = in ? ce : -INFINITY;
            }
            for (int a = 0; a < 6; ++a) {
                v[a] = wave_min(v[a]);
                v[6 + a] = wave_max(v[6 + a]);
            }
            if (lane == 0) {
                BNode &C = bnodes[s.child[side]];
                for (int a = 0; a < 3; ++a) {
                    atomicMin(&C.box[a], enc(v[a]));
                    atomicMin(&C.cbox[a], enc(v[3 + a]));
                    atomicMax(&C.box[3 + a], enc(v[6 + a]));
                    atomicMax(&C.cbox[3 + a], enc(v[9 + a]));
                }
            }
        }
    }
}

// ---- 3. bottom phase: one wave per subtree root of at most kSmall primitives
struct SmallLds {
    float4 lo[kSmall], hi[kSmall];
    uint32_t bins[kBinSlots * kBinWords];
    uint32_t st_range[kStack];  // local start << 16 | count
    int32_t st_parent[kStack];
    uint32_t st_ds[kStack];     // depth << 2 | bounds known << 1 | side
    float st_b[kStack][12];     // box lo hi, centroid lo hi (when known)
};

__global__ __launch_bounds__(256) void small_kernel(PBox *__restrict__ pb0, const PBox *__restrict__ pb1,
                                                    const uint32_t *__restrict__ small_list, uint32_t n_small,
                                                    const BNode *__restrict__ bnodes, GpuNode *__restrict__ nodes,
                                                    uint8_t *__restrict__ used, Ctr *__restrict__ ctr, int max_prims) {
    __shared__ SmallLds lds[4];
    const uint32_t item = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (item >= n_small) return;
    SmallLds &L = lds[threadIdx.x / 64];
    const BNode R = bnodes[small_list[item]];
    const uint32_t s0 = R.start, m0 = R.count;
    const PBox *src = (R.depth_side >> 1) & 1u ? pb1 : pb0;
    for (uint32_t j = lane; j < m0; j += 64) {
        const PBox p = src[s0 + j];
        L.lo[j] = p.lo;
        L.hi[j] = p.hi;
    }
    // the root's bounds: from its parent's bins, the scatter's reduction or
    // the primitive pass (every top-phase child has them)
    float nb[12];
    for (int x = 0; x < 6; ++x) {
        nb[x] = dec(R.box[x]);
        nb[6 + x] = dec(R.cbox[x]);
    }
    uint32_t s = 0, m = m0, ds = (R.depth_side >> 2) << 2 | 2u | (R.depth_side & 1u);
    int32_t parent = R.parent;
    int sp = 0;
    // the subtree's node slots: at most m0 - 1 inner nodes
    uint32_t base = 0, n_inner = 0;
    if (lane == 0 && m0 > 1) base = atomicAdd(&ctr->alloc, m0 - 1);
    base = __shfl(base, 0);
    wave_lds_sync();
    uint32_t leaves = 0, max_depth = 0;
    bool overflow = false;
    for (;;) {
        const uint32_t depth = ds >> 2;
        // this lane's primitives: local s + lane and s + 64 + lane
        PBox p[2];
        bool valid[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t j = (uint32_t)lane + 64u * h;
            valid[h] = j < m;
            p[h].lo = valid[h] ? L.lo[s + j] : make_float4(0.f, 0.f, 0.f, 0.f);
            p[h].hi = valid[h] ? L.hi[s + j] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        if (!((ds >> 1) & 1u)) {
            // bounds not known (a count split or an object median made the
            // node): reduce them over its primitives
            float v[12];
            for (int x = 0; x < 6; ++x) v[x] = INFINITY, v[6 + x] = -INFINITY;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (!valid[h]) continue;
                const float lo[3] = {p[h].lo.x, p[h].lo.y, p[h].lo.z}, hi[3] = {p[h].hi.x, p[h].hi.y, p[h].hi.z};
                for (int a = 0; a < 3; ++a) {
                    const float c = lo[a] * 0.5f + hi[a] * 0.5f;
                    v[a] = fminf(v[a], lo[a]);
                    v[3 + a] = fminf(v[3 + a], c);
                    v[6 + a] = fmaxf(v[6 + a], hi[a]);
                    v[9 + a] = fmaxf(v[9 + a], c);
                }
            }
            for (int x = 0; x < 6; ++x) {
                v[x] = wave_min(v[x]);
                v[6 + x] = wave_max(v[6 + x]);
            }
            // -> box lo hi, centroid lo hi
            for (int a = 0; a < 3; ++a) {
                nb[a] = v[a];
                nb[3 + a] = v[6 + a];
                nb[6 + a] = v[3 + a];
                nb[9 + a] = v[9 + a];
            }
        }
        const float *box = nb;
        const float ext3[3] = {nb[9] - nb[6], nb[10] - nb[7], nb[11] - nb[8]};
        // decision, in the host builder's order
        int mode = -1;  // -1 leaf, 0 bins, 1 first half, 2 object median
        int axis = 0, bin = 0;
        float k = 0.f, lo_ax = 0.f;
        float cv[2][12];
        if (m == 1) {
            mode = -1;
        } else if (depth >= (uint32_t)kSahMedianDepth) {
            mode = m <= (uint32_t)max_prims ? -1 : 2;
        } else {
            for (int i = lane; i < kBinSlots; i += 64) clear_bin(L.bins + i * kBinWords);
            wave_lds_sync();
            float kk[3];
            for (int a = 0; a < 3; ++a) kk[a] = ext3[a] > 0 ? (float)kSahBins / ext3[a] : 0.f;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (!valid[h]) continue;
                uint32_t e[12];
                prim_words(p[h], e);
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    if (ext3[a] > 0)
                        bin_add(L.bins + (a * kSahBins + bin_of(centroid(p[h], a), nb[6 + a], kk[a])) * kBinWords, e);
            }
            wave_lds_sync();
            const float parent_area = sah_area(box[3] - box[0], box[4] - box[1], box[5] - box[2]);
            float lb[12], rb[12];
            const Best best = sah_best(L.bins, kBinWords, ext3, parent_area, lane, lb, rb);
            wave_lds_sync();
            if (best.cost == INFINITY) {
                mode = m <= (uint32_t)max_prims ? -1 : 1;
            } else if (m <= (uint32_t)max_prims && best.cost >= (float)m) {
                mode = -1;
            } else {
                mode = 0;
                axis = best.lane / (kSahBins - 1);
                bin = best.lane % (kSahBins - 1);
                lo_ax = pick3(axis, nb[6], nb[7], nb[8]);
                k = (float)kSahBins / pick3(axis, ext3[0], ext3[1], ext3[2]);
#pragma unroll
                for (int x = 0; x < 12; ++x) {
                    cv[0][x] = bcast(lb[x], best.lane);
                    cv[1][x] = bcast(rb[x], best.lane);
                }
            }
        }
        bool pop = false;
        if (mode < 0) {
            if (lane == 0 && parent >= 0) set_child(nodes, parent, ds & 1u, box, make_leaf_ref(s0 + s, m));
            ++leaves;
            max_depth = max(max_depth, depth);
            pop = true;
        } else {
            const int32_t slot = (int32_t)(base + n_inner++);
            if (lane == 0) {
                used[slot] = 1;
                if (parent >= 0) set_child(nodes, parent, ds & 1u, box, slot);
            }
            // flags
            bool left[2] = {false, false};
            int med_axis = 0;
            if (mode == 2) {
                float ext = -1.f;
                for (int a = 0; a < 3; ++a)
                    if (ext3[a] > ext) {
                        ext = ext3[a];
                        med_axis = a;
                    }
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (!valid[h]) continue;
                const uint32_t j = (uint32_t)lane + 64u * h;
                if (mode == 0) {
                    left[h] = bin_of(centroid(p[h], axis), lo_ax, k) <= bin;
                } else if (mode == 1) {
                    left[h] = j < m / 2;
                } else {
                    // object median: rank by (centroid, position)
                    const float cj = centroid(p[h], med_axis);
                    uint32_t rank = 0;
                    for (uint32_t t = 0; t < m; ++t) {
                        PBox q;
                        q.lo = L.lo[s + t];
                        q.hi = L.hi[s + t];
                        const float ct = centroid(q, med_axis);
                        rank += (ct < cj || (ct == cj && t < j)) ? 1u : 0u;
                    }
                    left[h] = rank < m / 2;
                }
            }
            const uint64_t l0 = __ballot(valid[0] && left[0]), l1 = __ballot(valid[1] && left[1]);
            const uint64_t r0 = __ballot(valid[0] && !left[0]), r1 = __ballot(valid[1] && !left[1]);
            const uint32_t nl0 = (uint32_t)__popcll(l0), nl = nl0 + (uint32_t)__popcll(l1);
            const uint32_t nr0 = (uint32_t)__popcll(r0);
            wave_lds_sync();
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (!valid[h]) continue;
                uint32_t np;
                if (left[h]) np = h == 0 ? lanes_below(l0, lane) : nl0 + lanes_below(l1, lane);
                else np = nl + (h == 0 ? lanes_below(r0, lane) : nr0 + lanes_below(r1, lane));
                L.lo[s + np] = p[h].lo;
                L.hi[s + np] = p[h].hi;
            }
            wave_lds_sync();
            // right child pending, continue with the left
            if (sp >= kStack) {
                overflow = true;
                break;
            }
            const uint32_t known = mode == 0 ? 2u : 0u;
            if (lane == 0) {
                L.st_range[sp] = (s + nl) << 16 | (m - nl);
                L.st_parent[sp] = slot;
                L.st_ds[sp] = (depth + 1) << 2 | known | 1u;
            }
            if (mode == 0) {
#pragma unroll
                for (int x = 0; x < 12; ++x) {
                    if (lane == 0) L.st_b[sp][x] = cv[1][x];
                    nb[x] = cv[0][x];
                }
            }
            ++sp;
            wave_lds_sync();
            m = nl;
            parent = slot;
            ds = (depth + 1) << 2 | known;
        }
        if (pop) {
            if (sp == 0) break;
            --sp;
            s = L.st_range[sp] >> 16;
            m = L.st_range[sp] & 0xffffu;
            parent = L.st_parent[sp];
            ds = L.st_ds[sp];
#pragma unroll
            for (int x = 0; x < 12; ++x) nb[x] = L.st_b[sp][x];
            wave_lds_sync();
        }
    }
    for (uint32_t j = lane; j < m0; j += 64) {
        PBox p;
        p.lo = L.lo[j];
        p.hi = L.hi[j];
        pb0[s0 + j] = p;
    }
    if (lane == 0) {
        atomicAdd(&ctr->inner, n_inner);
        atomicAdd(&ctr->leaves, leaves);
        atomicMax(&ctr->depth, max_depth);
        if (overflow) atomicOr(&ctr->overflow, 1u);
    }
}

// ---- 4. records and the four-child collapse
__global__ void records_kernel(const float *__restrict__ verts, const uint32_t *__restrict__ idx, uint32_t n_tris,
                               const float *__restrict__ spheres, uint32_t n, const PBox *__restrict__ pb,
                               PrimRecord *__restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t p = __float_as_uint(pb[k].lo.w);
    PrimRecord r;
    for (int a = 0; a < 4; ++a) r.a[a] = r.b[a] = r.c[a] = 0.f;
    if (p < n_tris) {
        for (int a = 0; a < 3; ++a) {
            r.a[a] = verts[3 * (size_t)idx[3 * (size_t)p] + a];
            r.b[a] = verts[3 * (size_t)idx[3 * (size_t)p + 1] + a];
            r.c[a] = verts[3 * (size_t)idx[3 * (size_t)p + 2] + a];
        }
        r.a[3] = __uint_as_float(p);
    } else {
        const float *s = spheres + 4 * (size_t)(p - n_tris);
        r.a[0] = s[0];
        r.a[1] = s[1];
        r.a[2] = s[2];
        r.a[3] = __uint_as_float(p | kSphereBit);
        r.b[0] = s[3];
    }
    out[k] = r;
}

struct Kid {
    float b[6];
    int32_t ref;
};
__device__ __forceinline__ Kid kid_of(const GpuNode &g, int c) {
    Kid k;
    if (c == 0) {
        k.b[0] = g.c0_lo_x; k.b[1] = g.c0_lo_y; k.b[2] = g.c0_lo_z;
        k.b[3] = g.c0_hi_x; k.b[4] = g.c0_hi_y; k.b[5] = g.c0_hi_z;
        k.ref = g.child0;
    } else {
        k.b[0] = g.c1_lo_x; k.b[1] = g.c1_lo_y; k.b[2] = g.c1_lo_z;
        k.b[3] = g.c1_hi_x; k.b[4] = g.c1_hi_y; k.b[5] = g.c1_hi_z;
        k.ref = g.child1;
    }
    return k;
}

// One thread per four-child node of this level (item = binary node, slot):
// the children by the host collapse's rule, slots for the inner ones
// (consecutive), the quantized node.
// k slots from *ctr for every lane of the wave, one atomic per wave: the
// lane's first slot (its k may be 0). Every lane of the wave calls it.
__device__ __forceinline__ uint32_t wave_alloc(uint32_t *ctr, uint32_t k, int lane) {
    uint32_t x = k;  // inclusive prefix sum over the lanes
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    const uint32_t total = __shfl(x, 63);
    uint32_t base = 0;
    if (lane == 0 && total) base = atomicAdd(ctr, total);
    return __shfl(base, 0) + x - k;
}

// cap2 bounds the binary node indices (the slots handed out), cap4 the items,
// the next items and the four-child slots (the binary tree's inner-node
// count); a malformed tree sets overflow bit 4 instead of reading or writing
// past them.
__global__ __launch_bounds__(256) void collapse_kernel(const GpuNode *__restrict__ n2, const uint2 *__restrict__ items,
                                                       uint32_t n_items, uint2 *__restrict__ next,
                                                       GpuNode4Q *__restrict__ out, uint32_t cap2, uint32_t cap4,
                                                       Ctr *__restrict__ ctr) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x % 64;
    bool ok = i < n_items;
    const uint2 it = ok ? items[i] : make_uint2(0u, 0u);
    if (ok && (it.x >= cap2 || it.y >= cap4)) {
        atomicOr(&ctr->overflow, 4u);
        ok = false;
    }
    Kid kids[4];
    int nk = 0;
    if (ok) {
        nk = 2;
        kids[0] = kid_of(n2[it.x], 0);
        kids[1] = kid_of(n2[it.x], 1);
        while (nk < 4) {
            int best = -1;
            float best_area = -1.f;
            for (int c = 0; c < nk; ++c)
                if (kids[c].ref >= 0 && collapse_area(kids[c].b) > best_area) {
                    best_area = collapse_area(kids[c].b);
                    best = c;
                }
            if (best < 0) break;
            if ((uint32_t)kids[best].ref >= cap2) {
                atomicOr(&ctr->overflow, 4u);
                ok = false;
                break;
            }
            const GpuNode g = n2[kids[best].ref];
            kids[best] = kid_of(g, 0);
            kids[nk++] = kid_of(g, 1);
        }
    }
    uint32_t n_inner = 0;
    if (ok)
        for (int c = 0; c < nk; ++c) n_inner += kids[c].ref >= 0 ? 1u : 0u;
    const uint32_t base = wave_alloc(&ctr->n4, n_inner, lane);
    const uint32_t nbase = wave_alloc(&ctr->next_front, n_inner, lane);
    if (!ok) return;
    GpuNode4 g4;
    uint32_t used = 0;
    for (int c = 0; c < 4; ++c) {
        g4.pad[c] = 0;
        if (c >= nk) {
            g4.lo_x[c] = g4.hi_x[c] = g4.lo_y[c] = g4.hi_y[c] = g4.lo_z[c] = g4.hi_z[c] = INFINITY;
            g4.child[c] = make_leaf_ref(0, 0);
            continue;
        }
        const Kid &k = kids[c];
        g4.lo_x[c] = k.b[0]; g4.lo_y[c] = k.b[1]; g4.lo_z[c] = k.b[2];
        g4.hi_x[c] = k.b[3]; g4.hi_y[c] = k.b[4]; g4.hi_z[c] = k.b[5];
        if (k.ref < 0) {
            g4.child[c] = k.ref;
        } else {
            g4.child[c] = (int32_t)(base + used);
            if (nbase + used < cap4 && base + used < cap4) next[nbase + used] = make_uint2((uint32_t)k.ref, base + used);
            else atomicOr(&ctr->overflow, 4u);
            ++used;
        }
    }
    GpuNode4Q q;
    if (!quantize_node(g4, &q)) atomicOr(&ctr->overflow, 2u);
    out[it.y] = q;
}

// Width 2: the used slots, in slot order, renumbered by the exclusive scan
// `idx` of the used marks (child references follow).
__global__ void compact_kernel(const GpuNode *__restrict__ n2, const uint8_t *__restrict__ used,
                               const uint32_t *__restrict__ idx, uint32_t n_slots, GpuNode *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_slots || !used[i]) return;
    GpuNode g = n2[i];
    if (g.child0 >= 0) g.child0 = (int32_t)idx[g.child0];
    if (g.child1 >= 0) g.child1 = (int32_t)idx[g.child1];
    out[idx[i]] = g;
}

__global__ void set_u32_kernel(uint32_t *p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// The collapse's first item (the root, slot 0) and its counters.
__global__ void collapse_init_kernel(uint2 *items, Ctr *ctr) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    items[0] = make_uint2(0u, 0u);
    Ctr c{};
    c.n4 = 1;
    *ctr = c;
}

// The build's temporaries: stream-ordered allocations on the build's stream,
// freed in stream order (hipFree would wait for the whole device, so builds
// of several scenes on their own streams could not overlap).
struct DevBuf {
    hipStream_t st;
    std::vector<void *> ptrs;
    explicit DevBuf(hipStream_t s) : st(s) {}
    ~DevBuf() {
        for (void *p : ptrs) (void)hipFreeAsync(p, st);
    }
    template <typename T>
    hipError_t alloc(T **p, size_t count) {
        hipError_t e = hipMallocAsync((void **)p, std::max<size_t>(count, 1) * sizeof(T), st);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

}  // namespace

hipError_t sah_gpu_build(const float *verts, uint32_t n_verts, const uint32_t *idx, uint32_t n_tris,
                         const float *spheres, uint32_t n_spheres, int max_prims, int width, hipStream_t st,
                         SahGpuResult *res) {
    const uint32_t n = n_tris + n_spheres;
    *res = SahGpuResult();
    DevBuf tmp(st);
    float *d_verts = nullptr, *d_sph = nullptr;
    uint32_t *d_idx = nullptr;
    hipError_t e = hipSuccess;
    if (n_tris) {
        if ((e = tmp.alloc(&d_verts, 3 * (size_t)n_verts)) != hipSuccess) return e;
        if ((e = tmp.alloc(&d_idx, 3 * (size_t)n_tris)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(d_verts, verts, 12 * (size_t)n_verts, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(d_idx, idx, 12 * (size_t)n_tris, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    }
    if (n_spheres) {
        if ((e = tmp.alloc(&d_sph, 4 * (size_t)n_spheres)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(d_sph, spheres, 16 * (size_t)n_spheres, hipMemcpyHostToDevice, st)) != hipSuccess)
            return e;
    }
    // frontier nodes of one level hold more than kSmall primitives each
    const uint32_t front_cap = n / (kSmall + 1) + 1;
    PBox *pb[2];
    uint32_t *flags, *excl, *front[2], *next_chunks, *chunk_end, *small_list, *bins;
    uint8_t *used;
    Split *split;
    Ctr *ctr;
    GpuNode *n2;
    if ((e = tmp.alloc(&pb[0], n)) != hipSuccess || (e = tmp.alloc(&pb[1], n)) != hipSuccess ||
        (e = tmp.alloc(&flags, n)) != hipSuccess || (e = tmp.alloc(&excl, n)) != hipSuccess ||
        (e = tmp.alloc(&front[0], front_cap)) != hipSuccess || (e = tmp.alloc(&front[1], front_cap)) != hipSuccess ||
        (e = tmp.alloc(&next_chunks, front_cap)) != hipSuccess || (e = tmp.alloc(&chunk_end, front_cap)) != hipSuccess ||
        (e = tmp.alloc(&small_list, n)) != hipSuccess ||
        (e = tmp.alloc(&bins, (size_t)front_cap * kBinSlots * kGBinWords)) != hipSuccess ||
        (e = tmp.alloc(&split, front_cap)) != hipSuccess || (e = tmp.alloc(&ctr, 1)) != hipSuccess ||
        (e = tmp.alloc(&used, n)) != hipSuccess)
        return e;
    if ((e = hipMemsetAsync(used, 0, n, st)) != hipSuccess) return e;
    // BNodes: the root, then two per top-phase split; grown on demand
    size_t bnode_cap = 1 + 2 * (size_t)front_cap * 4;
    BNode *bnodes = nullptr;
    if ((e = hipMallocAsync((void **)&bnodes, bnode_cap * sizeof(BNode), st)) != hipSuccess) return e;
    struct BNodeGuard {
        BNode **p;
        hipStream_t st;
        ~BNodeGuard() { (void)hipFreeAsync(*p, st); }
    } bguard{&bnodes, st};
    if ((e = tmp.alloc(&n2, (size_t)std::max<uint32_t>(n - 1, 1))) != hipSuccess) return e;
    if ((e = hipMemsetAsync(n2, 0, (size_t)std::max<uint32_t>(n - 1, 1) * sizeof(GpuNode), st)) != hipSuccess) return e;
    size_t scan_bytes = 0, cscan_bytes = 0;
    if ((e = rocprim::exclusive_scan(nullptr, scan_bytes, flags, excl, 0u, n, rocprim::plus<uint32_t>(), st)) !=
            hipSuccess ||
        (e = rocprim::inclusive_scan(nullptr, cscan_bytes, next_chunks, chunk_end, front_cap,
                                     rocprim::plus<uint32_t>(), st)) != hipSuccess)
        return e;
    scan_bytes = std::max(scan_bytes, cscan_bytes);
    void *scan_tmp = nullptr;
    if ((e = tmp.alloc((char **)&scan_tmp, scan_bytes)) != hipSuccess) return e;

    init_root_kernel<<<1, 64, 0, st>>>(bnodes, n, front[0], chunk_end, small_list, ctr);
    prim_kernel<<<blocks_for(n, 256), 256, 0, st>>>(d_verts, d_idx, n_tris, d_sph, n, pb[0], bnodes);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    Ctr hc;
    if ((e = hipMemcpyAsync(&hc, ctr, sizeof(Ctr), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    uint32_t cur = 0;
    uint32_t nf = hc.next_front;
    for (int level = 0; nf > 0; ++level) {
        if (level >= kMaxTopLevels) return hipErrorUnknown;
        // room for this level's children
        if (hc.bnodes + 2 * (size_t)nf > bnode_cap) {
            const size_t cap2 = std::max(bnode_cap * 2, hc.bnodes + 2 * (size_t)nf);
            BNode *b2 = nullptr;
            if ((e = hipMallocAsync((void **)&b2, cap2 * sizeof(BNode), st)) != hipSuccess) return e;
            if ((e = hipMemcpyAsync(b2, bnodes, hc.bnodes * sizeof(BNode), hipMemcpyDeviceToDevice, st)) != hipSuccess) {
                (void)hipFreeAsync(b2, st);
                return e;
            }
            (void)hipFreeAsync(bnodes, st);  // after the copy, in stream order
            bnodes = b2;
            bnode_cap = cap2;
        }
        const uint32_t nxt = cur ^ 1u;
        // chunks: at most ceil(count / kChunk) per node, so this many blocks cover them
        const uint32_t grid = n / kChunk + nf + 1;
        set_u32_kernel<<<1, 64, 0, st>>>(&ctr->next_front, 0u);
        clear_bins_kernel<<<blocks_for((uint64_t)nf * kBinSlots, 256), 256, 0, st>>>(bins, nf * kBinSlots);
        bin_kernel<<<grid, 256, 0, st>>>(pb[cur], chunk_end, nf, front[cur], bnodes, bins);
        split_kernel<<<blocks_for(nf, 4), 256, 0, st>>>(front[cur], nf, bnodes, bins, split, front[nxt], next_chunks,
                                                       small_list, n2, used, ctr, nxt);
        flag_kernel<<<grid, 256, 0, st>>>(pb[cur], chunk_end, nf, front[cur], bnodes, split, flags);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = rocprim::exclusive_scan(scan_tmp, scan_bytes, flags, excl, 0u, n, rocprim::plus<uint32_t>(), st)) !=
            hipSuccess)
            return e;
        scatter_kernel<<<grid, 256, 0, st>>>(pb[cur], chunk_end, nf, front[cur], bnodes, split, flags, excl, pb[nxt]);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // the next level's chunk table (entries past its frontier are never read)
        if ((e = rocprim::inclusive_scan(scan_tmp, scan_bytes, next_chunks, chunk_end, front_cap,
                                         rocprim::plus<uint32_t>(), st)) != hipSuccess)
            return e;
        if ((e = hipMemcpyAsync(&hc, ctr, sizeof(Ctr), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        cur = nxt;
        nf = hc.next_front;
        res->top_levels = level + 1;
    }
    if (hc.small > 0)
        small_kernel<<<blocks_for(hc.small, 4), 256, 0, st>>>(pb[0], pb[1], small_list, hc.small, bnodes, n2, used,
                                                              ctr, max_prims);
    PrimRecord *prims = nullptr;
    if ((e = hipMalloc(&prims, (size_t)n * sizeof(PrimRecord))) != hipSuccess) return e;
    records_kernel<<<blocks_for(n, 256), 256, 0, st>>>(d_verts, d_idx, n_tris, d_sph, n, pb[0], prims);
    if ((e = hipGetLastError()) == hipSuccess) e = hipMemcpyAsync(&hc, ctr, sizeof(Ctr), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && (hc.overflow || hc.inner != hc.leaves - 1 || hc.alloc > std::max<uint32_t>(n - 1, 1)))
        e = hipErrorUnknown;
    // the root's box
    if (e == hipSuccess) {
        uint32_t rb[6];
        e = hipMemcpyAsync(rb, bnodes, sizeof(rb), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        for (int a = 0; a < 6; ++a) {
            const uint32_t o = rb[a];
            const uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
            std::memcpy(&res->bounds[a], &u, 4);
        }
    }
    if (e != hipSuccess) {
        (void)hipFree(prims);
        return e;
    }
    res->prims = prims;
    res->n_leaves = hc.leaves;
    res->binary_depth = hc.depth;
    res->n_nodes = hc.inner;
    if (width == 2) {
        // the used slots, compacted in slot order
        uint32_t *cidx = nullptr;
        GpuNode *out = nullptr;
        size_t cbytes = 0;
        void *ctmp = nullptr;
        e = tmp.alloc(&cidx, hc.alloc);
        if (e == hipSuccess)
            e = rocprim::exclusive_scan(nullptr, cbytes, used, cidx, 0u, hc.alloc, rocprim::plus<uint32_t>(), st);
        if (e == hipSuccess) e = tmp.alloc((char **)&ctmp, cbytes);
        if (e == hipSuccess)
            e = rocprim::exclusive_scan(ctmp, cbytes, used, cidx, 0u, hc.alloc, rocprim::plus<uint32_t>(), st);
        if (e == hipSuccess) e = hipMalloc(&out, (size_t)hc.inner * sizeof(GpuNode));
        if (e == hipSuccess) {
            compact_kernel<<<blocks_for(hc.alloc, 256), 256, 0, st>>>(n2, used, cidx, hc.alloc, out);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            (void)hipFree(out);
            (void)hipFree(prims);
            res->prims = nullptr;
            return e;
        }
        res->nodes = out;
        res->depth = hc.depth;
        return hipSuccess;
    }
    // four-child collapse, one launch per four-child level
    GpuNode4Q *n4 = nullptr;
    uint2 *items[2];
    if ((e = tmp.alloc(&items[0], hc.inner)) != hipSuccess || (e = tmp.alloc(&items[1], hc.inner)) != hipSuccess ||
        (e = hipMalloc(&n4, (size_t)hc.inner * sizeof(GpuNode4Q))) != hipSuccess) {
        (void)hipFree(prims);
        res->prims = nullptr;
        return e;
    }
    Ctr c4{};
    collapse_init_kernel<<<1, 64, 0, st>>>(items[0], ctr);
    uint32_t n_items = 1, levels = 0;
    int ic = 0;
    while (e == hipSuccess && n_items > 0) {
        set_u32_kernel<<<1, 64, 0, st>>>(&ctr->next_front, 0u);
        collapse_kernel<<<blocks_for(n_items, 256), 256, 0, st>>>(n2, items[ic], n_items, items[ic ^ 1], n4,
                                                                 hc.alloc, hc.inner, ctr);
        if ((e = hipGetLastError()) != hipSuccess) break;
        if ((e = hipMemcpyAsync(&c4, ctr, sizeof(Ctr), hipMemcpyDeviceToHost, st)) != hipSuccess) break;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) break;
        n_items = c4.next_front;
        ic ^= 1;
        ++levels;
        if ((c4.overflow & 4u) || c4.n4 > hc.inner || n_items > hc.inner || levels > kMaxTopLevels) {
            e = hipErrorUnknown;  // not a tree the collapse can take
            break;
        }
    }
    if (e == hipSuccess && (c4.overflow & 2u)) e = hipErrorInvalidValue;  // a box could not be quantized
    if (e != hipSuccess) {
        (void)hipFree(n4);
        (void)hipFree(prims);
        res->prims = nullptr;
        return e;
    }
    res->nodes4 = n4;
    res->n_nodes = c4.n4;
    res->depth = levels;
    return hipSuccess;
}

}  // namespace illume
