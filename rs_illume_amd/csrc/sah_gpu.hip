// Binned SAH BVH built on the device: BVHAccel::new with SplitMethod::SAH
// (accelerators/bvh.rs:32-247, recursive_build 95-247) under the host
// builder's split rule (bvh_builder.h: kSahBins bins on each axis, split cost
// kSahTraversalCost + (n0 A0 + n1 A1) / A against leaf cost n, count splits
// for coincident centroids, object medians from depth kSahMedianDepth on).
// Bins hold exact min/max boxes and integer counts, so every split decision is
// the host's, bit for bit, and the tree is the host tree: the same inner
// nodes with the same boxes, the same leaves (only the order of primitives
// inside a leaf may differ). The one exception is an object median of a node
// of more than kSmall primitives below depth kSahMedianDepth, split here by
// count in the current order (the host selects by centroid).
//
//   1. primitive boxes (one thread per primitive), the root's box and
//      centroid bounds;
//   2. top phase, one pass per level over nodes of more than kSmall
//      primitives: binning (block-level bins in LDS for the block's first
//      node, global atomics otherwise), one wave per node evaluating the 3 x 15
//      candidate splits, a flag per primitive, a device-wide exclusive scan,
//      and a stable scatter into the children's ranges that also reduces the
//      children's boxes and centroid bounds;
//   3. bottom phase: one wave per node of at most kSmall primitives builds
//      its whole subtree with the primitives in LDS (binning with LDS
//      atomics, partition by ballot), depth first;
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

constexpr uint32_t kNone = 0xffffffffu;
constexpr int kSmall = 128;  // nodes of at most this many primitives: one wave builds the subtree
constexpr int kStack = 64;   // pending nodes of a bottom-phase subtree (depth bound: 40 + 7 + 1)
constexpr int kBinSlots = 3 * kSahBins;
constexpr int kBinWords = 7;  // count, lo xyz, hi xyz (order-preserving encodings)
constexpr int kGBinWords = 8;
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

struct BNode {          // a node of the top phase or the root of a bottom-phase subtree
    uint32_t box[6];    // encoded lo xyz, hi xyz
    uint32_t cbox[6];   // encoded centroid bounds
    uint32_t start, count;
    int32_t parent;     // two-child node slot of the parent, -1 for the root
    uint32_t depth_side;  // depth << 1 | which child of the parent
};
static_assert(sizeof(BNode) == 64, "BNode is 64 bytes");

struct Split {          // decision for a top-phase node
    uint32_t mode;      // 0: by bin along axis, 1: first lcnt in the current order
    uint32_t axis, bin;
    float k, lo;        // bin = clamp((int)((c - lo) * k), 0, kSahBins - 1)
    uint32_t start, lcnt;
    uint32_t child[2];        // BNode ids
    uint32_t child_front[2];  // next frontier index of a child of more than kSmall primitives, else kNone
    uint32_t pad;
};

struct Ctr {
    uint32_t inner, bnodes, next_front, small, leaves, depth, overflow, n4;
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

__device__ __forceinline__ float centroid(const PBox &p, int a) {
    const float lo = a == 0 ? p.lo.x : (a == 1 ? p.lo.y : p.lo.z);
    const float hi = a == 0 ? p.hi.x : (a == 1 ? p.hi.y : p.hi.z);
    return lo * 0.5f + hi * 0.5f;
}
__device__ __forceinline__ int bin_of(float c, float lo, float k) {
    const int b = (int)((c - lo) * k);
    return min(max(b, 0), kSahBins - 1);
}

// Best split over bins (kBinSlots entries of `stride` words: count, lo xyz,
// hi xyz): lane 15 a + i evaluates the split after bin i on axis a, as the
// host loop does; the wave's minimum cost, ties to the lowest lane (the host
// keeps the first strictly smaller cost in axis-then-bin order).
struct Best {
    float cost;
    int lane;
    uint32_t lcnt;
};
__device__ __forceinline__ Best sah_best(const uint32_t *bins, int stride, const float *cb_lo, const float *cb_hi,
                                         float parent_area, int lane) {
    float cost = INFINITY;
    uint32_t lcnt = 0;
    if (lane < 3 * (kSahBins - 1)) {
        const int a = lane / (kSahBins - 1), i = lane % (kSahBins - 1);
        const float ext = cb_hi[a] - cb_lo[a];
        if (ext > 0) {
            float l[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
            float r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
            uint32_t rcnt = 0;
            for (int b = 0; b < kSahBins; ++b) {
                const uint32_t *e = bins + (a * kSahBins + b) * stride;
                const uint32_t cnt = e[0];
                if (cnt == 0) continue;
                float *acc = b <= i ? l : r;
                if (b <= i) lcnt += cnt;
                else rcnt += cnt;
                for (int x = 0; x < 3; ++x) {
                    acc[x] = fminf(acc[x], dec(e[1 + x]));
                    acc[3 + x] = fmaxf(acc[3 + x], dec(e[4 + x]));
                }
            }
            if (lcnt != 0 && rcnt != 0)
                cost = kSahTraversalCost + ((float)lcnt * sah_area(l[3] - l[0], l[4] - l[1], l[5] - l[2]) +
                                            (float)rcnt * sah_area(r[3] - r[0], r[4] - r[1], r[5] - r[2])) /
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
    return b;  // b.cost == INFINITY: no candidate
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
                                                   PBox *__restrict__ pb, uint32_t *__restrict__ node_of,
                                                   uint32_t root_front, BNode *__restrict__ root) {
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
        node_of[p] = root_front;
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

__global__ void init_root_kernel(BNode *root, uint32_t n, uint32_t *front, uint32_t *small_list, Ctr *ctr) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int a = 0; a < 3; ++a) {
        root->box[a] = root->cbox[a] = kEncPosInf;
        root->box[3 + a] = root->cbox[3 + a] = kEncNegInf;
    }
    root->start = 0;
    root->count = n;
    root->parent = -1;
    root->depth_side = 0;
    Ctr c{0, 1, 0, 0, 0, 0, 0, 0};
    if (n > (uint32_t)kSmall) {
        front[0] = 0;
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
    uint32_t *e = bins + (size_t)i * kGBinWords;
    e[0] = 0;
    e[1] = e[2] = e[3] = kEncPosInf;
    e[4] = e[5] = e[6] = kEncNegInf;
    e[7] = 0;
}

// ---- 2. top phase
__global__ __launch_bounds__(256) void bin_kernel(const PBox *__restrict__ pb, const uint32_t *__restrict__ node_of,
                                                  uint32_t n, const uint32_t *__restrict__ front,
                                                  const BNode *__restrict__ bnodes, uint32_t *__restrict__ bins) {
    __shared__ uint32_t lb[kBinSlots * kBinWords];
    const uint32_t first = blockIdx.x * 256;
    const uint32_t f0 = node_of[first];
    for (int i = threadIdx.x; i < kBinSlots; i += 256) {
        lb[i * kBinWords] = 0;
        for (int x = 0; x < 3; ++x) {
            lb[i * kBinWords + 1 + x] = kEncPosInf;
            lb[i * kBinWords + 4 + x] = kEncNegInf;
        }
    }
    __syncthreads();
    const uint32_t pos = first + threadIdx.x;
    const uint32_t f = pos < n ? node_of[pos] : kNone;
    if (f != kNone) {
        const BNode &B = bnodes[front[f]];
        const PBox p = pb[pos];
        const uint32_t e[6] = {enc(p.lo.x), enc(p.lo.y), enc(p.lo.z), enc(p.hi.x), enc(p.hi.y), enc(p.hi.z)};
        for (int a = 0; a < 3; ++a) {
            const float lo = dec(B.cbox[a]), ext = dec(B.cbox[3 + a]) - lo;
            if (!(ext > 0)) continue;
            const float k = (float)kSahBins / ext;
            const int b = bin_of(centroid(p, a), lo, k);
            uint32_t *dst = f == f0 ? lb + (a * kSahBins + b) * kBinWords
                                    : bins + ((size_t)f * kBinSlots + a * kSahBins + b) * kGBinWords;
            atomicAdd(dst, 1u);
            for (int x = 0; x < 3; ++x) {
                atomicMin(dst + 1 + x, e[x]);
                atomicMax(dst + 4 + x, e[3 + x]);
            }
        }
    }
    __syncthreads();
    if (f0 != kNone && threadIdx.x < kBinSlots) {
        const uint32_t *src = lb + threadIdx.x * kBinWords;
        if (src[0]) {
            uint32_t *dst = bins + ((size_t)f0 * kBinSlots + threadIdx.x) * kGBinWords;
            atomicAdd(dst, src[0]);
            for (int x = 0; x < 3; ++x) {
                atomicMin(dst + 1 + x, src[1 + x]);
                atomicMax(dst + 4 + x, src[4 + x]);
            }
        }
    }
}

// One wave per top-phase node: the split, the node's slot and its entry in
// the parent, the two children's records.
__global__ __launch_bounds__(256) void split_kernel(const uint32_t *__restrict__ front, uint32_t nf,
                                                    BNode *__restrict__ bnodes, const uint32_t *__restrict__ bins,
                                                    Split *__restrict__ split, uint32_t *__restrict__ next_front,
                                                    uint32_t *__restrict__ small_list, GpuNode *__restrict__ nodes,
                                                    Ctr *__restrict__ ctr) {
    const uint32_t f = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (f >= nf) return;
    const uint32_t id = front[f];
    const BNode B = bnodes[id];
    float box[6], cb_lo[3], cb_hi[3];
    for (int a = 0; a < 3; ++a) {
        box[a] = dec(B.box[a]);
        box[3 + a] = dec(B.box[3 + a]);
        cb_lo[a] = dec(B.cbox[a]);
        cb_hi[a] = dec(B.cbox[3 + a]);
    }
    const uint32_t n = B.count, depth = B.depth_side >> 1;
    const float parent_area = sah_area(box[3] - box[0], box[4] - box[1], box[5] - box[2]);
    const Best best = sah_best(bins + (size_t)f * kBinSlots * kGBinWords, kGBinWords, cb_lo, cb_hi, parent_area, lane);
    if (lane != 0) return;
    Split s;
    s.start = B.start;
    s.pad = 0;
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
        s.lo = cb_lo[s.axis];
        s.k = (float)kSahBins / (cb_hi[s.axis] - cb_lo[s.axis]);
        s.lcnt = best.lcnt;
    }
    const int32_t slot = (int32_t)atomicAdd(&ctr->inner, 1u);
    if (B.parent >= 0) set_child(nodes, B.parent, B.depth_side & 1u, box, slot);
    const uint32_t c0 = atomicAdd(&ctr->bnodes, 2u);
    for (uint32_t k = 0; k < 2; ++k) {
        BNode c;
        for (int a = 0; a < 3; ++a) {
            c.box[a] = c.cbox[a] = kEncPosInf;
            c.box[3 + a] = c.cbox[3 + a] = kEncNegInf;
        }
        c.start = k ? s.start + s.lcnt : s.start;
        c.count = k ? n - s.lcnt : s.lcnt;
        c.parent = slot;
        c.depth_side = (depth + 1) << 1 | k;
        bnodes[c0 + k] = c;
        s.child[k] = c0 + k;
        if (c.count > (uint32_t)kSmall) {
            const uint32_t cf = atomicAdd(&ctr->next_front, 1u);
            next_front[cf] = c0 + k;
            s.child_front[k] = cf;
        } else {
            small_list[atomicAdd(&ctr->small, 1u)] = c0 + k;
            s.child_front[k] = kNone;
        }
    }
    split[f] = s;
}

__device__ __forceinline__ uint32_t goes_left(const Split &s, const PBox &p, uint32_t pos) {
    if (s.mode == 0) return bin_of(centroid(p, (int)s.axis), s.lo, s.k) <= (int)s.bin ? 1u : 0u;
    return pos - s.start < s.lcnt ? 1u : 0u;
}

__global__ __launch_bounds__(256) void flag_kernel(const PBox *__restrict__ pb, const uint32_t *__restrict__ node_of,
                                                   uint32_t n, const Split *__restrict__ split,
                                                   uint32_t *__restrict__ flags) {
    const uint32_t pos = blockIdx.x * 256 + threadIdx.x;
    if (pos >= n) return;
    const uint32_t f = node_of[pos];
    flags[pos] = f == kNone ? 0u : goes_left(split[f], pb[pos], pos);
}

// Stable partition of every top-phase node's range by its flags; primitives of
// other nodes keep their positions. Children of more than kSmall primitives
// get their boxes and centroid bounds here (wave-reduced atomics).
__global__ __launch_bounds__(256) void scatter_kernel(const PBox *__restrict__ pb, const uint32_t *__restrict__ node_of,
                                                      uint32_t n, const Split *__restrict__ split,
                                                      const uint32_t *__restrict__ flags,
                                                      const uint32_t *__restrict__ excl, PBox *__restrict__ pb2,
                                                      uint32_t *__restrict__ node_of2, BNode *__restrict__ bnodes) {
    const uint32_t pos = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x % 64;
    uint32_t child = kNone;
    PBox p;
    p.lo = p.hi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (pos < n) {
        const uint32_t f = node_of[pos];
        p = pb[pos];
        uint32_t np = pos, cf = kNone;
        if (f != kNone) {
            const Split s = split[f];
            const uint32_t before = excl[pos] - excl[s.start];
            const uint32_t k = flags[pos] ? 0u : 1u;
            np = k == 0 ? s.start + before : s.start + s.lcnt + (pos - s.start - before);
            cf = s.child_front[k];
            if (cf != kNone) child = s.child[k];
        }
        pb2[np] = p;
        node_of2[np] = cf;
    }
    uint64_t pending = __ballot(child != kNone);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t c = __shfl(child, leader);
        const uint64_t mask = __ballot(child == c);
        const bool in = child == c;
        float v[12];
        for (int a = 0; a < 3; ++a) {
            const float lo = a == 0 ? p.lo.x : (a == 1 ? p.lo.y : p.lo.z);
            const float hi = a == 0 ? p.hi.x : (a == 1 ? p.hi.y : p.hi.z);
            const float ce = lo * 0.5f + hi * 0.5f;
            v[a] = in ? lo : INFINITY;
            v[3 + a] = in ? ce : INFINITY;
            v[6 + a] = in ? hi : -INFINITY;
            v[9 + a] = in ? ce : -INFINITY;
        }
        for (int a = 0; a < 6; ++a) {
            v[a] = wave_min(v[a]);
            v[6 + a] = wave_max(v[6 + a]);
        }
        if (lane == leader) {
            BNode &B = bnodes[c];
            for (int a = 0; a < 3; ++a) {
                atomicMin(&B.box[a], enc(v[a]));
                atomicMin(&B.cbox[a], enc(v[3 + a]));
                atomicMax(&B.box[3 + a], enc(v[6 + a]));
                atomicMax(&B.cbox[3 + a], enc(v[9 + a]));
            }
        }
        pending &= ~mask;
    }
}

// ---- 3. bottom phase: one wave per subtree root of at most kSmall primitives
struct SmallLds {
    float4 lo[kSmall], hi[kSmall];
    uint32_t bins[kBinSlots * kBinWords];
    uint32_t st_range[kStack];  // local start << 16 | count
    int32_t st_parent[kStack];
    uint32_t st_ds[kStack];     // depth << 1 | side
};

__global__ __launch_bounds__(256) void small_kernel(PBox *__restrict__ pb, const uint32_t *__restrict__ small_list,
                                                    uint32_t n_small, const BNode *__restrict__ bnodes,
                                                    GpuNode *__restrict__ nodes, Ctr *__restrict__ ctr,
                                                    int max_prims) {
    __shared__ SmallLds lds[4];
    const uint32_t item = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (item >= n_small) return;
    SmallLds &L = lds[threadIdx.x / 64];
    const BNode R = bnodes[small_list[item]];
    const uint32_t s0 = R.start, m0 = R.count;
    for (uint32_t j = lane; j < m0; j += 64) {
        const PBox p = pb[s0 + j];
        L.lo[j] = p.lo;
        L.hi[j] = p.hi;
    }
    int sp = 0;
    if (lane == 0) {
        L.st_range[0] = m0;
        L.st_parent[0] = R.parent;
        L.st_ds[0] = R.depth_side;
    }
    sp = 1;
    wave_lds_sync();
    uint32_t leaves = 0, max_depth = 0;
    bool overflow = false;
    while (sp > 0 && !overflow) {
        --sp;
        uint32_t s = L.st_range[sp] >> 16, m = L.st_range[sp] & 0xffffu;
        int32_t parent = L.st_parent[sp];
        uint32_t ds = L.st_ds[sp];
        wave_lds_sync();
        for (;;) {
            const uint32_t depth = ds >> 1;
            // this lane's primitives: local s + lane and s + 64 + lane
            PBox p[2];
            bool valid[2];
            float v[12];
            for (int a = 0; a < 6; ++a) v[a] = INFINITY, v[6 + a] = -INFINITY;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t j = (uint32_t)lane + 64u * h;
                valid[h] = j < m;
                if (valid[h]) {
                    p[h].lo = L.lo[s + j];
                    p[h].hi = L.hi[s + j];
                    const float lo[3] = {p[h].lo.x, p[h].lo.y, p[h].lo.z}, hi[3] = {p[h].hi.x, p[h].hi.y, p[h].hi.z};
                    for (int a = 0; a < 3; ++a) {
                        const float c = lo[a] * 0.5f + hi[a] * 0.5f;
                        v[a] = fminf(v[a], lo[a]);
                        v[3 + a] = fminf(v[3 + a], c);
                        v[6 + a] = fmaxf(v[6 + a], hi[a]);
                        v[9 + a] = fmaxf(v[9 + a], c);
                    }
                } else {
                    p[h].lo = p[h].hi = make_float4(0.f, 0.f, 0.f, 0.f);
                }
            }
            for (int a = 0; a < 6; ++a) {
                v[a] = wave_min(v[a]);
                v[6 + a] = wave_max(v[6 + a]);
            }
            const float box[6] = {v[0], v[1], v[2], v[6], v[7], v[8]};
            const float cb_lo[3] = {v[3], v[4], v[5]}, cb_hi[3] = {v[9], v[10], v[11]};
            // decision, in the host builder's order
            int mode = -1;  // -1 leaf, 0 bins, 1 first half, 2 object median
            int axis = 0, bin = 0;
            float k = 0.f, lo_ax = 0.f;
            if (m == 1) {
                mode = -1;
            } else if (depth >= (uint32_t)kSahMedianDepth) {
                mode = m <= (uint32_t)max_prims ? -1 : 2;
            } else {
                for (int i = lane; i < kBinSlots; i += 64) {
                    L.bins[i * kBinWords] = 0;
                    for (int x = 0; x < 3; ++x) {
                        L.bins[i * kBinWords + 1 + x] = kEncPosInf;
                        L.bins[i * kBinWords + 4 + x] = kEncNegInf;
                    }
                }
                wave_lds_sync();
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (!valid[h]) continue;
                    const uint32_t e[6] = {enc(p[h].lo.x), enc(p[h].lo.y), enc(p[h].lo.z),
                                           enc(p[h].hi.x), enc(p[h].hi.y), enc(p[h].hi.z)};
                    for (int a = 0; a < 3; ++a) {
                        const float ext = cb_hi[a] - cb_lo[a];
                        if (!(ext > 0)) continue;
                        const int b = bin_of(centroid(p[h], a), cb_lo[a], (float)kSahBins / ext);
                        uint32_t *dst = L.bins + (a * kSahBins + b) * kBinWords;
                        atomicAdd(dst, 1u);
                        for (int x = 0; x < 3; ++x) {
                            atomicMin(dst + 1 + x, e[x]);
                            atomicMax(dst + 4 + x, e[3 + x]);
                        }
                    }
                }
                wave_lds_sync();
                const float parent_area = sah_area(box[3] - box[0], box[4] - box[1], box[5] - box[2]);
                const Best best = sah_best(L.bins, kBinWords, cb_lo, cb_hi, parent_area, lane);
                wave_lds_sync();
                if (best.cost == INFINITY) {
                    mode = m <= (uint32_t)max_prims ? -1 : 1;
                } else if (m <= (uint32_t)max_prims && best.cost >= (float)m) {
                    mode = -1;
                } else {
                    mode = 0;
                    axis = best.lane / (kSahBins - 1);
                    bin = best.lane % (kSahBins - 1);
                    lo_ax = cb_lo[axis];
                    k = (float)kSahBins / (cb_hi[axis] - cb_lo[axis]);
                }
            }
            if (mode < 0) {
                if (lane == 0 && parent >= 0) set_child(nodes, parent, ds & 1u, box, make_leaf_ref(s0 + s, m));
                ++leaves;
                max_depth = max(max_depth, depth);
                break;
            }
            int32_t slot = 0;
            if (lane == 0) {
                slot = (int32_t)atomicAdd(&ctr->inner, 1u);
                if (parent >= 0) set_child(nodes, parent, ds & 1u, box, slot);
            }
            slot = __shfl(slot, 0);
            // flags
            bool left[2] = {false, false};
            int med_axis = 0;
            if (mode == 2) {
                float ext = -1.f;
                for (int a = 0; a < 3; ++a) {
                    const float x = cb_hi[a] - cb_lo[a];
                    if (x > ext) {
                        ext = x;
                        med_axis = a;
                    }
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
                    // object median: rank by centroid, ties by position
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
            if (lane == 0) {
                L.st_range[sp] = (s + nl) << 16 | (m - nl);
                L.st_parent[sp] = slot;
                L.st_ds[sp] = (depth + 1) << 1 | 1u;
            }
            ++sp;
            wave_lds_sync();
            m = nl;
            parent = slot;
            ds = (depth + 1) << 1;
        }
    }
    for (uint32_t j = lane; j < m0; j += 64) {
        PBox p;
        p.lo = L.lo[j];
        p.hi = L.hi[j];
        pb[s0 + j] = p;
    }
    if (lane == 0) {
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
// cap bounds the items, the next items and the four-child slots (the binary
// tree's inner-node count); a malformed tree sets overflow bit 4 instead of
// writing past it.
__global__ __launch_bounds__(256) void collapse_kernel(const GpuNode *__restrict__ n2, const uint2 *__restrict__ items,
                                                       uint32_t n_items, uint2 *__restrict__ next,
                                                       GpuNode4Q *__restrict__ out, uint32_t cap,
                                                       Ctr *__restrict__ ctr) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n_items) return;
    const uint2 it = items[i];
    Kid kids[4];
    int nk = 2;
    if (it.x >= cap || it.y >= cap) {
        atomicOr(&ctr->overflow, 4u);
        return;
    }
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
        if ((uint32_t)kids[best].ref >= cap) {
            atomicOr(&ctr->overflow, 4u);
            return;
        }
        const GpuNode g = n2[kids[best].ref];
        kids[best] = kid_of(g, 0);
        kids[nk++] = kid_of(g, 1);
    }
    uint32_t n_inner = 0;
    for (int c = 0; c < nk; ++c) n_inner += kids[c].ref >= 0 ? 1u : 0u;
    const uint32_t base = n_inner ? atomicAdd(&ctr->n4, n_inner) : 0u;
    const uint32_t nbase = n_inner ? atomicAdd(&ctr->next_front, n_inner) : 0u;
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
            if (nbase + used < cap) next[nbase + used] = make_uint2((uint32_t)k.ref, base + used);
            else atomicOr(&ctr->overflow, 4u);
            ++used;
        }
    }
    GpuNode4Q q;
    if (!quantize_node(g4, &q)) atomicOr(&ctr->overflow, 2u);
    out[it.y] = q;
}

__global__ void set_u32_kernel(uint32_t *p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// The collapse's first item (the root, slot 0) and its counters.
__global__ void collapse_init_kernel(uint2 *items, Ctr *ctr) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    items[0] = make_uint2(0u, 0u);
    Ctr c{0, 0, 0, 0, 0, 0, 0, 1};
    *ctr = c;
}

struct DevBuf {
    std::vector<void *> ptrs;
    ~DevBuf() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    hipError_t alloc(T **p, size_t count) {
        hipError_t e = hipMalloc(p, std::max<size_t>(count, 1) * sizeof(T));
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
    DevBuf tmp;
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
    uint32_t *node_of[2], *flags, *excl, *front[2], *small_list, *bins;
    Split *split;
    Ctr *ctr;
    GpuNode *n2;
    if ((e = tmp.alloc(&pb[0], n)) != hipSuccess || (e = tmp.alloc(&pb[1], n)) != hipSuccess ||
        (e = tmp.alloc(&node_of[0], n)) != hipSuccess || (e = tmp.alloc(&node_of[1], n)) != hipSuccess ||
        (e = tmp.alloc(&flags, n)) != hipSuccess || (e = tmp.alloc(&excl, n)) != hipSuccess ||
        (e = tmp.alloc(&front[0], front_cap)) != hipSuccess || (e = tmp.alloc(&front[1], front_cap)) != hipSuccess ||
        (e = tmp.alloc(&small_list, n)) != hipSuccess ||
        (e = tmp.alloc(&bins, (size_t)front_cap * kBinSlots * kGBinWords)) != hipSuccess ||
        (e = tmp.alloc(&split, front_cap)) != hipSuccess || (e = tmp.alloc(&ctr, 1)) != hipSuccess)
        return e;
    // BNodes: the root, then two per top-phase split; grown on demand
    size_t bnode_cap = 1 + 2 * (size_t)front_cap * 4;
    BNode *bnodes = nullptr;
    if ((e = hipMalloc(&bnodes, bnode_cap * sizeof(BNode))) != hipSuccess) return e;
    struct BNodeGuard {
        BNode **p;
        ~BNodeGuard() { (void)hipFree(*p); }
    } bguard{&bnodes};
    if ((e = hipMalloc(&n2, (size_t)std::max<uint32_t>(n - 1, 1) * sizeof(GpuNode))) != hipSuccess) return e;
    struct NodeGuard {
        GpuNode **p;
        bool keep = false;
        ~NodeGuard() {
            if (!keep) (void)hipFree(*p);
        }
    } nguard{&n2};
    if ((e = hipMemsetAsync(n2, 0, (size_t)std::max<uint32_t>(n - 1, 1) * sizeof(GpuNode), st)) != hipSuccess) return e;
    size_t scan_bytes = 0;
    if ((e = rocprim::exclusive_scan(nullptr, scan_bytes, flags, excl, 0u, n, rocprim::plus<uint32_t>(), st)) !=
        hipSuccess)
        return e;
    void *scan_tmp = nullptr;
    if ((e = tmp.alloc((char **)&scan_tmp, scan_bytes)) != hipSuccess) return e;

    init_root_kernel<<<1, 64, 0, st>>>(bnodes, n, front[0], small_list, ctr);
    prim_kernel<<<blocks_for(n, 256), 256, 0, st>>>(d_verts, d_idx, n_tris, d_sph, n, pb[0], node_of[0],
                                                   n > (uint32_t)kSmall ? 0u : kNone, bnodes);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    Ctr hc;
    if ((e = hipMemcpyAsync(&hc, ctr, sizeof(Ctr), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    int cur = 0;
    uint32_t nf = hc.next_front;
    for (int level = 0; nf > 0; ++level) {
        if (level >= kMaxTopLevels) return hipErrorUnknown;
        // room for this level's children
        if (hc.bnodes + 2 * (size_t)nf > bnode_cap) {
            const size_t cap2 = std::max(bnode_cap * 2, hc.bnodes + 2 * (size_t)nf);
            BNode *b2 = nullptr;
            if ((e = hipMalloc(&b2, cap2 * sizeof(BNode))) != hipSuccess) return e;
            if ((e = hipMemcpyAsync(b2, bnodes, hc.bnodes * sizeof(BNode), hipMemcpyDeviceToDevice, st)) != hipSuccess ||
                (e = hipStreamSynchronize(st)) != hipSuccess) {
                (void)hipFree(b2);
                return e;
            }
            (void)hipFree(bnodes);
            bnodes = b2;
            bnode_cap = cap2;
        }
        set_u32_kernel<<<1, 64, 0, st>>>(&ctr->next_front, 0u);
        clear_bins_kernel<<<blocks_for((uint64_t)nf * kBinSlots, 256), 256, 0, st>>>(bins, nf * kBinSlots);
        bin_kernel<<<blocks_for(n, 256), 256, 0, st>>>(pb[cur], node_of[cur], n, front[cur], bnodes, bins);
        split_kernel<<<blocks_for(nf, 4), 256, 0, st>>>(front[cur], nf, bnodes, bins, split, front[cur ^ 1],
                                                       small_list, n2, ctr);
        flag_kernel<<<blocks_for(n, 256), 256, 0, st>>>(pb[cur], node_of[cur], n, split, flags);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = rocprim::exclusive_scan(scan_tmp, scan_bytes, flags, excl, 0u, n, rocprim::plus<uint32_t>(), st)) !=
            hipSuccess)
            return e;
        scatter_kernel<<<blocks_for(n, 256), 256, 0, st>>>(pb[cur], node_of[cur], n, split, flags, excl,
                                                          pb[cur ^ 1], node_of[cur ^ 1], bnodes);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(&hc, ctr, sizeof(Ctr), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        cur ^= 1;
        nf = hc.next_front;
        res->top_levels = level + 1;
    }
    if (hc.small > 0)
        small_kernel<<<blocks_for(hc.small, 4), 256, 0, st>>>(pb[cur], small_list, hc.small, bnodes, n2, ctr,
                                                              max_prims);
    PrimRecord *prims = nullptr;
    if ((e = hipMalloc(&prims, (size_t)n * sizeof(PrimRecord))) != hipSuccess) return e;
    records_kernel<<<blocks_for(n, 256), 256, 0, st>>>(d_verts, d_idx, n_tris, d_sph, n, pb[cur], prims);
    if ((e = hipGetLastError()) == hipSuccess) e = hipMemcpyAsync(&hc, ctr, sizeof(Ctr), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && (hc.overflow || hc.inner != hc.leaves - 1)) e = hipErrorUnknown;
    // the root's box
    if (e == hipSuccess) {
        uint32_t rb[6];
        e = hipMemcpy(rb, bnodes, sizeof(rb), hipMemcpyDeviceToHost);
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
        nguard.keep = true;
        res->nodes = n2;
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
                                                                 hc.inner, ctr);
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
