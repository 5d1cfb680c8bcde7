// BVH construction on the device: a linear BVH (Karras, "Maximizing
// parallelism in the construction of BVHs, octrees, and k-d trees", HPG 2012)
// over 63-bit Morton codes of the primitive centroids. The replacement for
// BVHAccel::new's recursive build (accelerators/bvh.rs:32-93; the reference's
// own HLBVH, bvh.rs:249-502, also starts from Morton codes) when build time
// matters more than tree quality: the closest hit does not depend on the
// tree, so both builders give the same hits.
//
//   1. primitive boxes and centroids (one thread per primitive),
//   2. centroid bounds (block reduction, then one block),
//   3. Morton codes, radix sort of (code, primitive) pairs (rocPRIM),
//   4. hierarchy: internal node i finds its key range and split by binary
//      search over common-prefix lengths (one thread per internal node),
//   5. boxes bottom-up: one thread per leaf climbs while it is the second
//      child to arrive at a node (device-scope fences between the writes of
//      the first arrival and the reads of the second),
//   6. emission of the two-child node layout: subtrees of at most
//      max_prims primitives become leaves (their primitives are contiguous in
//      Morton order), and the leaf-ordered primitive records.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <cmath>
#include <vector>
#include "layout.h"
#include "lbvh.h"

namespace illume {
namespace {

constexpr uint32_t kLeafBit = 0x80000000u;  // child encoding before emission: leaf k = k | kLeafBit
// Climbs stop after this many levels (a tree that deep is rejected by the
// depth check anyway; the bound keeps every loop finite).
constexpr int kMaxLevels = 1024;

__global__ void prim_box_kernel(const float *__restrict__ verts, const uint32_t *__restrict__ idx, uint32_t n_tris,
                                const float *__restrict__ spheres, uint32_t n, float *__restrict__ box,
                                float *__restrict__ cen) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float lo[3], hi[3];
    if (p < n_tris) {
        for (int a = 0; a < 3; ++a) lo[a] = hi[a] = verts[3 * (size_t)idx[3 * (size_t)p] + a];
        for (int k = 1; k < 3; ++k)
            for (int a = 0; a < 3; ++a) {
                const float v = verts[3 * (size_t)idx[3 * (size_t)p + k] + a];
                lo[a] = fminf(lo[a], v);
                hi[a] = fmaxf(hi[a], v);
            }
    } else {
        // one ulp outward, as the host builder does
        const float *s = spheres + 4 * (size_t)(p - n_tris);
        for (int a = 0; a < 3; ++a) {
            lo[a] = nextafterf(s[a] + (-s[3]), -INFINITY);
            hi[a] = nextafterf(s[a] + s[3], INFINITY);
        }
    }
    for (int a = 0; a < 3; ++a) {
        box[6 * (size_t)p + a] = lo[a];
        box[6 * (size_t)p + 3 + a] = hi[a];
        cen[3 * (size_t)p + a] = lo[a] * 0.5f + hi[a] * 0.5f;
    }
}

// Per-block min/max of `vals` ([n][dim], dim 3 or 6 with the lower half
// minimised and the upper half maximised for boxes) -> partial[block][6].
__global__ __launch_bounds__(256) void bounds_partial_kernel(const float *__restrict__ vals, int dim, uint32_t n,
                                                             float *__restrict__ partial) {
    __shared__ float s[6][256];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], vals[(size_t)i * dim + a]);
            hi[a] = fmaxf(hi[a], vals[(size_t)i * dim + (dim == 6 ? 3 : 0) + a]);
        }
    for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = lo[a];
        s[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int a = 0; a < 3; ++a) {
                s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + w]);
                s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void bounds_final_kernel(const float *__restrict__ partial, uint32_t blocks, float *__restrict__ out) {
    if (threadIdx.x != 0) return;
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t k = 0; k < blocks; ++k)
        for (int a = 0; a < 3; ++a) {
            b[a] = fminf(b[a], partial[6 * k + a]);
            b[3 + a] = fmaxf(b[3 + a], partial[6 * k + 3 + a]);
        }
    for (int a = 0; a < 6; ++a) out[a] = b[a];
}

__device__ __forceinline__ uint64_t spread3(uint32_t x) {  // 21 bits -> every third bit of 63
    uint64_t v = x & 0x1fffffu;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

__global__ void morton_kernel(const float *__restrict__ cen, uint32_t n, const float *__restrict__ cb,
                              uint64_t *__restrict__ keys, uint32_t *__restrict__ ids) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    // one scale for all axes (the largest centroid extent): per-axis scales
    // would stretch flat scenes and split their thin axis too early
    const float ext = fmaxf(fmaxf(cb[3] - cb[0], cb[4] - cb[1]), cb[5] - cb[2]);
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
        const float u = ext > 0.f ? (cen[3 * (size_t)p + a] - cb[a]) / ext : 0.f;
        q[a] = (uint32_t)fminf(fmaxf(u * 2097152.f, 0.f), 2097151.f);
    }
    keys[p] = spread3(q[0]) << 2 | spread3(q[1]) << 1 | spread3(q[2]);
    ids[p] = p;
}

// Common-prefix length of sorted keys i and j; ties broken by the index so
// that every key is distinct; -1 outside [0, n).
__device__ __forceinline__ int delta(const uint64_t *__restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = k[i], b = k[j];
    if (a != b) return __clzll((long long)(a ^ b));
    return 64 + __clz((int)((uint32_t)i ^ (uint32_t)j));
}

__global__ void hierarchy_kernel(const uint64_t *__restrict__ keys, int n, uint32_t *__restrict__ left,
                                 uint32_t *__restrict__ right, uint32_t *__restrict__ first,
                                 uint32_t *__restrict__ last, int32_t *__restrict__ parent_inner,
                                 int32_t *__restrict__ parent_leaf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1) > 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    for (int t = (l + 1) / 2, step = 0; step < 64; t = (t + 1) / 2, ++step) {  // t = ceil(l/2), ceil(l/4), .., 1
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int lo = min(i, j), hi = max(i, j);
    const uint32_t lc = lo == gamma ? ((uint32_t)gamma | kLeafBit) : (uint32_t)gamma;
    const uint32_t rc = hi == gamma + 1 ? ((uint32_t)(gamma + 1) | kLeafBit) : (uint32_t)(gamma + 1);
    left[i] = lc;
    right[i] = rc;
    first[i] = (uint32_t)lo;
    last[i] = (uint32_t)hi;
    if (lc & kLeafBit) parent_leaf[gamma] = i;
    else parent_inner[gamma] = i;
    if (rc & kLeafBit) parent_leaf[gamma + 1] = i;
    else parent_inner[gamma + 1] = i;
}

__device__ __forceinline__ void child_box(uint32_t c, const float *__restrict__ pbox, const uint32_t *__restrict__ ids,
                                          const float *nbox, float *b) {
    const float *src = (c & kLeafBit) ? pbox + 6 * (size_t)ids[c & ~kLeafBit] : nbox + 6 * (size_t)c;
    for (int a = 0; a < 6; ++a) b[a] = src[a];
}

// One thread per leaf climbs towards the root; at every node only the
// second arrival continues, after the first has published its subtree box.
__global__ void refit_kernel(int n, const uint32_t *__restrict__ left, const uint32_t *__restrict__ right,
                             const int32_t *__restrict__ parent_inner, const int32_t *__restrict__ parent_leaf,
                             const float *__restrict__ pbox, const uint32_t *__restrict__ ids, float *nbox,
                             uint32_t *visits) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int node = parent_leaf[k];
    for (int level = 0; node >= 0 && level < kMaxLevels; ++level) {
        __threadfence();
        if (atomicAdd(&visits[node], 1u) == 0u) return;
        __threadfence();
        float a[6], b[6];
        child_box(left[node], pbox, ids, nbox, a);
        child_box(right[node], pbox, ids, nbox, b);
        for (int x = 0; x < 3; ++x) {
            nbox[6 * (size_t)node + x] = fminf(a[x], b[x]);
            nbox[6 * (size_t)node + 3 + x] = fmaxf(a[3 + x], b[3 + x]);
        }
        node = parent_inner[node];
    }
}

// Depth of each leaf (inner nodes on its path) -> max.
__global__ void depth_kernel(int n, const int32_t *__restrict__ parent_inner, const int32_t *__restrict__ parent_leaf,
                             uint32_t *__restrict__ max_depth) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t d = 0;
    for (int node = parent_leaf[k]; node >= 0 && d < (uint32_t)kMaxLevels; node = parent_inner[node]) ++d;
    atomicMax(max_depth, d);
}

__global__ void emit_kernel(int n, int max_prims, const uint32_t *__restrict__ left, const uint32_t *__restrict__ right,
                            const uint32_t *__restrict__ first, const uint32_t *__restrict__ last,
                            const float *__restrict__ pbox, const uint32_t *__restrict__ ids,
                            const float *__restrict__ nbox, GpuNode *__restrict__ nodes, uint32_t *__restrict__ n_leaves) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const uint32_t cs[2] = {left[i], right[i]};
    int32_t refs[2];
    float b[2][6];
    uint32_t leaves = 0;
    for (int c = 0; c < 2; ++c) {
        const uint32_t ch = cs[c];
        child_box(ch, pbox, ids, nbox, b[c]);
        if (ch & kLeafBit) {
            refs[c] = make_leaf_ref(ch & ~kLeafBit, 1);
            ++leaves;
        } else if (last[ch] - first[ch] + 1 <= (uint32_t)max_prims) {
            refs[c] = make_leaf_ref(first[ch], last[ch] - first[ch] + 1);
            ++leaves;
        } else {
            refs[c] = (int32_t)ch;
        }
    }
    GpuNode g;
    g.c0_lo_x = b[0][0]; g.c0_hi_x = b[0][3]; g.c0_lo_y = b[0][1]; g.c0_hi_y = b[0][4];
    g.c1_lo_x = b[1][0]; g.c1_hi_x = b[1][3]; g.c1_lo_y = b[1][1]; g.c1_hi_y = b[1][4];
    g.c0_lo_z = b[0][2]; g.c0_hi_z = b[0][5]; g.c1_lo_z = b[1][2]; g.c1_hi_z = b[1][5];
    g.child0 = refs[0];
    g.child1 = refs[1];
    g.pad0 = g.pad1 = 0;
    nodes[i] = g;
    // inner nodes inside a collapsed subtree are never reached; count only
    // the leaves of reachable nodes (the root, or count > max_prims)
    if (leaves && (i == 0 || last[i] - first[i] + 1 > (uint32_t)max_prims)) atomicAdd(n_leaves, leaves);
}

// Node i spans more than one treelet: the codes at both ends of its sorted
// range differ within the top kTreeletBits of the 63-bit code.
__global__ void upper_kernel(int n, const uint64_t *__restrict__ keys, const uint32_t *__restrict__ first,
                             const uint32_t *__restrict__ last, uint8_t *__restrict__ upper) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    upper[i] = ((keys[first[i]] ^ keys[last[i]]) >> (63 - kTreeletBits)) != 0 ? 1 : 0;
}

__global__ void records_kernel(const float *__restrict__ verts, const uint32_t *__restrict__ idx, uint32_t n_tris,
                               const float *__restrict__ spheres, uint32_t n, const uint32_t *__restrict__ ids,
                               PrimRecord *__restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t p = ids[k];
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

struct DevBuf {
    std::vector<void *> ptrs;
    ~DevBuf() {
        for (void *p : ptrs) hipFree(p);
    }
    template <typename T>
    hipError_t alloc(T **p, size_t count) {
        hipError_t e = hipMalloc(p, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

}  // namespace

hipError_t lbvh_build(const float *verts, uint32_t n_verts, const uint32_t *idx, uint32_t n_tris,
                      const float *spheres, uint32_t n_spheres, int max_prims, hipStream_t st, LbvhResult *res) {
    const uint32_t n = n_tris + n_spheres;
    res->nodes = nullptr;
    res->prims = nullptr;
    DevBuf tmp;
    float *d_verts = nullptr, *d_sph = nullptr, *pbox, *cen, *partial, *cb, *sb, *nbox;
    uint32_t *d_idx = nullptr, *ids_in, *ids, *left, *right, *first, *last, *visits, *counters;
    uint8_t *upper;
    uint64_t *keys_in, *keys;
    int32_t *parent_inner, *parent_leaf;
    hipError_t e = hipSuccess;
    const uint32_t blocks = (n + 255) / 256, rblocks = std::min<uint32_t>(blocks, 1024);
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
    if ((e = tmp.alloc(&pbox, 6 * (size_t)n)) != hipSuccess || (e = tmp.alloc(&cen, 3 * (size_t)n)) != hipSuccess ||
        (e = tmp.alloc(&partial, 6 * (size_t)rblocks)) != hipSuccess || (e = tmp.alloc(&cb, 6)) != hipSuccess ||
        (e = tmp.alloc(&sb, 6)) != hipSuccess || (e = tmp.alloc(&keys_in, n)) != hipSuccess ||
        (e = tmp.alloc(&keys, n)) != hipSuccess || (e = tmp.alloc(&ids_in, n)) != hipSuccess ||
        (e = tmp.alloc(&ids, n)) != hipSuccess || (e = tmp.alloc(&left, n)) != hipSuccess ||
        (e = tmp.alloc(&right, n)) != hipSuccess || (e = tmp.alloc(&first, n)) != hipSuccess ||
        (e = tmp.alloc(&last, n)) != hipSuccess || (e = tmp.alloc(&parent_inner, n)) != hipSuccess ||
        (e = tmp.alloc(&parent_leaf, n)) != hipSuccess || (e = tmp.alloc(&nbox, 6 * (size_t)n)) != hipSuccess ||
        (e = tmp.alloc(&visits, n)) != hipSuccess || (e = tmp.alloc(&counters, 2)) != hipSuccess ||
        (e = tmp.alloc(&upper, n)) != hipSuccess)
        return e;
    prim_box_kernel<<<blocks, 256, 0, st>>>(d_verts, d_idx, n_tris, d_sph, n, pbox, cen);
    bounds_partial_kernel<<<rblocks, 256, 0, st>>>(cen, 3, n, partial);
    bounds_final_kernel<<<1, 64, 0, st>>>(partial, rblocks, cb);
    bounds_partial_kernel<<<rblocks, 256, 0, st>>>(pbox, 6, n, partial);
    bounds_final_kernel<<<1, 64, 0, st>>>(partial, rblocks, sb);
    morton_kernel<<<blocks, 256, 0, st>>>(cen, n, cb, keys_in, ids_in);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t sort_bytes = 0;
    if ((e = rocprim::radix_sort_pairs(nullptr, sort_bytes, keys_in, keys, ids_in, ids, n, 0, 63, st)) != hipSuccess)
        return e;
    void *sort_tmp = nullptr;
    if ((e = tmp.alloc((char **)&sort_tmp, sort_bytes)) != hipSuccess) return e;
    if ((e = rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys_in, keys, ids_in, ids, n, 0, 63, st)) != hipSuccess)
        return e;
    if ((e = hipMemsetAsync(parent_inner, 0xff, 4 * (size_t)n, st)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(visits, 0, 4 * (size_t)n, st)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(counters, 0, 8, st)) != hipSuccess) return e;
    hierarchy_kernel<<<blocks, 256, 0, st>>>(keys, (int)n, left, right, first, last, parent_inner, parent_leaf);
    refit_kernel<<<blocks, 256, 0, st>>>((int)n, left, right, parent_inner, parent_leaf, pbox, ids, nbox, visits);
    depth_kernel<<<blocks, 256, 0, st>>>((int)n, parent_inner, parent_leaf, counters);
    GpuNode *nodes = nullptr;
    PrimRecord *prims = nullptr;
    if ((e = hipMalloc(&nodes, (size_t)(n - 1) * sizeof(GpuNode))) != hipSuccess) return e;
    if ((e = hipMalloc(&prims, (size_t)n * sizeof(PrimRecord))) != hipSuccess) {
        hipFree(nodes);
        return e;
    }
    emit_kernel<<<blocks, 256, 0, st>>>((int)n, max_prims, left, right, first, last, pbox, ids, nbox, nodes,
                                        counters + 1);
    records_kernel<<<blocks, 256, 0, st>>>(d_verts, d_idx, n_tris, d_sph, n, ids, prims);
    upper_kernel<<<blocks, 256, 0, st>>>((int)n, keys, first, last, upper);
    uint32_t hc[2];
    res->upper.assign(n - 1, 0);
    if ((e = hipGetLastError()) == hipSuccess) e = hipMemcpyAsync(hc, counters, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(res->upper.data(), upper, n - 1, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(res->bounds, sb, 24, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        hipFree(nodes);
        hipFree(prims);
        return e;
    }
    res->nodes = nodes;
    res->prims = prims;
    res->n_nodes = n - 1;
    res->max_depth = hc[0];
    res->n_leaves = hc[1];
    return hipSuccess;
}

}  // namespace illume
