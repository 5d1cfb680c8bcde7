// Device BVH builder (lbvh.hip) and the host collapse of a two-child BVH into
// the four-child layout.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "layout.h"

namespace illume {

struct LbvhResult {
    GpuNode *nodes = nullptr;      // device, n_nodes two-child nodes, root at 0
    PrimRecord *prims = nullptr;   // device, leaf (Morton) order
    uint32_t n_nodes = 0, n_leaves = 0, max_depth = 0;
    float bounds[6] = {0, 0, 0, 0, 0, 0};
    // host: per node, 1 if its primitives span more than one treelet (their
    // Morton codes differ within the top kTreeletBits bits)
    std::vector<uint8_t> upper;
};

// Treelets of the SAH top: primitives whose Morton codes share the top 12 bits
// (4 per axis), as the reference's HLBVH clusters them (bvh.rs:276-300).
constexpr int kTreeletBits = 12;

// The SAH top (the reference's build_upper_sah, bvh.rs:406-502, over pbrt's
// 12 buckets): the nodes above the treelets (upper[i] != 0, reachable from the
// root through upper nodes only) are rebuilt by a binned SAH over the treelet
// roots' boxes, in the same node slots (root at 0; T treelets need T - 1
// nodes either way). Returns the binary tree's depth after the rebuild.
uint32_t sah_top(std::vector<GpuNode> &n2, const std::vector<uint8_t> &upper);
// Inner-node depth of a two-child BVH (root at 0).
uint32_t binary_depth(const std::vector<GpuNode> &n2);

// Host input arrays (as for build_bvh), already validated; n_tris + n_spheres
// >= 2. The caller owns res->nodes and res->prims (hipFree).
hipError_t lbvh_build(const float *verts, uint32_t n_verts, const uint32_t *idx, uint32_t n_tris,
                      const float *spheres, uint32_t n_spheres, int max_prims, hipStream_t st, LbvhResult *res);

// Binned SAH build on the device (sah_gpu.hip), the host builder's split rule.
struct SahGpuResult {
    GpuNode *nodes = nullptr;      // width 2: device, n_nodes two-child nodes, root at 0
    GpuNode4Q *nodes4 = nullptr;   // width 4: device, n_nodes quantized four-child nodes, root at 0
    PrimRecord *prims = nullptr;   // device, leaf order
    uint32_t n_nodes = 0, n_leaves = 0;
    uint32_t depth = 0;            // levels of the emitted tree (two- or four-child)
    uint32_t binary_depth = 0;     // inner nodes on the longest root-to-leaf path of the binary tree
    uint32_t top_levels = 0;       // level passes of the top phase
    float bounds[6] = {0, 0, 0, 0, 0, 0};
};

// Host input arrays, already validated; n_tris + n_spheres >= 2. The caller
// owns res->nodes / res->nodes4 and res->prims (hipFree). Complete on return.
hipError_t sah_gpu_build(const float *verts, uint32_t n_verts, const uint32_t *idx, uint32_t n_tris,
                         const float *spheres, uint32_t n_spheres, int max_prims, int width, hipStream_t st,
                         SahGpuResult *res);

// Four-child nodes from a two-child BVH (any node order, root at 0): each
// four-child node takes its binary node's two children and opens the inner
// child of largest surface area until it has four. Depth-first order; leaf
// references are kept. Returns the four-child depth.
uint32_t collapse_to_width4(const std::vector<GpuNode> &n2, std::vector<GpuNode4> *n4);

// The 64-byte quantized form of four-child nodes (same topology and child
// references). False if a box is not finite or cannot be covered.
bool quantize_width4(const std::vector<GpuNode4> &n4, std::vector<GpuNode4Q> *out);

// Validation shared by both builders: "" or a description of the problem.
std::string validate_prims(const float *verts, uint32_t n_verts, const uint32_t *idx, uint32_t n_tris,
                           const float *spheres, uint32_t n_spheres);

}  // namespace illume
