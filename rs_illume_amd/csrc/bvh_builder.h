// Host-side BVH builder: binned SAH over triangles and spheres, emitted in the
// 64-byte two-child node layout of layout.h.
#pragma once
#include <string>
#include <vector>
#include "layout.h"

namespace illume {

// Split rule of the binned SAH build, shared by the host builder and the
// device builder (sah_gpu.hip), which make the same tree: kSahBins bins on
// each axis over the node's centroid bounds, split cost kSahTraversalCost +
// (n0 A0 + n1 A1) / A against leaf cost n, object medians from depth
// kSahMedianDepth on.
constexpr int kSahBins = 16;
constexpr int kSahMedianDepth = 40;
constexpr float kSahTraversalCost = 0.25f;  // a node step against one primitive test

// Surface area of a box of extents dx, dy, dz (0 for an empty box), in the
// operation order both builders use.
__host__ __device__ inline float sah_area(float dx, float dy, float dz) {
    if (!(dx >= 0 && dy >= 0 && dz >= 0)) return 0.f;
    return 2.f * (dx * dy + dx * dz + dy * dz);
}

struct BuildInput {
    const float *verts = nullptr;  // 3 floats per vertex
    uint32_t n_verts = 0;
    const uint32_t *idx = nullptr;  // 3 per triangle
    uint32_t n_tris = 0;
    const float *spheres = nullptr;  // cx, cy, cz, r
    uint32_t n_spheres = 0;
    int max_prims_in_node = 4;
    int width = 2;  // 2: GpuNode, 4: GpuNode4 (binary tree collapsed)
};

struct BuildOutput {
    std::vector<GpuNode> nodes;
    std::vector<GpuNode4> nodes4;
    std::vector<PrimRecord> prims;  // leaf order
    uint32_t n_leaves = 0;
    uint32_t max_depth = 0;
    float bounds[6] = {0, 0, 0, 0, 0, 0};
};

// Returns an empty string on success, else a description of the invalid input.
std::string build_bvh(const BuildInput &in, BuildOutput *out);

}  // namespace illume
