// ORACLE — test infrastructure only (see pbrt.h). pbrt-v3 BVHAccel restated
// from accelerators/bvh.rs: SAH build (12 buckets), depth-first flattening
// into 32-byte nodes, stack traversal for closest hit and any hit.
#pragma once
#include <vector>
#include "shapes.h"

namespace orc {

// A scene's primitives: triangles (world-space vertex array + index triples)
// first, then spheres. Primitive id p < n_tris is triangle p, otherwise sphere
// p - n_tris. This is the caller's original order (bvh.rs:58 reorders
// internally; ids returned to callers are the original ones).
struct PrimSet {
    std::vector<V3> verts;
    std::vector<uint32_t> idx;  // 3 per triangle
    std::vector<SphereShape> spheres;
    bool reverse_orientation = false;
    uint32_t n_tris() const { return (uint32_t)(idx.size() / 3); }
    uint32_t n_prims() const { return n_tris() + (uint32_t)spheres.size(); }
    Bounds3 world_bound(uint32_t p) const;
    // GeometricPrimitive::intersect (geometric_primitive.rs:42-65): on a hit
    // the ray's t_max becomes t_hit.
    bool intersect(uint32_t p, Ray &ray, SurfaceInteraction *si) const;
    bool intersect_p(uint32_t p, const Ray &ray) const;  // geometric_primitive.rs:67-69
};

// bvh.rs:665-673 with pbrt-v3's 32-byte layout (divergence 5: the reference
// aligns to 64 bytes).
struct LinearBVHNode {
    Bounds3 bounds;
    int32_t offset;  // primitives offset (leaf) or second child offset (interior)
    uint16_t n_primitives;
    uint8_t axis;
    uint8_t pad;
};

struct TraversalStats {
    uint64_t node_tests = 0;  // Bounds3::intersect_p calls
    uint64_t prim_tests = 0;  // primitive intersect / intersect_p calls
};

struct BVH {
    const PrimSet *prims = nullptr;
    int max_prims_in_node = 4;
    std::vector<uint32_t> ordered;  // leaf-order -> original primitive id
    std::vector<LinearBVHNode> nodes;
    int max_depth = 0;
    // bvh.rs:32-93
    void build(const PrimSet *p, int max_prims);
    // bvh.rs:563-612; returns the primitive id or UINT32_MAX on a miss.
    uint32_t intersect(Ray &ray, SurfaceInteraction *si, TraversalStats *st = nullptr) const;
    // bvh.rs:614-662
    bool intersect_p(const Ray &ray, TraversalStats *st = nullptr) const;
};

// O(N) references over all primitives (test-only cross-checks).
uint32_t brute_intersect(const PrimSet &ps, Ray &ray, SurfaceInteraction *si);
bool brute_intersect_p(const PrimSet &ps, const Ray &ray);

}  // namespace orc
