// ORACLE — test infrastructure only (see pbrt.h).
#include "bvh.h"
#include <stdexcept>

namespace orc {

Bounds3 PrimSet::world_bound(uint32_t p) const {
    uint32_t nt = n_tris();
    if (p < nt) {
        // triangle.rs:126-131
        const V3 &p0 = verts[idx[3 * p]], &p1 = verts[idx[3 * p + 1]], &p2 = verts[idx[3 * p + 2]];
        return bunion(Bounds3(p0, p1), p2);
    }
    return spheres[p - nt].world_bound();
}

bool PrimSet::intersect(uint32_t p, Ray &ray, SurfaceInteraction *si) const {
    uint32_t nt = n_tris();
    float t_hit;
    bool hit;
    if (p < nt)
        hit = triangle_intersect(verts[idx[3 * p]], verts[idx[3 * p + 1]], verts[idx[3 * p + 2]], reverse_orientation, ray,
                                 si, &t_hit);
    else
        hit = sphere_intersect(spheres[p - nt], ray, si, &t_hit);
    if (!hit) return false;
    ray.t_max = t_hit;
    si->prim = p;
    return true;
}

bool PrimSet::intersect_p(uint32_t p, const Ray &ray) const {
    uint32_t nt = n_tris();
    if (p < nt) return triangle_intersect_p(verts[idx[3 * p]], verts[idx[3 * p + 1]], verts[idx[3 * p + 2]], ray);
    return sphere_intersect_p(spheres[p - nt], ray);
}

namespace {
// bvh.rs:722-738
struct PrimInfo {
    uint32_t prim;
    Bounds3 bounds;
    V3 centroid;
};
struct BuildNode {
    Bounds3 bounds;
    BuildNode *children[2] = {nullptr, nullptr};
    int split_axis = 0, first_prim_offset = 0, n_primitives = 0;
};
struct Bucket {
    int count = 0;
    Bounds3 bounds;
};

struct Builder {
    int max_prims;
    std::vector<PrimInfo> &info;
    std::vector<uint32_t> &ordered;
    std::vector<BuildNode *> all;
    int total_nodes = 0;
    int max_depth = 0;
    Builder(int mp, std::vector<PrimInfo> &i, std::vector<uint32_t> &o) : max_prims(mp), info(i), ordered(o) {}
    ~Builder() {
        for (auto *n : all) delete n;
    }
    BuildNode *leaf(BuildNode *node, int start, int end, const Bounds3 &bounds) {
        node->first_prim_offset = (int)ordered.size();
        for (int i = start; i < end; ++i) ordered.push_back(info[i].prim);
        node->n_primitives = end - start;
        node->bounds = bounds;
        return node;
    }
    // bvh.rs:95-247, SAH branch, with pbrt-v3 semantics: centroid bounds from
    // centroids (divergence 2) and the bucket index computed as
    // int(n_buckets * offset) (divergence 3). std::partition and
    // std::nth_element stand in for SliceExtension::partition/partial_sort.
    BuildNode *build(int start, int end, int depth) {
        if (depth > max_depth) max_depth = depth;
        BuildNode *node = new BuildNode();
        all.push_back(node);
        ++total_nodes;
        Bounds3 bounds;
        for (int i = start; i < end; ++i) bounds = bunion(bounds, info[i].bounds);
        int n = end - start;
        if (n == 1) return leaf(node, start, end, bounds);
        Bounds3 cb;
        for (int i = start; i < end; ++i) cb = bunion(cb, info[i].centroid);
        int dim = cb.maximum_extent();
        int mid = (start + end) / 2;
        if (cb.pmax[dim] == cb.pmin[dim]) return leaf(node, start, end, bounds);
        if (n <= 2) {
            mid = (start + end) / 2;
            std::nth_element(&info[start], &info[mid], &info[end - 1] + 1,
                             [dim](const PrimInfo &a, const PrimInfo &b) { return a.centroid[dim] < b.centroid[dim]; });
        } else {
            const int nb = 12;
            Bucket buckets[nb];
            for (int i = start; i < end; ++i) {
                int b = (int)(nb * cb.offset(info[i].centroid)[dim]);
                if (b == nb) b = nb - 1;
                buckets[b].count++;
                buckets[b].bounds = bunion(buckets[b].bounds, info[i].bounds);
            }
            float cost[nb - 1];
            for (int i = 0; i < nb - 1; ++i) {
                Bounds3 b0, b1;
                int c0 = 0, c1 = 0;
                for (int j = 0; j <= i; ++j) {
                    b0 = bunion(b0, buckets[j].bounds);
                    c0 += buckets[j].count;
                }
                for (int j = i + 1; j < nb; ++j) {
                    b1 = bunion(b1, buckets[j].bounds);
                    c1 += buckets[j].count;
                }
                cost[i] = 1 + ((float)c0 * b0.surface_area() + (float)c1 * b1.surface_area()) / bounds.surface_area();
            }
            float min_cost = cost[0];
            int min_bucket = 0;
            for (int i = 1; i < nb - 1; ++i)
                if (cost[i] < min_cost) {
                    min_cost = cost[i];
                    min_bucket = i;
                }
            float leaf_cost = (float)n;
            if (n > max_prims || min_cost < leaf_cost) {
                PrimInfo *pmid = std::partition(&info[start], &info[end - 1] + 1, [=](const PrimInfo &pi) {
                    int b = (int)(nb * cb.offset(pi.centroid)[dim]);
                    if (b == nb) b = nb - 1;
                    return b <= min_bucket;
                });
                mid = (int)(pmid - &info[0]);
            } else {
                return leaf(node, start, end, bounds);
            }
        }
        node->children[0] = build(start, mid, depth + 1);
        node->children[1] = build(mid, end, depth + 1);
        node->bounds = bunion(node->children[0]->bounds, node->children[1]->bounds);
        node->split_axis = dim;
        node->n_primitives = 0;
        return node;
    }
};

// bvh.rs:504-537
int flatten(BuildNode *node, std::vector<LinearBVHNode> &nodes, int *offset) {
    LinearBVHNode &ln = nodes[*offset];
    int my = (*offset)++;
    ln.bounds = node->bounds;
    if (node->n_primitives > 0) {
        if (node->n_primitives >= 65536) throw std::runtime_error("leaf too large");
        nodes[my].offset = node->first_prim_offset;
        nodes[my].n_primitives = (uint16_t)node->n_primitives;
        nodes[my].axis = 0;
    } else {
        nodes[my].axis = (uint8_t)node->split_axis;
        nodes[my].n_primitives = 0;
        flatten(node->children[0], nodes, offset);
        nodes[my].offset = flatten(node->children[1], nodes, offset);
    }
    return my;
}
}  // namespace

void BVH::build(const PrimSet *p, int mp) {
    prims = p;
    max_prims_in_node = std::min(255, mp);
    nodes.clear();
    ordered.clear();
    uint32_t n = p->n_prims();
    if (n == 0) return;
    std::vector<PrimInfo> info(n);
    for (uint32_t i = 0; i < n; ++i) {
        info[i].prim = i;
        info[i].bounds = p->world_bound(i);
        info[i].centroid = info[i].bounds.centroid();
    }
    ordered.reserve(n);
    Builder b(max_prims_in_node, info, ordered);
    BuildNode *root = b.build(0, (int)n, 0);
    max_depth = b.max_depth;
    nodes.resize(b.total_nodes);
    int off = 0;
    flatten(root, nodes, &off);
}

// bvh.rs:563-612 with pbrt-v3's push in both branches (divergence 1).
uint32_t BVH::intersect(Ray &ray, SurfaceInteraction *si, TraversalStats *st) const {
    if (nodes.empty()) return 0xffffffffu;
    uint32_t hit = 0xffffffffu;
    V3 inv_dir(1 / ray.d.x, 1 / ray.d.y, 1 / ray.d.z);
    int dir_is_neg[3] = {inv_dir.x < 0, inv_dir.y < 0, inv_dir.z < 0};
    int to_visit = 0, cur = 0;
    int stack[64];
    while (true) {
        const LinearBVHNode &node = nodes[cur];
        if (st) st->node_tests++;
        if (node.bounds.intersect_p(ray, inv_dir, dir_is_neg)) {
            if (node.n_primitives > 0) {
                for (int i = 0; i < node.n_primitives; ++i) {
                    uint32_t p = ordered[node.offset + i];
                    if (st) st->prim_tests++;
                    if (prims->intersect(p, ray, si)) hit = p;
                }
                if (to_visit == 0) break;
                cur = stack[--to_visit];
            } else {
                if (dir_is_neg[node.axis]) {
                    stack[to_visit++] = cur + 1;
                    cur = node.offset;
                } else {
                    stack[to_visit++] = node.offset;
                    cur = cur + 1;
                }
            }
        } else {
            if (to_visit == 0) break;
            cur = stack[--to_visit];
        }
    }
    return hit;
}

// bvh.rs:614-662 (divergence 1 fixed as above)
bool BVH::intersect_p(const Ray &ray, TraversalStats *st) const {
    if (nodes.empty()) return false;
    V3 inv_dir(1 / ray.d.x, 1 / ray.d.y, 1 / ray.d.z);
    int dir_is_neg[3] = {inv_dir.x < 0, inv_dir.y < 0, inv_dir.z < 0};
    int to_visit = 0, cur = 0;
    int stack[64];
    while (true) {
        const LinearBVHNode &node = nodes[cur];
        if (st) st->node_tests++;
        if (node.bounds.intersect_p(ray, inv_dir, dir_is_neg)) {
            if (node.n_primitives > 0) {
                for (int i = 0; i < node.n_primitives; ++i) {
                    if (st) st->prim_tests++;
                    if (prims->intersect_p(ordered[node.offset + i], ray)) return true;
                }
                if (to_visit == 0) break;
                cur = stack[--to_visit];
            } else {
                if (dir_is_neg[node.axis]) {
                    stack[to_visit++] = cur + 1;
                    cur = node.offset;
                } else {
                    stack[to_visit++] = node.offset;
                    cur = cur + 1;
                }
            }
        } else {
            if (to_visit == 0) break;
            cur = stack[--to_visit];
        }
    }
    return false;
}

uint32_t brute_intersect(const PrimSet &ps, Ray &ray, SurfaceInteraction *si) {
    uint32_t hit = 0xffffffffu;
    for (uint32_t p = 0; p < ps.n_prims(); ++p)
        if (ps.intersect(p, ray, si)) hit = p;
    return hit;
}

bool brute_intersect_p(const PrimSet &ps, const Ray &ray) {
    for (uint32_t p = 0; p < ps.n_prims(); ++p)
        if (ps.intersect_p(p, ray)) return true;
    return false;
}

}  // namespace orc
