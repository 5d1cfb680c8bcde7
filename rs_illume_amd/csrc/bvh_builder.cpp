// Binned SAH builder (replaces BVHAccel::new / recursive_build,
// accelerators/bvh.rs:32-247, for SplitMethod::SAH).
//
// Differences from the reference's build that do not change hit results:
//  * 16 bins on each of the 3 axes (the reference: 12 buckets on the longest
//    centroid axis); split cost kTraversalCost + (n0*SA0 + n1*SA1)/SA against
//    leaf cost n, where the reference prices the traversal step at 1
//    (bvh.rs:199-200). On the GPU traversal a node step costs about as much as
//    a primitive test and steps of the two kinds run side by side, so smaller
//    leaves pay: kTraversalCost 0.25 gives closest hit +3-4% on the
//    2^20-triangle soup and +4% on a terrain view against 1 (0 to 0.5 within
//    2%; 1.5 and 2 lose 11% and 21%).
//  * A node whose primitives all share one centroid is split by count rather
//    than made into an oversized leaf (bvh.rs:134-142), so every leaf holds at
//    most max_prims_in_node primitives.
//  * Below depth kMedianDepth splits are object medians, which bounds the
//    tree depth and so the traversal stack.
//  * Sphere boxes are widened by one ulp outwards so box culling stays
//    conservative for hits computed with the sphere's interval arithmetic.
//  * Partitions are stable (each side keeps its order) and object medians
//    take the n/2 smallest (centroid, position) pairs, so the device build
//    (sah_gpu.hip) reproduces the leaves with their primitives in the same
//    order: the sphere test's interval check makes the reported hit depend on
//    the order in which a leaf's primitives are tested when a sphere and
//    another primitive are hit at nearly the same distance.
// Closest-hit ids do not depend on the tree shape up to ties (two primitives
// hit within an ulp or two: which one a traversal reports depends on the
// order it reaches them), which is why the tree may differ from the
// reference's.
#include "bvh_builder.h"
#include "lbvh.h"
#include "quantize.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <tuple>

namespace illume {
namespace {

constexpr int kBins = kSahBins;
constexpr int kMedianDepth = kSahMedianDepth;
constexpr float kTraversalCost = kSahTraversalCost;

struct Box {
    float lo[3], hi[3];
    Box() {
        for (int i = 0; i < 3; ++i) {
            lo[i] = std::numeric_limits<float>::infinity();
            hi[i] = -std::numeric_limits<float>::infinity();
        }
    }
    void grow(const Box &b) {
        for (int i = 0; i < 3; ++i) {
            lo[i] = std::min(lo[i], b.lo[i]);
            hi[i] = std::max(hi[i], b.hi[i]);
        }
    }
    void grow(const float *p) {
        for (int i = 0; i < 3; ++i) {
            lo[i] = std::min(lo[i], p[i]);
            hi[i] = std::max(hi[i], p[i]);
        }
    }
    float area() const {
        return sah_area(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
    }
};

struct TmpNode {
    Box box;
    int32_t left = -1, right = -1;  // children (TmpNode indices) or -1 for a leaf
    uint32_t start = 0, count = 0;  // leaf range in `ids`
};

struct Builder {
    const BuildInput &in;
    std::vector<Box> pbox;
    std::vector<float> cen;  // 3 per prim
    std::vector<uint32_t> ids;
    std::vector<uint32_t> scratch;
    std::vector<TmpNode> tmp;
    uint32_t max_depth = 0;
    int max_prims;

    explicit Builder(const BuildInput &i) : in(i), max_prims(i.max_prims_in_node) {}

    int leaf(const Box &b, uint32_t s, uint32_t e) {
        TmpNode n;
        n.box = b;
        n.start = s;
        n.count = e - s;
        tmp.push_back(n);
        return (int)tmp.size() - 1;
    }

    // Reorders ids[s, e): the entries with left[i] first, then the others,
    // each side in its current order (a stable partition); returns the split.
    uint32_t stable_split(uint32_t s, uint32_t e, const std::vector<uint8_t> &left) {
        std::vector<uint32_t> &t = scratch;
        t.clear();
        for (uint32_t i = s; i < e; ++i)
            if (left[i - s]) t.push_back(ids[i]);
        const uint32_t mid = s + (uint32_t)t.size();
        for (uint32_t i = s; i < e; ++i)
            if (!left[i - s]) t.push_back(ids[i]);
        std::copy(t.begin(), t.end(), ids.begin() + s);
        return mid;
    }

    // Object median along the longest centroid axis: the n/2 smallest
    // (centroid, position) pairs go left.
    uint32_t median_split(uint32_t s, uint32_t e, const Box &cb) {
        int axis = 0;
        float ext = -1;
        for (int a = 0; a < 3; ++a) {
            float x = cb.hi[a] - cb.lo[a];
            if (x > ext) {
                ext = x;
                axis = a;
            }
        }
        const uint32_t n = e - s, half = n / 2;
        std::vector<std::pair<float, uint32_t>> key(n);
        for (uint32_t i = 0; i < n; ++i) key[i] = {cen[3 * ids[s + i] + axis], i};
        std::vector<std::pair<float, uint32_t>> sel(key);
        std::nth_element(sel.begin(), sel.begin() + half, sel.end());
        const std::pair<float, uint32_t> pivot = sel[half];
        std::vector<uint8_t> left(n);
        for (uint32_t i = 0; i < n; ++i) left[i] = key[i] < pivot ? 1 : 0;
        return stable_split(s, e, left);
    }

    int build(uint32_t s, uint32_t e, uint32_t depth) {
        max_depth = std::max(max_depth, depth);
        Box b, cb;
        for (uint32_t i = s; i < e; ++i) {
            b.grow(pbox[ids[i]]);
            cb.grow(&cen[3 * ids[i]]);
        }
        uint32_t n = e - s;
        if (n == 1) return leaf(b, s, e);
        uint32_t mid = 0;
        bool split = false;
        if (depth >= (uint32_t)kMedianDepth) {
            if (n <= (uint32_t)max_prims) return leaf(b, s, e);
            mid = median_split(s, e, cb);
            split = true;
        } else {
            float best_cost = std::numeric_limits<float>::infinity();
            int best_axis = -1, best_bin = -1;
            float parent_area = b.area();
            for (int a = 0; a < 3; ++a) {
                float ext = cb.hi[a] - cb.lo[a];
                if (!(ext > 0)) continue;
                float k = kBins / ext;
                Box bb[kBins];
                uint32_t bc[kBins] = {0};
                for (uint32_t i = s; i < e; ++i) {
                    uint32_t p = ids[i];
                    int bi = (int)((cen[3 * p + a] - cb.lo[a]) * k);
                    bi = std::min(std::max(bi, 0), kBins - 1);
                    bc[bi]++;
                    bb[bi].grow(pbox[p]);
                }
                float right_area[kBins];
                uint32_t right_cnt[kBins];
                Box acc;
                uint32_t cnt = 0;
                for (int i = kBins - 1; i > 0; --i) {
                    acc.grow(bb[i]);
                    cnt += bc[i];
                    right_area[i] = acc.area();
                    right_cnt[i] = cnt;
                }
                Box lacc;
                uint32_t lcnt = 0;
                for (int i = 0; i < kBins - 1; ++i) {
                    lacc.grow(bb[i]);
                    lcnt += bc[i];
                    if (lcnt == 0 || right_cnt[i + 1] == 0) continue;
                    float cost = kTraversalCost + ((float)lcnt * lacc.area() + (float)right_cnt[i + 1] * right_area[i + 1]) /
                                           std::max(parent_area, 1e-30f);
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = a;
                        best_bin = i;
                    }
                }
            }
            if (best_axis < 0) {
                // all centroids coincide
                if (n <= (uint32_t)max_prims) return leaf(b, s, e);
                mid = s + n / 2;
                split = true;
            } else {
                if (n <= (uint32_t)max_prims && best_cost >= (float)n) return leaf(b, s, e);
                int a = best_axis;
                float k = kBins / (cb.hi[a] - cb.lo[a]);
                float lo = cb.lo[a];
                std::vector<uint8_t> left(n);
                for (uint32_t i = s; i < e; ++i) {
                    int bi = (int)((cen[3 * ids[i] + a] - lo) * k);
                    bi = std::min(std::max(bi, 0), kBins - 1);
                    left[i - s] = bi <= best_bin ? 1 : 0;
                }
                mid = stable_split(s, e, left);
                if (mid == s || mid == e) mid = median_split(s, e, cb);
                split = true;
            }
        }
        (void)split;
        int self = (int)tmp.size();
        tmp.push_back(TmpNode());
        tmp[self].box = b;
        int l = build(s, mid, depth + 1);
        int r = build(mid, e, depth + 1);
        tmp[self].left = l;
        tmp[self].right = r;
        return self;
    }
};

inline bool finite3(const float *p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

}  // namespace

std::string build_bvh(const BuildInput &in, BuildOutput *out) {
    out->nodes.clear();
    out->prims.clear();
    out->n_leaves = 0;
    out->max_depth = 0;
    if (in.max_prims_in_node < 1 || in.max_prims_in_node > kMaxLeafPrims)
        return "max_prims_in_node must be in 1..15";
    uint64_t n_prims = (uint64_t)in.n_tris + in.n_spheres;
    if (n_prims >= kMaxPrimSlots) return "too many primitives (limit 2^27)";
    if (in.n_tris && (!in.verts || !in.idx)) return "null vertex or index array";
    if (in.n_spheres && !in.spheres) return "null sphere array";

    Builder B(in);
    B.pbox.resize(n_prims);
    B.cen.resize(3 * n_prims);
    B.ids.resize(n_prims);
    for (uint32_t v = 0; v < in.n_verts && in.n_tris; ++v)
        if (!finite3(in.verts + 3 * (size_t)v)) return "non-finite vertex " + std::to_string(v);
    for (uint32_t t = 0; t < in.n_tris; ++t) {
        Box b;
        for (int k = 0; k < 3; ++k) {
            uint32_t vi = in.idx[3 * (size_t)t + k];
            if (vi >= in.n_verts) return "triangle " + std::to_string(t) + " indexes vertex out of range";
            b.grow(in.verts + 3 * (size_t)vi);
        }
        B.pbox[t] = b;
    }
    for (uint32_t i = 0; i < in.n_spheres; ++i) {
        const float *sp = in.spheres + 4 * (size_t)i;
        if (!finite3(sp) || !std::isfinite(sp[3]) || !(sp[3] > 0)) return "invalid sphere " + std::to_string(i);
        Box b;
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = std::nextafter(sp[a] + (-sp[3]), -std::numeric_limits<float>::infinity());
            b.hi[a] = std::nextafter(sp[a] + sp[3], std::numeric_limits<float>::infinity());
        }
        B.pbox[in.n_tris + i] = b;
    }
    for (uint64_t p = 0; p < n_prims; ++p) {
        for (int a = 0; a < 3; ++a) B.cen[3 * p + a] = B.pbox[p].lo[a] * 0.5f + B.pbox[p].hi[a] * 0.5f;
        B.ids[p] = (uint32_t)p;
    }
    if (n_prims == 0) return "";
    B.tmp.reserve(2 * n_prims);
    int root = B.build(0, (uint32_t)n_prims, 0);
    out->max_depth = B.max_depth;
    const Box &rb = B.tmp[root].box;
    for (int a = 0; a < 3; ++a) {
        out->bounds[a] = rb.lo[a];
        out->bounds[3 + a] = rb.hi[a];
    }

    // Emit: depth-first over inner nodes; leaves append their primitive
    // records in visit order.
    out->prims.reserve(n_prims);
    auto emit_prims = [&](const TmpNode &t) -> int32_t {
        uint32_t first = (uint32_t)out->prims.size();
        for (uint32_t i = 0; i < t.count; ++i) {
            uint32_t p = B.ids[t.start + i];
            PrimRecord r;
            std::memset(&r, 0, sizeof(r));
            if (p < in.n_tris) {
                for (int k = 0; k < 3; ++k) {
                    const float *v = in.verts + 3 * (size_t)in.idx[3 * (size_t)p + k];
                    float *dst = k == 0 ? r.a : (k == 1 ? r.b : r.c);
                    dst[0] = v[0];
                    dst[1] = v[1];
                    dst[2] = v[2];
                }
                std::memcpy(&r.a[3], &p, 4);
            } else {
                const float *sp = in.spheres + 4 * (size_t)(p - in.n_tris);
                r.a[0] = sp[0];
                r.a[1] = sp[1];
                r.a[2] = sp[2];
                uint32_t tag = p | kSphereBit;
                std::memcpy(&r.a[3], &tag, 4);
                r.b[0] = sp[3];
            }
            out->prims.push_back(r);
        }
        out->n_leaves++;
        return make_leaf_ref(first, t.count);
    };
    auto set_box = [](GpuNode &g, int which, const Box &b) {
        if (which == 0) {
            g.c0_lo_x = b.lo[0]; g.c0_hi_x = b.hi[0]; g.c0_lo_y = b.lo[1]; g.c0_hi_y = b.hi[1];
            g.c0_lo_z = b.lo[2]; g.c0_hi_z = b.hi[2];
        } else {
            g.c1_lo_x = b.lo[0]; g.c1_hi_x = b.hi[0]; g.c1_lo_y = b.lo[1]; g.c1_hi_y = b.hi[1];
            g.c1_lo_z = b.lo[2]; g.c1_hi_z = b.hi[2];
        }
    };
    const float inf = std::numeric_limits<float>::infinity();
    Box never;  // a box no ray enters: all planes at +inf
    for (int a = 0; a < 3; ++a) never.lo[a] = never.hi[a] = inf;

    if (in.width == 4) {
        // Collapse: each 4-wide node takes its binary node's two children and
        // repeatedly opens the inner child of largest area until it has four.
        auto set4 = [](GpuNode4 &g, int c, const Box &b) {
            g.lo_x[c] = b.lo[0]; g.hi_x[c] = b.hi[0];
            g.lo_y[c] = b.lo[1]; g.hi_y[c] = b.hi[1];
            g.lo_z[c] = b.lo[2]; g.hi_z[c] = b.hi[2];
        };
        out->nodes4.reserve(n_prims / 2 + 1);
        out->nodes4.push_back(GpuNode4());
        std::vector<std::tuple<int, int, uint32_t>> st4;  // (tmp index, node slot, depth)
        st4.push_back({root, 0, 1});
        out->max_depth = 0;
        while (!st4.empty()) {
            auto [ti, gi, dep] = st4.back();
            st4.pop_back();
            out->max_depth = std::max(out->max_depth, dep);
            int kids[4], nk = 0;
            if (B.tmp[ti].left < 0) {
                kids[nk++] = ti;  // the root is a leaf
            } else {
                kids[nk++] = B.tmp[ti].left;
                kids[nk++] = B.tmp[ti].right;
                while (nk < 4) {
                    int best = -1;
                    float best_area = -1.f;
                    for (int c = 0; c < nk; ++c) {
                        const TmpNode &k = B.tmp[kids[c]];
                        if (k.left >= 0 && k.box.area() > best_area) {
                            best_area = k.box.area();
                            best = c;
                        }
                    }
                    if (best < 0) break;
                    const TmpNode &k = B.tmp[kids[best]];
                    kids[best] = k.left;
                    kids[nk++] = k.right;
                }
            }
            GpuNode4 g;
            std::memset(&g, 0, sizeof(g));
            for (int c = 0; c < 4; ++c) {
                if (c >= nk) {
                    set4(g, c, never);
                    g.child[c] = make_leaf_ref(0, 0);
                    continue;
                }
                const TmpNode &k = B.tmp[kids[c]];
                set4(g, c, k.box);
                if (k.left < 0) {
                    g.child[c] = emit_prims(k);
                } else {
                    int slot = (int)out->nodes4.size();
                    out->nodes4.push_back(GpuNode4());
                    g.child[c] = slot;
                    st4.push_back({kids[c], slot, dep + 1});
                }
            }
            out->nodes4[gi] = g;
        }
        return "";
    }

    if (B.tmp[root].left < 0) {
        GpuNode g;
        std::memset(&g, 0, sizeof(g));
        set_box(g, 0, B.tmp[root].box);
        set_box(g, 1, never);
        g.child0 = emit_prims(B.tmp[root]);
        g.child1 = make_leaf_ref(0, 0);  // unreachable: its box is never entered
        out->nodes.push_back(g);
        return "";
    }
    // iterative DFS: (tmp index, gpu slot)
    std::vector<std::pair<int, int>> stack;
    out->nodes.reserve(n_prims);
    out->nodes.push_back(GpuNode());
    stack.push_back({root, 0});
    while (!stack.empty()) {
        auto [ti, gi] = stack.back();
        stack.pop_back();
        const TmpNode t = B.tmp[ti];
        GpuNode g;
        std::memset(&g, 0, sizeof(g));
        int kids[2] = {t.left, t.right};
        int32_t refs[2];
        for (int c = 0; c < 2; ++c) {
            const TmpNode &k = B.tmp[kids[c]];
            set_box(g, c, k.box);
            if (k.left < 0) {
                refs[c] = emit_prims(k);
            } else {
                int slot = (int)out->nodes.size();
                out->nodes.push_back(GpuNode());
                refs[c] = slot;
            }
        }
        g.child0 = refs[0];
        g.child1 = refs[1];
        out->nodes[gi] = g;
        // push right first so the left subtree is laid out next
        if (B.tmp[t.right].left >= 0) stack.push_back({t.right, refs[1]});
        if (B.tmp[t.left].left >= 0) stack.push_back({t.left, refs[0]});
    }
    return "";
}

}  // namespace illume

namespace illume {

std::string validate_prims(const float *verts, uint32_t n_verts, const uint32_t *idx, uint32_t n_tris,
                           const float *spheres, uint32_t n_spheres) {
    if ((uint64_t)n_tris + n_spheres >= kMaxPrimSlots) return "too many primitives (limit 2^27)";
    if (n_tris && (!verts || !idx)) return "null vertex or index array";
    if (n_spheres && !spheres) return "null sphere array";
    for (uint32_t v = 0; v < n_verts && n_tris; ++v)
        if (!finite3(verts + 3 * (size_t)v)) return "non-finite vertex " + std::to_string(v);
    for (uint64_t t = 0; t < 3 * (uint64_t)n_tris; ++t)
        if (idx[t] >= n_verts) return "triangle " + std::to_string(t / 3) + " indexes vertex out of range";
    for (uint32_t i = 0; i < n_spheres; ++i) {
        const float *sp = spheres + 4 * (size_t)i;
        if (!finite3(sp) || !std::isfinite(sp[3]) || !(sp[3] > 0)) return "invalid sphere " + std::to_string(i);
    }
    return "";
}

namespace {
struct TopItem {
    int32_t ref;  // treelet root: node index or leaf reference
    float b[6];   // its box (lo xyz, hi xyz)
    float c[3];   // box centre
};
void box_union(float *acc, const float *b) {
    for (int a = 0; a < 3; ++a) {
        acc[a] = std::min(acc[a], b[a]);
        acc[3 + a] = std::max(acc[3 + a], b[3 + a]);
    }
}
float box_area(const float *b) {
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return 2.f * (dx * dy + dy * dz + dz * dx);
}
void child_of(const GpuNode &g, int c, int32_t *ref, float *b) {
    if (c == 0) {
        *ref = g.child0;
        b[0] = g.c0_lo_x; b[1] = g.c0_lo_y; b[2] = g.c0_lo_z; b[3] = g.c0_hi_x; b[4] = g.c0_hi_y; b[5] = g.c0_hi_z;
    } else {
        *ref = g.child1;
        b[0] = g.c1_lo_x; b[1] = g.c1_lo_y; b[2] = g.c1_lo_z; b[3] = g.c1_hi_x; b[4] = g.c1_hi_y; b[5] = g.c1_hi_z;
    }
}
void set_child(GpuNode &g, int c, int32_t ref, const float *b) {
    if (c == 0) {
        g.child0 = ref;
        g.c0_lo_x = b[0]; g.c0_lo_y = b[1]; g.c0_lo_z = b[2]; g.c0_hi_x = b[3]; g.c0_hi_y = b[4]; g.c0_hi_z = b[5];
    } else {
        g.child1 = ref;
        g.c1_lo_x = b[0]; g.c1_lo_y = b[1]; g.c1_lo_z = b[2]; g.c1_hi_x = b[3]; g.c1_hi_y = b[4]; g.c1_hi_z = b[5];
    }
}

// Binary SAH tree over items[begin, end) written into the next free slots;
// returns the subtree's reference and box. One item is the item itself.
struct TopBuilder {
    std::vector<TopItem> &items;
    std::vector<GpuNode> &nodes;
    const std::vector<int32_t> &slots;
    size_t next = 0;
    int32_t build(size_t begin, size_t end, float *box) {
        if (end - begin == 1) {
            std::memcpy(box, items[begin].b, sizeof(items[begin].b));
            return items[begin].ref;
        }
        const int32_t slot = slots[next++];
        float cb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (size_t i = begin; i < end; ++i) {
            const float c6[6] = {items[i].c[0], items[i].c[1], items[i].c[2], items[i].c[0], items[i].c[1],
                                 items[i].c[2]};
            box_union(cb, c6);
        }
        int dim = 0;
        for (int a = 1; a < 3; ++a)
            if (cb[3 + a] - cb[a] > cb[3 + dim] - cb[dim]) dim = a;
        size_t mid = begin + (end - begin) / 2;
        const float ext = cb[3 + dim] - cb[dim];
        if (ext > 0.f) {
            // 12 buckets along dim; the split with the least count x area on both sides
            constexpr int kB = 12;
            int cnt[kB] = {0};
            float bb[kB][6];
            for (int k = 0; k < kB; ++k)
                for (int a = 0; a < 3; ++a) {
                    bb[k][a] = INFINITY;
                    bb[k][3 + a] = -INFINITY;
                }
            auto bucket = [&](const TopItem &it) {
                return std::min(kB - 1, (int)((float)kB * ((it.c[dim] - cb[dim]) / ext)));
            };
            for (size_t i = begin; i < end; ++i) {
                const int k = bucket(items[i]);
                ++cnt[k];
                box_union(bb[k], items[i].b);
            }
            float best = INFINITY;
            int best_k = -1;
            for (int k = 0; k < kB - 1; ++k) {
                float l[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY}, r[6];
                std::memcpy(r, l, sizeof(l));
                int nl = 0, nr = 0;
                for (int j = 0; j <= k; ++j) {
                    nl += cnt[j];
                    if (cnt[j]) box_union(l, bb[j]);
                }
                for (int j = k + 1; j < kB; ++j) {
                    nr += cnt[j];
                    if (cnt[j]) box_union(r, bb[j]);
                }
                if (nl == 0 || nr == 0) continue;
                const float cost = (float)nl * box_area(l) + (float)nr * box_area(r);
                if (cost < best) {
                    best = cost;
                    best_k = k;
                }
            }
            if (best_k >= 0) {
                auto it = std::partition(items.begin() + begin, items.begin() + end,
                                         [&](const TopItem &x) { return bucket(x) <= best_k; });
                mid = (size_t)(it - items.begin());
            }
        }
        if (mid == begin || mid == end) mid = begin + (end - begin) / 2;
        float bl[6], br[6];
        const int32_t rl = build(begin, mid, bl);
        const int32_t rr = build(mid, end, br);
        GpuNode g;
        std::memset(&g, 0, sizeof(g));
        set_child(g, 0, rl, bl);
        set_child(g, 1, rr, br);
        nodes[slot] = g;
        std::memcpy(box, bl, sizeof(bl));
        box_union(box, br);
        return slot;
    }
};
}  // namespace

uint32_t binary_depth(const std::vector<GpuNode> &n2) {
    if (n2.empty()) return 0;
    uint32_t best = 0;
    std::vector<std::pair<int32_t, uint32_t>> st{{0, 1}};
    while (!st.empty()) {
        auto [i, d] = st.back();
        st.pop_back();
        best = std::max(best, d);
        if (n2[i].child0 >= 0) st.push_back({n2[i].child0, d + 1});
        if (n2[i].child1 >= 0) st.push_back({n2[i].child1, d + 1});
    }
    return best;
}

uint32_t sah_top(std::vector<GpuNode> &n2, const std::vector<uint8_t> &upper) {
    if (n2.empty() || upper.size() < n2.size() || !upper[0]) return binary_depth(n2);
    // the upper nodes (their slots, root first) and the treelet roots below them
    std::vector<int32_t> slots;
    std::vector<TopItem> items;
    std::vector<int32_t> st{0};
    while (!st.empty()) {
        const int32_t i = st.back();
        st.pop_back();
        slots.push_back(i);
        for (int c = 0; c < 2; ++c) {
            TopItem it;
            child_of(n2[i], c, &it.ref, it.b);
            if (it.ref >= 0 && upper[it.ref]) {
                st.push_back(it.ref);
            } else {
                for (int a = 0; a < 3; ++a) it.c[a] = 0.5f * (it.b[a] + it.b[3 + a]);
                items.push_back(it);
            }
        }
    }
    // a binary tree over T treelet roots has T - 1 inner nodes: the same count
    if (items.size() != slots.size() + 1) return binary_depth(n2);
    TopBuilder tb{items, n2, slots};
    float box[6];
    tb.build(0, items.size(), box);
    return binary_depth(n2);
}

uint32_t collapse_to_width4(const std::vector<GpuNode> &n2, std::vector<GpuNode4> *n4) {
    struct Kid {
        float b[6];
        int32_t ref;
    };
    auto kid = [&](const GpuNode &g, int c) {
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
    };
    auto area = [](const Kid &k) { return collapse_area(k.b); };
    const float inf = std::numeric_limits<float>::infinity();
    n4->clear();
    n4->reserve(n2.size() / 2 + 1);
    n4->push_back(GpuNode4());
    std::vector<std::tuple<int32_t, int, uint32_t>> st;  // (binary node, slot, depth)
    st.push_back({0, 0, 1});
    uint32_t max_depth = 0;
    while (!st.empty()) {
        auto [bi, slot, dep] = st.back();
        st.pop_back();
        max_depth = std::max(max_depth, dep);
        Kid kids[4];
        int nk = 0;
        kids[nk++] = kid(n2[bi], 0);
        kids[nk++] = kid(n2[bi], 1);
        while (nk < 4) {
            int best = -1;
            float best_area = -1.f;
            for (int c = 0; c < nk; ++c)
                if (kids[c].ref >= 0 && area(kids[c]) > best_area) {
                    best_area = area(kids[c]);
                    best = c;
                }
            if (best < 0) break;
            const GpuNode &g = n2[kids[best].ref];
            kids[best] = kid(g, 0);
            kids[nk++] = kid(g, 1);
        }
        GpuNode4 g4;
        std::memset(&g4, 0, sizeof(g4));
        for (int c = 0; c < 4; ++c) {
            if (c >= nk) {
                g4.lo_x[c] = g4.hi_x[c] = g4.lo_y[c] = g4.hi_y[c] = g4.lo_z[c] = g4.hi_z[c] = inf;
                g4.child[c] = make_leaf_ref(0, 0);
                continue;
            }
            const Kid &k = kids[c];
            g4.lo_x[c] = k.b[0]; g4.lo_y[c] = k.b[1]; g4.lo_z[c] = k.b[2];
            g4.hi_x[c] = k.b[3]; g4.hi_y[c] = k.b[4]; g4.hi_z[c] = k.b[5];
            if (k.ref < 0) {
                g4.child[c] = k.ref;
            } else {
                const int s = (int)n4->size();
                n4->push_back(GpuNode4());
                g4.child[c] = s;
                st.push_back({k.ref, s, dep + 1});
            }
        }
        (*n4)[slot] = g4;
    }
    return max_depth;
}

bool quantize_width4(const std::vector<GpuNode4> &n4, std::vector<GpuNode4Q> *out) {
    out->resize(n4.size());
    for (size_t i = 0; i < n4.size(); ++i)
        if (!quantize_node(n4[i], &(*out)[i])) return false;
    return true;
}

}  // namespace illume
