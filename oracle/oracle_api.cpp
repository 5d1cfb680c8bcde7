// ORACLE — test infrastructure only (see pbrt.h). C entry points for the
// tests (ctypes) and for the CPU-baseline leg of bench.py.
#include <thread>
#include <atomic>
#include <vector>
#include "bvh.h"

using namespace orc;

namespace {
struct OrcRay {
    float o[3], d[3], t_max, time;
};
struct OrcHit {
    uint32_t prim;
    float t, b1, b2;
};
struct OrcScene {
    PrimSet ps;
    BVH bvh;
};
inline Ray to_ray(const OrcRay &r) { return Ray(V3(r.o[0], r.o[1], r.o[2]), V3(r.d[0], r.d[1], r.d[2]), r.t_max, r.time); }

template <typename F>
void parallel_chunks(uint64_t n, int n_threads, F f) {
    if (n_threads <= 1 || n < 4096) {
        f(0, n, 0);
        return;
    }
    std::atomic<uint64_t> next(0);
    const uint64_t chunk = 1 << 14;
    std::vector<std::thread> th;
    for (int t = 0; t < n_threads; ++t)
        th.emplace_back([&, t]() {
            while (true) {
                uint64_t s = next.fetch_add(chunk);
                if (s >= n) break;
                f(s, std::min(n, s + chunk), t);
            }
        });
    for (auto &x : th) x.join();
}
}  // namespace

extern "C" {

// spheres: 4 floats each (cx, cy, cz, radius)
void *orc_scene_create(const float *verts, uint32_t n_verts, const uint32_t *idx, uint32_t n_tris, const float *spheres,
                       uint32_t n_spheres, int max_prims_in_node, int reverse_orientation) {
    OrcScene *s = new OrcScene();
    s->ps.verts.resize(n_verts);
    for (uint32_t i = 0; i < n_verts; ++i) s->ps.verts[i] = V3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    s->ps.idx.assign(idx, idx + 3 * (size_t)n_tris);
    s->ps.spheres.resize(n_spheres);
    for (uint32_t i = 0; i < n_spheres; ++i)
        s->ps.spheres[i].init(V3(spheres[4 * i], spheres[4 * i + 1], spheres[4 * i + 2]), spheres[4 * i + 3]);
    s->ps.reverse_orientation = reverse_orientation != 0;
    s->bvh.build(&s->ps, max_prims_in_node);
    return s;
}

void orc_scene_destroy(void *h) { delete (OrcScene *)h; }

// out[0] = node count, out[1] = max depth, out[2] = primitive count
void orc_scene_info(void *h, int64_t *out) {
    OrcScene *s = (OrcScene *)h;
    out[0] = (int64_t)s->bvh.nodes.size();
    out[1] = s->bvh.max_depth;
    out[2] = s->ps.n_prims();
}

// Flattened nodes: bounds (6 floats), offset, n_primitives, axis.
void orc_scene_nodes(void *h, float *bounds, int32_t *offset, int32_t *n_prims, int32_t *axis, uint32_t *ordered) {
    OrcScene *s = (OrcScene *)h;
    for (size_t i = 0; i < s->bvh.nodes.size(); ++i) {
        const LinearBVHNode &n = s->bvh.nodes[i];
        bounds[6 * i + 0] = n.bounds.pmin.x; bounds[6 * i + 1] = n.bounds.pmin.y; bounds[6 * i + 2] = n.bounds.pmin.z;
        bounds[6 * i + 3] = n.bounds.pmax.x; bounds[6 * i + 4] = n.bounds.pmax.y; bounds[6 * i + 5] = n.bounds.pmax.z;
        offset[i] = n.offset;
        n_prims[i] = n.n_primitives;
        axis[i] = n.axis;
    }
    for (size_t i = 0; i < s->bvh.ordered.size(); ++i) ordered[i] = s->bvh.ordered[i];
}

// Closest hit for each ray (Scene::intersect, scene.rs:43-47). brute != 0
// loops over all primitives instead of the BVH. stats (may be null):
// [0] node tests, [1] primitive tests.
void orc_intersect(void *h, const void *rays, uint64_t n, void *hits, int brute, int n_threads, uint64_t *stats) {
    OrcScene *s = (OrcScene *)h;
    const OrcRay *r = (const OrcRay *)rays;
    OrcHit *o = (OrcHit *)hits;
    int nt = std::max(1, n_threads);
    std::vector<TraversalStats> st(nt);
    parallel_chunks(n, nt, [&](uint64_t b, uint64_t e, int t) {
        for (uint64_t i = b; i < e; ++i) {
            Ray ray = to_ray(r[i]);
            SurfaceInteraction si;
            uint32_t p = brute ? brute_intersect(s->ps, ray, &si) : s->bvh.intersect(ray, &si, &st[t]);
            o[i].prim = p;
            o[i].t = (p == 0xffffffffu) ? 0.f : si.t;
            o[i].b1 = (p == 0xffffffffu) ? 0.f : si.b1;
            o[i].b2 = (p == 0xffffffffu) ? 0.f : si.b2;
        }
    });
    if (stats) {
        stats[0] = stats[1] = 0;
        for (auto &x : st) {
            stats[0] += x.node_tests;
            stats[1] += x.prim_tests;
        }
    }
}

// Any hit for each ray (Scene::intersect_p, scene.rs:49-53).
void orc_intersect_p(void *h, const void *rays, uint64_t n, uint8_t *occluded, int brute, int n_threads,
                     uint64_t *stats) {
    OrcScene *s = (OrcScene *)h;
    const OrcRay *r = (const OrcRay *)rays;
    int nt = std::max(1, n_threads);
    std::vector<TraversalStats> st(nt);
    parallel_chunks(n, nt, [&](uint64_t b, uint64_t e, int t) {
        for (uint64_t i = b; i < e; ++i) {
            Ray ray = to_ray(r[i]);
            occluded[i] = (brute ? brute_intersect_p(s->ps, ray) : s->bvh.intersect_p(ray, &st[t])) ? 1 : 0;
        }
    });
    if (stats) {
        stats[0] = stats[1] = 0;
        for (auto &x : st) {
            stats[0] += x.node_tests;
            stats[1] += x.prim_tests;
        }
    }
}

// Single-primitive probes. Returns 1 on a hit; out = (t, b0, b1, b2).
int orc_triangle_test(const float *p9, const void *ray, float *out) {
    TriHit h;
    V3 p0(p9[0], p9[1], p9[2]), p1(p9[3], p9[4], p9[5]), p2(p9[6], p9[7], p9[8]);
    if (!triangle_test(p0, p1, p2, to_ray(*(const OrcRay *)ray), &h)) return 0;
    out[0] = h.t; out[1] = h.b0; out[2] = h.b1; out[3] = h.b2;
    return 1;
}
// out = (t, p.xyz, n.xyz, p_error.xyz)
int orc_sphere_intersect(const float *c4, const void *ray, float *out) {
    SphereShape s;
    s.init(V3(c4[0], c4[1], c4[2]), c4[3]);
    SurfaceInteraction si;
    float t;
    if (!sphere_intersect(s, to_ray(*(const OrcRay *)ray), &si, &t)) return 0;
    out[0] = t;
    out[1] = si.p.x; out[2] = si.p.y; out[3] = si.p.z;
    out[4] = si.n.x; out[5] = si.n.y; out[6] = si.n.z;
    out[7] = si.p_error.x; out[8] = si.p_error.y; out[9] = si.p_error.z;
    return 1;
}
// EFloat quadratic probe: returns 1 when real roots exist; out = t0 (v, lo, hi), t1 (v, lo, hi).
int orc_quadratic(const float *abc_v_err, float *out) {
    EFloat a(abc_v_err[0], abc_v_err[1]), b(abc_v_err[2], abc_v_err[3]), c(abc_v_err[4], abc_v_err[5]);
    EFloat t0, t1;
    if (!quadratic(a, b, c, &t0, &t1)) return 0;
    out[0] = t0.v; out[1] = t0.low; out[2] = t0.high;
    out[3] = t1.v; out[4] = t1.low; out[5] = t1.high;
    return 1;
}

// One primitive against one ray with closest-hit semantics
// (GeometricPrimitive::intersect); returns 1 and its t on a hit.
int orc_prim_intersect(void *h, uint32_t prim, const void *ray, float *t) {
    OrcScene *s = (OrcScene *)h;
    if (prim >= s->ps.n_prims()) return 0;
    Ray r = to_ray(*(const OrcRay *)ray);
    SurfaceInteraction si;
    if (!s->ps.intersect(prim, r, &si)) return 0;
    *t = si.t;
    return 1;
}

float orc_gamma(int n) { return gamma(n); }
float orc_next_float_up(float v) { return next_float_up(v); }
float orc_next_float_down(float v) { return next_float_down(v); }
void orc_offset_ray_origin(const float *p, const float *pe, const float *n, const float *w, float *out) {
    V3 r = offset_ray_origin(V3(p[0], p[1], p[2]), V3(pe[0], pe[1], pe[2]), V3(n[0], n[1], n[2]), V3(w[0], w[1], w[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// PCG32 (rng.rs): n outputs of uniform_uint32 from Rng::new(seq) after advance(skip).
void orc_pcg32(uint64_t seq, int64_t skip, uint32_t n, uint32_t *out) {
    Rng r(seq);
    if (skip) r.advance(skip);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.uniform_uint32();
}
void orc_pcg32_float(uint64_t seq, uint32_t n, float *out) {
    Rng r(seq);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.uniform_float();
}

int orc_hw_threads() { return (int)std::thread::hardware_concurrency(); }

}  // extern "C"
