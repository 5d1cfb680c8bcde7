// Path tracer: pbrt-v3's PathIntegrator (the reference has the
// SamplerIntegrator trait only, core/integrator.rs:31-35) with the
// reference's estimate_direct / uniform_sample_one_light MIS
// (integrator.rs:321-494), camera sampling (sampler.rs:78-84,
// perspective.rs:103-126) and the radiance guard of
// SamplerIntegratorBase::render (integrator.rs:113-122).
//
// One thread traces one camera path through all of its bounces (closest-hit
// extension rays, any-hit shadow rays and MIS rays share the traversal of
// trace.h). Lanes of a wave hold neighbouring pixels of the same sample
// index, so primary rays are coherent. Each finished path adds its radiance
// to the caller's float RGB sum buffer with float atomics; ray counts are
// reduced per block in LDS before one 64-bit atomic per block.
#include <hip/hip_runtime.h>
#include "kernels.h"
#include "shade.h"
#include "trace.h"

#pragma clang fp contract(off)

namespace illume {

namespace {

constexpr int kRenderBlock = 128;

struct PathCounters {
    uint32_t closest, any;
};

template <int BLOCK>
__device__ __forceinline__ bool scene_closest(const DevScene &sc, f3 o, f3 d, float t_max, TraceHit &h,
                                              TraceStack<BLOCK> &st, PathCounters &pc) {
    uint32_t nb = 0, np = 0;
    ++pc.closest;
    return trace_ray<false, false, BLOCK>(sc.nodes, sc.prims, o, d, t_max, h, st, nb, np);
}
template <int BLOCK>
__device__ __forceinline__ bool scene_any(const DevScene &sc, f3 o, f3 d, float t_max, TraceStack<BLOCK> &st,
                                          PathCounters &pc) {
    uint32_t nb = 0, np = 0;
    TraceHit h;
    ++pc.any;
    return trace_ray<true, false, BLOCK>(sc.nodes, sc.prims, o, d, t_max, h, st, nb, np);
}

// Builds the interaction of a closest hit.
__device__ __forceinline__ void hit_isect(const DevScene &sc, const TraceHit &h, f3 o, f3 d, Isect &it) {
    it.prim = h.prim;
    if (h.prim < sc.n_tris) {
        const float4 *v = sc.tri_verts + 3 * (size_t)h.prim;
        const float4 a = v[0], b = v[1], c = v[2];
        triangle_isect(mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), mk(c.x, c.y, c.z), h.b0, h.b1, h.b2, d, it);
    } else {
        const float4 s = sc.spheres[h.prim - sc.n_tris];
        sphere_isect(mk(s.x, s.y, s.z), s.w, o, d, h.t, it);
    }
}

// DiffuseAreaLight::L: one-sided emission along the geometric normal.
__device__ __forceinline__ f3 area_L(const DevLight &l, f3 n, f3 w) {
    return dot(n, w) > 0 ? ld3(l.L) : mk(0.f, 0.f, 0.f);
}

// sampling.rs:75-88 sample_discrete with pbrt-v3's FindInterval
__device__ __forceinline__ int light_sample_discrete(const DevScene &sc, float u, float &pdf) {
    int size = (int)sc.n_lights + 1, first = 0, len = size;
    while (len > 0) {
        int half = len >> 1, middle = first + half;
        if (sc.light_cdf[middle] <= u) {
            first = middle + 1;
            len -= half + 1;
        } else
            len = half;
    }
    int offset = min(max(first - 1, 0), size - 2);
    pdf = (sc.light_func_int > 0) ? sc.light_func[offset] / (sc.light_func_int * (float)sc.n_lights) : 0.f;
    return offset;
}

// estimate_direct (integrator.rs:353-494) for a one-sided triangle area light,
// with both MIS strategies (the BSDF-sampled term per pbrt-v3).
template <int BLOCK>
__device__ f3 estimate_direct(const DevScene &sc, const Isect &it, const Bsdf &bsdf, float us0, float us1, int li,
                              float ul0, float ul1, TraceStack<BLOCK> &st, PathCounters &pc) {
    const DevLight &light = sc.lights[li];
    const int flags = kAll & ~kSpecular;
    f3 ld = mk(0.f, 0.f, 0.f);
    const f3 p0 = ld3(light.p0), p1 = ld3(light.p1), p2 = ld3(light.p2);
    // Light::sample_li -> Triangle::sample (triangle.rs:616-642) and the
    // area-to-solid-angle conversion of shape.rs:43-57.
    float su0 = sqrtf(ul0);
    float b0 = 1 - su0, b1 = ul1 * su0;
    f3 lp = add(add(smul(b0, p0), smul(b1, p1)), smul(1 - b0 - b1, p2));
    f3 ln = normalize(cross(sub(p1, p0), sub(p2, p0)));
    f3 lpe = smul(gamma_f(6), add(add(fabs3(smul(b0, p0)), fabs3(smul(b1, p1))), fabs3(smul(1 - b0 - b1, p2))));
    float light_pdf = 1 / light.area;
    f3 w = sub(lp, it.p);
    if (length_squared(w) == 0)
        light_pdf = 0;
    else {
        w = normalize(w);
        light_pdf *= length_squared(sub(it.p, lp)) / abs_dot(ln, neg(w));
        if (__builtin_isinf(light_pdf)) light_pdf = 0.f;
    }
    f3 wi = mk(0.f, 0.f, 0.f);
    f3 li_v = mk(0.f, 0.f, 0.f);
    if (!(light_pdf == 0 || length_squared(sub(lp, it.p)) == 0)) {
        wi = normalize(sub(lp, it.p));
        li_v = area_L(light, ln, neg(wi));
    } else {
        light_pdf = 0;
    }
    float scattering_pdf = 0;
    if (light_pdf > 0 && !is_black(li_v)) {
        const f3 fv = bsdf.f(it.wo, wi, flags);
        const float ad = abs_dot(wi, it.shading_n);
        f3 f = mk(fv.x * ad, fv.y * ad, fv.z * ad);
        scattering_pdf = bsdf.pdf(it.wo, wi, flags);
        if (!is_black(f)) {
            f3 so, sd;
            spawn_ray_to(it, lp, lpe, ln, so, sd);
            if (scene_any<BLOCK>(sc, so, sd, 1.f - 0.0001f, st, pc)) li_v = mk(0.f, 0.f, 0.f);
            if (!is_black(li_v)) {
                float weight = power_heuristic(light_pdf, scattering_pdf);
                f3 c = cmul(f, li_v);
                ld = add(ld, mk(c.x * weight / light_pdf, c.y * weight / light_pdf, c.z * weight / light_pdf));
            }
        }
    }
    {
        int sampled = 0;
        f3 f = bsdf.sample_f(it.wo, wi, us0, us1, scattering_pdf, flags, sampled);
        float ad = abs_dot(wi, it.shading_n);
        f = mk(f.x * ad, f.y * ad, f.z * ad);
        bool sampled_specular = (sampled & kSpecular) != 0;
        if (!is_black(f) && scattering_pdf > 0) {
            float weight = 1;
            if (!sampled_specular) {
                // Shape::pdf_with_point (shape.rs:60-81): the light's own triangle only
                f3 so;
                spawn_ray(it, wi, so);
                float t, c0, c1, c2;
                if (!triangle_test(p0, p1, p2, so, wi, __builtin_inff(), t, c0, c1, c2) ||
                    triangle_degenerate(p0, p1, p2))
                    return ld;
                Isect il;
                triangle_isect(p0, p1, p2, c0, c1, c2, wi, il);
                float lpdf = length_squared(sub(it.p, il.p)) / (abs_dot(il.n, neg(wi)) * light.area);
                if (__builtin_isinf(lpdf)) lpdf = 0.f;
                light_pdf = lpdf;
                if (light_pdf == 0) return ld;
                weight = power_heuristic(scattering_pdf, light_pdf);
            }
            f3 so;
            spawn_ray(it, wi, so);
            TraceHit h;
            bool found = scene_closest<BLOCK>(sc, so, wi, __builtin_inff(), h, st, pc);
            f3 li2 = mk(0.f, 0.f, 0.f);
            if (found && sc.prim_light[h.prim] == li) {
                Isect lh;
                hit_isect(sc, h, so, wi, lh);
                li2 = area_L(light, lh.n, neg(wi));
            }
            if (!is_black(li2)) {
                f3 c = cmul(f, li2);
                ld = add(ld, mk(c.x * weight / scattering_pdf, c.y * weight / scattering_pdf,
                                c.z * weight / scattering_pdf));
            }
        }
    }
    return ld;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void render_kernel(DevScene sc, DevCamera cam, uint32_t spp, uint32_t max_depth,
                                                        uint64_t seed, uint32_t sample_offset, float *__restrict__ rgb,
                                                        unsigned long long *__restrict__ counters) {
    __shared__ int lds_stack[kStackLds * BLOCK];
    __shared__ uint32_t blk_counts[2];
    if (threadIdx.x < 2) blk_counts[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t n_pix = (uint64_t)cam.width * cam.height;
    const uint64_t gid = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    PathCounters pc{0, 0};
    if (gid < n_pix * spp) {
        const uint64_t pix = gid % n_pix;
        const uint32_t s = (uint32_t)(gid / n_pix) + sample_offset;
        const int x = (int)(pix % cam.width), y = (int)(pix / cam.width);
        TraceStack<BLOCK> st;
        st.reset(lds_stack + threadIdx.x);
        Pcg32 rng;
        rng.seed(seed, pix, s);
        // get_camera_sample: p_film (2D), time (1D), p_lens (2D)
        const float uf0 = rng.get_1d(), uf1 = rng.get_1d();
        const float time = rng.get_1d();
        (void)time;
        rng.get_1d();
        rng.get_1d();
        // PerspectiveCamera::generate_ray + camera_to_world applied to the ray
        const float fx = (float)x + uf0, fy = (float)y + uf1;
        const float *m = cam.raster_to_camera;
        float px = m[0] * fx + m[1] * fy + m[2] * 0.f + m[3];
        float py = m[4] * fx + m[5] * fy + m[6] * 0.f + m[7];
        float pz = m[8] * fx + m[9] * fy + m[10] * 0.f + m[11];
        float pw = m[12] * fx + m[13] * fy + m[14] * 0.f + m[15];
        f3 pcam = (pw == 1.f) ? mk(px, py, pz) : mk(px / pw, py / pw, pz / pw);
        f3 dc = normalize(pcam);
        const float *c = cam.camera_to_world;
        f3 o = mk((c[0] * 0.f + c[1] * 0.f) + (c[2] * 0.f + c[3]), (c[4] * 0.f + c[5] * 0.f) + (c[6] * 0.f + c[7]),
                  (c[8] * 0.f + c[9] * 0.f) + (c[10] * 0.f + c[11]));
        f3 oe = smul(gamma_f(3), mk(fabsf(c[0] * 0.f) + fabsf(c[1] * 0.f) + fabsf(c[2] * 0.f) + fabsf(c[3]),
                                    fabsf(c[4] * 0.f) + fabsf(c[5] * 0.f) + fabsf(c[6] * 0.f) + fabsf(c[7]),
                                    fabsf(c[8] * 0.f) + fabsf(c[9] * 0.f) + fabsf(c[10] * 0.f) + fabsf(c[11])));
        f3 d =mk(c[0] * dc.x + c[1] * dc.y + c[2] * dc.z, c[4] * dc.x + c[5] * dc.y + c[6] * dc.z,
                  c[8] * dc.x + c[9] * dc.y + c[10] * dc.z);
        float t_max = __builtin_inff();
        float ls = length_squared(d);
        if (ls > 0) {
            float dt = dot(fabs3(d), oe) / ls;
            o = add(o, mul(d, dt));
            t_max -= dt;
        }

        // PathIntegrator::li
        f3 L = mk(0.f, 0.f, 0.f), beta = mk(1.f, 1.f, 1.f);
        bool specular_bounce = false;
        float eta_scale = 1;
        for (uint32_t bounces = 0;; ++bounces) {
            TraceHit h;
            const bool found = scene_closest<BLOCK>(sc, o, d, t_max, h, st, pc);
            Isect it;
            if (found) hit_isect(sc, h, o, d, it);
            if (bounces == 0 || specular_bounce) {
                if (found) {
                    int li = sc.prim_light[h.prim];
                    if (li >= 0) L = add(L, cmul(beta, area_L(sc.lights[li], it.n, neg(d))));
                }
            }
            if (!found || bounces >= max_depth) break;
            Bsdf bsdf;
            bsdf.init(it, sc.materials[sc.prim_material[h.prim]]);
            if (bsdf.matches(kAll & ~kSpecular) && sc.n_lights > 0) {
                // uniform_sample_one_light (integrator.rs:321-350)
                float lpdf;
                const float ul = rng.get_1d();
                int ln = light_sample_discrete(sc, ul, lpdf);
                if (lpdf != 0) {
                    const float ul0 = rng.get_1d(), ul1 = rng.get_1d();
                    const float us0 = rng.get_1d(), us1 = rng.get_1d();
                    f3 ldir = estimate_direct<BLOCK>(sc, it, bsdf, us0, us1, ln, ul0, ul1, st, pc);
                    ldir = mk(ldir.x / lpdf, ldir.y / lpdf, ldir.z / lpdf);
                    L = add(L, cmul(beta, ldir));
                }
            }
            f3 wo = neg(d), wi = mk(0.f, 0.f, 0.f);
            float pdf = 0;
            int flags = 0;
            const float u0 = rng.get_1d(), u1 = rng.get_1d();
            f3 f = bsdf.sample_f(wo, wi, u0, u1, pdf, kAll, flags);
            if (is_black(f) || pdf == 0.f) break;
            float ad = abs_dot(wi, it.shading_n);
            f3 bf = cmul(beta, f);
            beta = mk(bf.x * ad / pdf, bf.y * ad / pdf, bf.z * ad / pdf);
            specular_bounce = (flags & kSpecular) != 0;
            if ((flags & kSpecular) && (flags & kTrans)) {
                float eta = bsdf.eta;
                eta_scale *= (dot(wo, it.n) > 0) ? (eta * eta) : 1 / (eta * eta);
            }
            spawn_ray(it, wi, o);
            d = wi;
            t_max = __builtin_inff();
            f3 rr = mk(beta.x * eta_scale, beta.y * eta_scale, beta.z * eta_scale);
            if (max_comp(rr) < 1.f && bounces > 3) {
                float q = fmaxf(0.05f, 1 - max_comp(rr));
                if (rng.get_1d() < q) break;
                beta = mk(beta.x / (1 - q), beta.y / (1 - q), beta.z / (1 - q));
            }
        }
        // radiance guard (integrator.rs:113-122)
        if (__builtin_isnan(L.x) || __builtin_isnan(L.y) || __builtin_isnan(L.z))
            L = mk(0.f, 0.f, 0.f);
        else if (luminance(L) < -1e-5f)
            L = mk(0.f, 0.f, 0.f);
        else if (__builtin_isinf(luminance(L)))
            L = mk(0.f, 0.f, 0.f);
        atomicAdd(rgb + 3 * pix, L.x);
        atomicAdd(rgb + 3 * pix + 1, L.y);
        atomicAdd(rgb + 3 * pix + 2, L.z);
    }
    if (counters) {
        atomicAdd(&blk_counts[0], pc.closest);
        atomicAdd(&blk_counts[1], pc.any);
        __syncthreads();
        if (threadIdx.x == 0) {
            atomicAdd(counters, (unsigned long long)blk_counts[0]);
            atomicAdd(counters + 1, (unsigned long long)blk_counts[1]);
        }
    }
}

}  // namespace

hipError_t launch_render(const DevScene &sc, const DevCamera &cam, uint32_t spp, uint32_t max_depth, uint64_t seed,
                         uint32_t sample_offset, float *rgb, unsigned long long *counters, hipStream_t stream) {
    const uint64_t n = (uint64_t)cam.width * cam.height * spp;
    if (n == 0) return hipSuccess;
    const uint64_t grid = (n + kRenderBlock - 1) / kRenderBlock;
    if (grid > 0x7fffffffULL) return hipErrorInvalidValue;
    render_kernel<kRenderBlock><<<(uint32_t)grid, kRenderBlock, 0, stream>>>(sc, cam, spp, max_depth, seed,
                                                                             sample_offset, rgb, counters);
    return hipGetLastError();
}

}  // namespace illume
