// Device-side path tracing shared by the megakernel (render.hip) and the
// wavefront kernels (wavefront.hip): scene queries, hit interactions,
// estimate_direct, camera rays and one bounce of PathIntegrator::li with an
// optional learned proposal.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include "shade.h"
#include "trace.h"

#pragma clang fp contract(off)

namespace illume {

struct PathCounters {
    uint32_t closest, any;
};

template <int BLOCK>
__device__ __forceinline__ bool scene_closest(const DevScene &sc, f3 o, f3 d, float t_max, TraceHit &h,
                                              TraceStack<BLOCK> &st, PathCounters &pc) {
    uint32_t nb = 0, np = 0;
    ++pc.closest;
    return trace_ray4<false, false, BLOCK>(sc.nodes, sc.prims, o, d, t_max, h, st, nb, np);
}
template <int BLOCK>
__device__ __forceinline__ bool scene_any(const DevScene &sc, f3 o, f3 d, float t_max, TraceStack<BLOCK> &st,
                                          PathCounters &pc) {
    uint32_t nb = 0, np = 0;
    TraceHit h;
    ++pc.any;
    return trace_ray4<true, false, BLOCK>(sc.nodes, sc.prims, o, d, t_max, h, st, nb, np);
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

// Analytic sky (an infinite area light with an analytic radiance):
// L(w) = zenith * t + horizon * (1 - t), t = sqrt(w.y), for w.y > 0 and 0
// below the horizon; sampled cosine-weighted about +y (pdf w.y / pi).
__device__ __forceinline__ f3 sky_L(const DevScene &sc, f3 w_in) {
    const f3 w = normalize(w_in);
    if (!(w.y > 0.f)) return mk(0.f, 0.f, 0.f);
    const float t = sqrtf(w.y);
    return mk(sc.sky_zenith[0] * t + sc.sky_horizon[0] * (1 - t), sc.sky_zenith[1] * t + sc.sky_horizon[1] * (1 - t),
              sc.sky_zenith[2] * t + sc.sky_horizon[2] * (1 - t));
}
__device__ __forceinline__ float sky_pdf(f3 wi) { return wi.y > 0.f ? wi.y * kInvPi : 0.f; }

// The two rays of estimate_direct (integrator.rs:353-494), found before either
// is traced: the light sample's shadow ray with the term it adds when
// unoccluded, and the BSDF sample's ray with its factor f |cos| and MIS
// weight; the radiance that ray finds is looked up once it is traced. The
// megakernel traces both at once (direct_resolve); the wavefront renderer
// queues them for a traversal kernel of their own.
struct DirectRays {
    f3 so, sd;       // shadow ray: any hit over (0, s_tmax)
    float s_tmax;
    f3 cs;           // the light sample's term when the shadow ray is unoccluded
    f3 mo, md;       // BSDF-sample ray: closest hit over (0, inf)
    f3 mf;           // f |cos| of the BSDF sample
    float mw, mpdf;  // its MIS weight and pdf
    int li;          // the sampled light (sc.sky_index for the sky)
    bool shadow, mis;
};

// The sky: the light sample's shadow ray leaves the scene (any hit over
// (0, inf)); the BSDF sample counts only if its ray escapes.
__device__ inline void direct_rays_sky(const DevScene &sc, const Isect &it, const Bsdf &bsdf, float us0, float us1,
                                       float ul0, float ul1, DirectRays &r) {
    const int flags = kAll & ~kSpecular;
    r.li = sc.sky_index;
    r.shadow = r.mis = false;
    const f3 dl = cosine_sample_hemisphere(ul0, ul1);
    f3 wi = mk(dl.x, dl.z, dl.y);
    float light_pdf = sky_pdf(wi);
    float scattering_pdf = 0.f;
    const f3 li_v = light_pdf > 0.f ? sky_L(sc, wi) : mk(0.f, 0.f, 0.f);
    if (light_pdf > 0 && !is_black(li_v)) {
        const f3 fv = bsdf.f(it.wo, wi, flags);
        const float ad = abs_dot(wi, it.shading_n);
        const f3 f = mk(fv.x * ad, fv.y * ad, fv.z * ad);
        scattering_pdf = bsdf.pdf(it.wo, wi, flags);
        if (!is_black(f)) {
            spawn_ray(it, wi, r.so);
            r.sd = wi;
            r.s_tmax = __builtin_inff();
            const float weight = power_heuristic(light_pdf, scattering_pdf);
            const f3 c = cmul(f, li_v);
            r.cs = mk(c.x * weight / light_pdf, c.y * weight / light_pdf, c.z * weight / light_pdf);
            r.shadow = true;
        }
    }
    int sampled = 0;
    f3 f = bsdf.sample_f(it.wo, wi, us0, us1, scattering_pdf, flags, sampled);
    const float ad = abs_dot(wi, it.shading_n);
    f = mk(f.x * ad, f.y * ad, f.z * ad);
    if (!is_black(f) && scattering_pdf > 0) {
        float weight = 1;
        if (!(sampled & kSpecular)) {
            light_pdf = sky_pdf(wi);
            if (light_pdf == 0) return;
            weight = power_heuristic(scattering_pdf, light_pdf);
        }
        spawn_ray(it, wi, r.mo);
        r.md = wi;
        r.mf = f;
        r.mw = weight;
        r.mpdf = scattering_pdf;
        r.mis = true;
    }
}

// A one-sided triangle area light, with both MIS strategies (the
// BSDF-sampled term per pbrt-v3).
__device__ inline void direct_rays(const DevScene &sc, const Isect &it, const Bsdf &bsdf, float us0, float us1, int li,
                                   float ul0, float ul1, DirectRays &r) {
    if (li == sc.sky_index) {
        direct_rays_sky(sc, it, bsdf, us0, us1, ul0, ul1, r);
        return;
    }
    r.li = li;
    r.shadow = r.mis = false;
    const DevLight &light = sc.lights[li];
    const int flags = kAll & ~kSpecular;
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
        const f3 f = mk(fv.x * ad, fv.y * ad, fv.z * ad);
        scattering_pdf = bsdf.pdf(it.wo, wi, flags);
        if (!is_black(f)) {
            spawn_ray_to(it, lp, lpe, ln, r.so, r.sd);
            r.s_tmax = 1.f - 0.0001f;
            const float weight = power_heuristic(light_pdf, scattering_pdf);
            const f3 c = cmul(f, li_v);
            r.cs = mk(c.x * weight / light_pdf, c.y * weight / light_pdf, c.z * weight / light_pdf);
            r.shadow = true;
        }
    }
    int sampled = 0;
    f3 f = bsdf.sample_f(it.wo, wi, us0, us1, scattering_pdf, flags, sampled);
    const float ad = abs_dot(wi, it.shading_n);
    f = mk(f.x * ad, f.y * ad, f.z * ad);
    if (!is_black(f) && scattering_pdf > 0) {
        float weight = 1;
        if (!(sampled & kSpecular)) {
            // Shape::pdf_with_point (shape.rs:60-81): the light's own triangle only
            f3 so;
            spawn_ray(it, wi, so);
            float t, c0, c1, c2;
            if (!triangle_test(p0, p1, p2, so, wi, __builtin_inff(), t, c0, c1, c2) ||
                triangle_degenerate(p0, p1, p2))
                return;
            Isect il;
            triangle_isect(p0, p1, p2, c0, c1, c2, wi, il);
            float lpdf = length_squared(sub(it.p, il.p)) / (abs_dot(il.n, neg(wi)) * light.area);
            if (__builtin_isinf(lpdf)) lpdf = 0.f;
            if (lpdf == 0) return;
            weight = power_heuristic(scattering_pdf, lpdf);
        }
        spawn_ray(it, wi, r.mo);
        r.md = wi;
        r.mf = f;
        r.mw = weight;
        r.mpdf = scattering_pdf;
        r.mis = true;
    }
}

// Traces the rays of direct_rays and returns estimate_direct's value: the
// light sample's term if its shadow ray is unoccluded, plus the BSDF sample's
// if its ray reaches the sampled light first (the sky: escapes the scene).
template <int BLOCK>
__device__ inline f3 direct_resolve(const DevScene &sc, const DirectRays &r, TraceStack<BLOCK> &st,
                                    PathCounters &pc) {
    f3 ld = mk(0.f, 0.f, 0.f);
    if (r.shadow && !scene_any<BLOCK>(sc, r.so, r.sd, r.s_tmax, st, pc)) ld = add(ld, r.cs);
    if (r.mis) {
        TraceHit h;
        const bool found = scene_closest<BLOCK>(sc, r.mo, r.md, __builtin_inff(), h, st, pc);
        f3 li2 = mk(0.f, 0.f, 0.f);
        if (r.li == sc.sky_index) {
            if (!found) li2 = sky_L(sc, r.md);
        } else if (found && sc.prim_light[h.prim] == r.li) {
            Isect lh;
            hit_isect(sc, h, r.mo, r.md, lh);
            li2 = area_L(sc.lights[r.li], lh.n, neg(r.md));
        }
        if (!is_black(li2)) {
            const f3 c = cmul(r.mf, li2);
            ld = add(ld, mk(c.x * r.mw / r.mpdf, c.y * r.mw / r.mpdf, c.z * r.mw / r.mpdf));
        }
    }
    return ld;
}

// uniform_sample_one_light's term added to the path radiance:
// L + beta (ld / lpdf).
__device__ __forceinline__ f3 add_direct(f3 L, f3 beta, f3 ld, float lpdf) {
    const f3 ldir = mk(ld.x / lpdf, ld.y / lpdf, ld.z / lpdf);
    return add(L, cmul(beta, ldir));
}

// A vertex's direct-lighting rays left for a later traversal stage, with what
// add_direct needs once they are traced.
struct DeferredDirect {
    DirectRays r;
    f3 beta;
    float lpdf;
    bool pending;
};


// PerspectiveCamera::generate_ray (perspective.rs:103-126, lens radius 0) for
// film position (fx, fy), then camera_to_world applied to the ray with the
// origin advanced past its error bound (transform.rs:365-389).
__device__ __forceinline__ void camera_ray(const DevCamera &cam, float fx, float fy, f3 &o, f3 &d, float &t_max) {
    const float *m = cam.raster_to_camera;
    float px = m[0] * fx + m[1] * fy + m[2] * 0.f + m[3];
    float py = m[4] * fx + m[5] * fy + m[6] * 0.f + m[7];
    float pz = m[8] * fx + m[9] * fy + m[10] * 0.f + m[11];
    float pw = m[12] * fx + m[13] * fy + m[14] * 0.f + m[15];
    f3 pcam = (pw == 1.f) ? mk(px, py, pz) : mk(px / pw, py / pw, pz / pw);
    f3 dc = normalize(pcam);
    const float *c = cam.camera_to_world;
    o = mk((c[0] * 0.f + c[1] * 0.f) + (c[2] * 0.f + c[3]), (c[4] * 0.f + c[5] * 0.f) + (c[6] * 0.f + c[7]),
           (c[8] * 0.f + c[9] * 0.f) + (c[10] * 0.f + c[11]));
    f3 oe = smul(gamma_f(3), mk(fabsf(c[0] * 0.f) + fabsf(c[1] * 0.f) + fabsf(c[2] * 0.f) + fabsf(c[3]),
                                fabsf(c[4] * 0.f) + fabsf(c[5] * 0.f) + fabsf(c[6] * 0.f) + fabsf(c[7]),
                                fabsf(c[8] * 0.f) + fabsf(c[9] * 0.f) + fabsf(c[10] * 0.f) + fabsf(c[11])));
    d = mk(c[0] * dc.x + c[1] * dc.y + c[2] * dc.z, c[4] * dc.x + c[5] * dc.y + c[6] * dc.z,
           c[8] * dc.x + c[9] * dc.y + c[10] * dc.z);
    t_max = __builtin_inff();
    float ls = length_squared(d);
    if (ls > 0) {
        float dt = dot(fabs3(d), oe) / ls;
        o = add(o, mul(d, dt));
        t_max -= dt;
    }
}

// ---------------------------------------------------------------- learned proposal
// The proposal's arithmetic uses the hardware exp2 / log2 / reciprocal /
// sin / cos (sin and cos take revolutions): its sampler and its pdf are the
// same formulas evaluated from the same decoded lobes, so the one-sample MIS
// estimator divides by the density the directions are drawn from up to float
// rounding, whatever the precision of the single operations.
constexpr float kLog2e = 1.44269504088896340736f;
constexpr float kLn2 = 0.69314718055994530942f;
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * kLog2e); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
// 1 - exp(-2k) for k >= 0, with the series where the difference cancels
__device__ __forceinline__ float one_minus_exp_m2k(float k) {
    return k < 1e-3f ? 2.f * k * (1.f - k) : 1.f - fast_exp(-2.f * k);
}
// vMF normalisation on the sphere: k / (2 pi (1 - exp(-2k))), 1 / (4 pi) at k = 0
__device__ __forceinline__ float vmf_norm(float k) {
    return k > 0.f ? k * fast_rcp(6.28318530717958647692f * one_minus_exp_m2k(k)) : 0.0795774715459476679f;
}

// 8-lobe vMF mixture decoded from the network's 40 outputs, stored as fp16
// (softmax weights, normalised means, concentrations exp(clamp(r, -4.6,
// 6.9))); c[j] = pi[j] * vmf_norm(kappa[j]), the lobe's density factor.
// The decoded lobes live in LDS, not registers: the shading code around them
// needs every register it has at three waves per SIMD. A block keeps
// [8][STRIDE] float4 (mu, kappa) and [8][STRIDE] float2 (pi, c), lobe-major
// and thread-minor, so a wave's accesses to one lobe are contiguous; each
// thread reaches only its own column.
template <int STRIDE>
struct Mixture {
    float4 *mk;  // this thread's column of the (mu, kappa) array
    float2 *pc;  // this thread's column of the (pi, c) array
    __device__ float4 &lobe(int j) const { return mk[j * STRIDE]; }
    __device__ float2 &weight(int j) const { return pc[j * STRIDE]; }
};
// LDS bytes of a block's mixtures
template <int STRIDE>
constexpr uint32_t kMixtureLds = 8 * STRIDE * (sizeof(float4) + sizeof(float2));

template <int STRIDE>
__device__ __forceinline__ bool decode_mixture(const __half *__restrict__ yh, const Mixture<STRIDE> &m) {
    float y[40];
    const uint4 *src = reinterpret_cast<const uint4 *>(yh);  // 80 bytes, 16-byte aligned rows
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const uint4 v = src[q];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            y[8 * q + 2 * e] = __half2float(__ushort_as_half((unsigned short)(w[e] & 0xffffu)));
            y[8 * q + 2 * e + 1] = __half2float(__ushort_as_half((unsigned short)(w[e] >> 16)));
        }
    }
    float mx = y[0];
#pragma unroll
    for (int j = 1; j < 8; ++j) mx = fmaxf(mx, y[5 * j]);
    float pi[8], s = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        pi[j] = fast_exp(y[5 * j] - mx);
        s += pi[j];
    }
    bool ok = s > 0.f && __builtin_isfinite(s);
    const float inv_s = fast_rcp(s);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float p = pi[j] * inv_s;
        const float a = y[5 * j + 1], b = y[5 * j + 2], c = y[5 * j + 3];
        const float len2 = a * a + b * b + c * c;
        const float inv = len2 > 0.f ? __builtin_amdgcn_rsqf(len2) : 0.f;
        const float kappa = fast_exp(fminf(fmaxf(y[5 * j + 4], -4.6f), 6.9f));
        m.lobe(j) = make_float4(a * inv, b * inv, len2 > 0.f ? c * inv : 1.f, kappa);
        m.weight(j) = make_float2(p, p * vmf_norm(kappa));
        ok = ok && __builtin_isfinite(p) && __builtin_isfinite(len2) && __builtin_isfinite(kappa);
    }
    return ok;
}
template <int STRIDE>
__device__ __forceinline__ float mixture_pdf(const Mixture<STRIDE> &m, f3 w) {
    float q = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float4 l = m.lobe(j);
        const float c = l.x * w.x + l.y * w.y + l.z * w.z;
        q += m.weight(j).y * fast_exp(l.w * (c - 1.f));
    }
    return q;
}
// Lobe by ul, then the vMF sample of Jakob (2012) around the lobe mean.
template <int STRIDE>
__device__ __forceinline__ f3 mixture_sample(const Mixture<STRIDE> &m, float ul, float u1, float u2) {
    // the first lobe whose cumulative weight exceeds ul (the last if none)
    int sel = 7;
    float c = 0.f;
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        c += m.weight(j).x;
        if (sel == 7 && ul < c) sel = j;
    }
    const float4 l = m.lobe(sel);
    const float k = l.w;
    const f3 mu = mk(l.x, l.y, l.z);
    float w = 1.f + __builtin_amdgcn_logf(u1 + (1.f - u1) * fast_exp(-2.f * k)) * kLn2 * fast_rcp(k);
    w = fminf(1.f, fmaxf(-1.f, w));
    const float s = __builtin_amdgcn_sqrtf(fmaxf(0.f, 1.f - w * w));
    f3 t1, t2;
    if (fabsf(mu.x) > fabsf(mu.y)) {
        const float il = __builtin_amdgcn_rsqf(mu.x * mu.x + mu.z * mu.z);
        t1 = mk(-mu.z * il, 0.f, mu.x * il);
    } else {
        const float il = __builtin_amdgcn_rsqf(mu.y * mu.y + mu.z * mu.z);
        t1 = mk(0.f, mu.z * il, -mu.y * il);
    }
    t2 = mk(mu.y * t1.z - mu.z * t1.y, mu.z * t1.x - mu.x * t1.z, mu.x * t1.y - mu.y * t1.x);
    // sin / cos of 2 pi u2 (the hardware takes revolutions)
    const float lx = s * __builtin_amdgcn_cosf(u2), ly = s * __builtin_amdgcn_sinf(u2);
    return normalize(add(add(smul(lx, t1), smul(ly, t2)), smul(w, mu)));
}

// Product of the mixture with a vMF fit of the vertex's cosine-weighted BSDF
// lobe (mean mu_b, concentration k_b). Each lobe's product is a vMF with
// k mu = k_j mu_j + k_b mu_b; its weight is scaled by the product's integral
// c(k_j) c(k_b) exp(k - k_j - k_b) / c(k), c(k) = vmf_norm(k). The result is
// again a normalised mixture, so its pdf is exact. False if the weights
// degenerate.
template <int STRIDE>
__device__ __forceinline__ bool mixture_times_lobe(const Mixture<STRIDE> &m, f3 mu_b, float k_b) {
    float s = 0.f;
    const float cb = vmf_norm(k_b);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float4 l = m.lobe(j);
        const float kj = l.w;
        const float vx = kj * l.x + k_b * mu_b.x, vy = kj * l.y + k_b * mu_b.y, vz = kj * l.z + k_b * mu_b.z;
        const float k = fmaxf(__builtin_amdgcn_sqrtf(vx * vx + vy * vy + vz * vz), 1e-4f);
        const float ck = vmf_norm(k);
        const float ik = fast_rcp(k);
        const float p = m.weight(j).x * (vmf_norm(kj) * cb * fast_exp(fminf(k - kj - k_b, 0.f)) * fast_rcp(ck));
        m.lobe(j) = make_float4(vx * ik, vy * ik, vz * ik, k);
        m.weight(j) = make_float2(p, ck);  // c times the normalised weight below
        s += p;
    }
    if (!(s > 0.f) || !__builtin_isfinite(s)) return false;
    const float inv = fast_rcp(s);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float2 wc = m.weight(j);
        const float p = wc.x * inv;
        m.weight(j) = make_float2(p, wc.y * p);
    }
    return true;
}

struct GuideCfg {
    const __half *out;        // [paths][40] network outputs (fp16), null = plain path tracing
    float frac;               // probability of sampling the proposal (one-sample MIS)
    float bounds_lo[3], inv_ext[3];
    float min_alpha;          // microfacet vertices below this roughness are not guided
    uint32_t product;         // sample the proposal times a vMF fit of the BSDF lobe
    int32_t net;              // the job's proposal network in a batch, -1 none
};

// Vertices where the proposal is queried and trained: diffuse, or microfacet
// with alpha >= min_alpha.
__device__ __forceinline__ bool guidable(const DevMaterial &m, float min_alpha) {
    if (m.type == kMetal) return m.alpha >= min_alpha;
    return m.type == kMatte && !(m.kd[0] == 0.f && m.kd[1] == 0.f && m.kd[2] == 0.f);
}

// One training vertex: proposal input (normalised position, normal), the
// sampled direction and its mixture pdf, the throughput after the vertex and
// the path radiance before anything arriving along that direction.
struct VertexRec {
    float in[6];
    float wi[3];
    float pdf;
    float beta[3];
    float lsnap[3];
};

struct PathState {
    f3 o, d;
    float t_max;
    f3 beta, L;
    float eta_scale;
    uint32_t bounces;
    bool specular, active;
    PathSampler rng;
};

// The camera-sample stage of SamplerIntegratorBase::render: sampler stream,
// get_camera_sample (film 2D, time 1D, lens 2D) and the camera ray, for
// sample s of the sample pixel (x, y) whose sampler stream is numbered
// `stream`; (fx, fy) gets the film position of the camera sample. With a
// low-discrepancy table the film offset is dimensions 0 and 1 of the pixel's
// global index (pbrt-v3 HaltonSampler / SobolSampler::SampleDimension) and
// later dimensions follow.
__device__ __forceinline__ void path_start_at(const DevCamera &cam, uint64_t seed, uint64_t stream, int x, int y,
                                              uint32_t s, PathState &ps, float &fx, float &fy,
                                              const SamplerLaunch &sl) {
    float uf0, uf1;
    if (sl.table) {
        uint64_t idx;
        if (sl.kind == kSamplerHalton) {
            idx = (uint64_t)halton_index(sl, x, y, (int64_t)s);
            uf0 = radical_inverse_base2(idx >> sl.exp0);
            uf1 = radical_inverse_prime(3, idx / (uint64_t)sl.scale1);
        } else {
            idx = sobol_index(sl.table, sl.log2_res, s, x - sl.x0, y - sl.y0);
            const float sx = sobol_value(sl.table, idx, 0) * sl.res + (float)sl.x0;
            const float sy = sobol_value(sl.table, idx, 1) * sl.res + (float)sl.y0;
            uf0 = fminf(fmaxf(sx - (float)x, 0.f), kOneMinusEpsilon);
            uf1 = fminf(fmaxf(sy - (float)y, 0.f), kOneMinusEpsilon);
        }
        ps.rng.lds = sl.table;
        ps.rng.state = idx;
        ps.rng.inc = 2;
    } else {
        ps.rng.seed(seed, stream, s);
        uf0 = ps.rng.get_1d();
        uf1 = ps.rng.get_1d();
    }
    ps.rng.get_1d();
    ps.rng.get_1d();
    ps.rng.get_1d();
    fx = (float)x + uf0;
    fy = (float)y + uf1;
    camera_ray(cam, fx, fy, ps.o, ps.d, ps.t_max);
    ps.beta = mk(1.f, 1.f, 1.f);
    ps.L = mk(0.f, 0.f, 0.f);
    ps.eta_scale = 1.f;
    ps.bounces = 0;
    ps.specular = false;
    ps.active = true;
}

// One bounce of PathIntegrator::li after the closest-hit query (found, h).
// With a proposal (gout != null, valid mixture, non-specular BSDF) the
// continuation is one-sample MIS: the mixture with probability frac, else the
// BSDF, weighted by frac * q + (1 - frac) * p_bsdf. recs (may be null) gets a
// training vertex per sampled continuation; n_recs counts them. mix holds
// this thread's mixture columns in LDS, required when gout is not null.
// DEFER: the direct-lighting rays are not traced here but returned in *dd
// (dd->pending), for the caller to trace and add with add_direct; st and pc
// are then unused.
template <int BLOCK, bool DEFER = false>
__device__ __forceinline__ void path_shade(const DevScene &sc, uint32_t max_depth, PathState &ps, bool found, const TraceHit &h,
                           const __half *__restrict__ gout, const GuideCfg &gc, VertexRec *recs, uint32_t &n_recs,
                           TraceStack<BLOCK> &st, PathCounters &pc, DeferredDirect *dd = nullptr,
                           Mixture<BLOCK> mix = {nullptr, nullptr}) {
    if (DEFER) dd->pending = false;
    if (recs && n_recs > 0 && recs[n_recs - 1].lsnap[0] < 0.f) {
        // the radiance gathered so far excludes everything arriving along the
        // last recorded direction
        VertexRec &r = recs[n_recs - 1];
        r.lsnap[0] = ps.L.x;
        r.lsnap[1] = ps.L.y;
        r.lsnap[2] = ps.L.z;
    }
    Isect it;
    if (found) hit_isect(sc, h, ps.o, ps.d, it);
    if (ps.bounces == 0 || ps.specular) {
        if (found) {
            int li = sc.prim_light[h.prim];
            if (li >= 0) ps.L = add(ps.L, cmul(ps.beta, area_L(sc.lights[li], it.n, neg(ps.d))));
        } else if (sc.sky_index >= 0) {
            ps.L = add(ps.L, cmul(ps.beta, sky_L(sc, ps.d)));  // Le of the escaped ray
        }
    }
    if (!found || ps.bounces >= max_depth) {
        ps.active = false;
        return;
    }
    Bsdf bsdf;
    bsdf.init(it, sc.materials[sc.prim_material[h.prim]]);
    const bool nonspec = bsdf.matches(kAll & ~kSpecular);
    if (nonspec && sc.n_lights > 0) {
        // uniform_sample_one_light (integrator.rs:321-350)
        float lpdf;
        const float ul = ps.rng.get_1d();
        int ln = light_sample_discrete(sc, ul, lpdf);
        if (lpdf != 0) {
            const float ul0 = ps.rng.get_1d(), ul1 = ps.rng.get_1d();
            const float us0 = ps.rng.get_1d(), us1 = ps.rng.get_1d();
            if (DEFER) {
                direct_rays(sc, it, bsdf, us0, us1, ln, ul0, ul1, dd->r);
                dd->beta = ps.beta;
                dd->lpdf = lpdf;
                dd->pending = true;
            } else {
                DirectRays r;
                direct_rays(sc, it, bsdf, us0, us1, ln, ul0, ul1, r);
                ps.L = add_direct(ps.L, ps.beta, direct_resolve<BLOCK>(sc, r, st, pc), lpdf);
            }
        }
    }
    f3 wo = neg(ps.d), wi = mk(0.f, 0.f, 0.f);
    float pdf = 0;
    int flags = 0;
    f3 f;
    bool guided = gout && nonspec && decode_mixture(gout, mix);
    if (guided && gc.product) {
        // the lobe about the mirror direction (microfacet: concentration 0.8 /
        // alpha^2, the half-maximum width of the reflected Trowbridge-Reitz
        // lobe) or about the normal (Lambertian: 1.4, the cosine's half maximum)
        const f3 ns = faceforward(it.shading_n, wo);
        const bool metal = bsdf.m.type == kMetal;
        const f3 mu_b = metal ? sub(smul(2.f * dot(wo, ns), ns), wo) : ns;
        const float k_b = metal ? fminf(0.8f / fmaxf(bsdf.m.alpha * bsdf.m.alpha, 1e-6f), 1000.f) : 1.4f;
        guided = mixture_times_lobe(mix, mu_b, k_b);
    }
    if (guided) {
        const float usel = ps.rng.get_1d();
        float p_bsdf = 0.f;
        if (usel < gc.frac) {
            const float ul = ps.rng.get_1d(), u1 = ps.rng.get_1d(), u2 = ps.rng.get_1d();
            wi = mixture_sample(mix, ul, u1, u2);
            f = bsdf.f(wo, wi, kAll);
            p_bsdf = bsdf.pdf(wo, wi, kAll);
            flags = bsdf.type & ~kSpecular;
        } else {
            const float u0 = ps.rng.get_1d(), u1 = ps.rng.get_1d();
            f = bsdf.sample_f(wo, wi, u0, u1, p_bsdf, kAll, flags);
            if (p_bsdf == 0.f) f = mk(0.f, 0.f, 0.f);
        }
        const float q = mixture_pdf(mix, wi);
        pdf = gc.frac * q + (1.f - gc.frac) * p_bsdf;
        if (!(pdf > 0.f) || !__builtin_isfinite(pdf)) pdf = 0.f;
    } else {
        const float u0 = ps.rng.get_1d(), u1 = ps.rng.get_1d();
        f = bsdf.sample_f(wo, wi, u0, u1, pdf, kAll, flags);
    }
    if (is_black(f) || pdf == 0.f) {
        ps.active = false;
        return;
    }
    float ad = abs_dot(wi, it.shading_n);
    f3 bf = cmul(ps.beta, f);
    ps.beta = mk(bf.x * ad / pdf, bf.y * ad / pdf, bf.z * ad / pdf);
    ps.specular = (flags & kSpecular) != 0;
    if ((flags & kSpecular) && (flags & kTrans)) {
        float eta = bsdf.eta;
        ps.eta_scale *= (dot(wo, it.n) > 0) ? (eta * eta) : 1 / (eta * eta);
    }
    if (recs && nonspec && n_recs < max_depth && guidable(bsdf.m, gc.min_alpha)) {
        VertexRec &r = recs[n_recs++];
        r.in[0] = (it.p.x - gc.bounds_lo[0]) * gc.inv_ext[0];
        r.in[1] = (it.p.y - gc.bounds_lo[1]) * gc.inv_ext[1];
        r.in[2] = (it.p.z - gc.bounds_lo[2]) * gc.inv_ext[2];
        r.in[3] = it.n.x;
        r.in[4] = it.n.y;
        r.in[5] = it.n.z;
        r.wi[0] = wi.x;
        r.wi[1] = wi.y;
        r.wi[2] = wi.z;
        r.pdf = pdf;
        r.beta[0] = ps.beta.x;
        r.beta[1] = ps.beta.y;
        r.beta[2] = ps.beta.z;
        r.lsnap[0] = r.lsnap[1] = r.lsnap[2] = -1.f;  // set when the next vertex is reached
    }
    spawn_ray(it, wi, ps.o);
    ps.d = wi;
    ps.t_max = __builtin_inff();
    f3 rr = mk(ps.beta.x * ps.eta_scale, ps.beta.y * ps.eta_scale, ps.beta.z * ps.eta_scale);
    ++ps.bounces;
    if (max_comp(rr) < 1.f && ps.bounces - 1 > 3) {
        float q = fmaxf(0.05f, 1 - max_comp(rr));
        if (ps.rng.get_1d() < q) {
            ps.active = false;
            return;
        }
        ps.beta = mk(ps.beta.x / (1 - q), ps.beta.y / (1 - q), ps.beta.z / (1 - q));
    }
}

// SamplerIntegratorBase::render's radiance guard (integrator.rs:113-122).
__device__ __forceinline__ f3 guard_radiance(f3 L) {
    if (__builtin_isnan(L.x) || __builtin_isnan(L.y) || __builtin_isnan(L.z)) return mk(0.f, 0.f, 0.f);
    if (luminance(L) < -1e-5f) return mk(0.f, 0.f, 0.f);
    if (__builtin_isinf(luminance(L))) return mk(0.f, 0.f, 0.f);
    return L;
}

}  // namespace illume
