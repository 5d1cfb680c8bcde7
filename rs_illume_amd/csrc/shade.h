// Device-side path-tracing pieces: the counter-based sampler, surface
// interactions of triangle and sphere hits, the three BSDFs (Lambertian,
// Trowbridge-Reitz conductor, Fresnel-specular dielectric), one-sided
// triangle area lights and the power-heuristic MIS of pbrt-v3.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "geom.h"
#include "lowdisc.h"
#include "layout.h"
#include "scene_types.h"

#pragma clang fp contract(off)

namespace illume {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInv2Pi = 0.15915494309189533577f;  // 1 / (2 pi) (pbrt.rs:35 has 2 / pi)
constexpr float kOneMinusEpsilon = 0x1.fffffep-1f;

// ---------------------------------------------------------------- sampler
// The sampler of one camera path. Random (lds == null): PCG32
// (core/rng.rs:9-63), one stream per (pixel, sample index) selected by a
// splitmix64 hash of (seed, pixel, sample); dimensions are consumed
// sequentially in pbrt's RandomSampler order. Low-discrepancy (lds = the
// Halton or Sobol' table): state is the path's global index, inc the next
// dimension (lowdisc.h).
__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
struct PathSampler {
    uint64_t state, inc;
    const uint32_t *lds;
    __device__ __forceinline__ uint32_t next() {
        uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    __device__ __forceinline__ void seed(uint64_t seed, uint64_t pixel, uint64_t sample) {
        lds = nullptr;
        uint64_t seq = splitmix64(seed ^ splitmix64((pixel << 24) | (sample & 0xffffff)));
        state = 0u;
        inc = (seq << 1u) | 1u;
        next();
        state += 0x853c49e6748fea9bULL;
        next();
    }
    __device__ __forceinline__ float get_1d() {
        if (lds) return lowdisc_dim(lds, state, (uint32_t)inc++);
        return fminf(kOneMinusEpsilon, (float)next() * 0x1p-32f);
    }
};

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ f3 cmul(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ bool is_black(f3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
__device__ __forceinline__ float luminance(f3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }
__device__ __forceinline__ float max_comp(f3 c) { return fmaxf(c.x, fmaxf(c.y, c.z)); }
__device__ __forceinline__ float abs_dot(f3 a, f3 b) { return fabsf(dot(a, b)); }
__device__ __forceinline__ f3 faceforward(f3 n, f3 v) { return (dot(n, v) < 0.f) ? neg(n) : n; }
__device__ __forceinline__ f3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }

// core/geometry/ray.rs:79-102
__device__ __forceinline__ f3 offset_ray_origin(f3 p, f3 p_error, f3 n, f3 w) {
    float d = dot(fabs3(n), p_error);
    f3 offset = smul(d, n);
    if (dot(w, n) < 0) offset = neg(offset);
    f3 po = add(p, offset);
    po.x = offset.x > 0 ? next_float_up(po.x) : (offset.x < 0 ? next_float_down(po.x) : po.x);
    po.y = offset.y > 0 ? next_float_up(po.y) : (offset.y < 0 ? next_float_down(po.y) : po.y);
    po.z = offset.z > 0 ? next_float_up(po.z) : (offset.z < 0 ? next_float_down(po.z) : po.z);
    return po;
}

// The shading-relevant part of SurfaceInteraction (interaction.rs:233-316).
struct Isect {
    f3 p, p_error, n, wo, shading_n, dpdu;
    uint32_t prim;
};

// Triangle hit (triangle.rs:258-345, default uvs): p from the barycentrics,
// p_error = gamma(7) * sum |b_i p_i|, n = normalize(dp02 x dp12),
// dpdu = p1 - p0 (rounded as (-1)*dp02 - (-1)*dp12).
__device__ __forceinline__ void triangle_isect(f3 p0, f3 p1, f3 p2, float b0, float b1, float b2, f3 d, Isect &it) {
    f3 dp02 = sub(p0, p2), dp12 = sub(p1, p2);
    f3 dpdu = mk((-1.f * dp02.x - (-1.f) * dp12.x) * 1.f, (-1.f * dp02.y - (-1.f) * dp12.y) * 1.f,
                 (-1.f * dp02.z - (-1.f) * dp12.z) * 1.f);
    float xs = fabsf(b0 * p0.x) + fabsf(b1 * p1.x) + fabsf(b2 * p2.x);
    float ys = fabsf(b0 * p0.y) + fabsf(b1 * p1.y) + fabsf(b2 * p2.y);
    float zs = fabsf(b0 * p0.z) + fabsf(b1 * p1.z) + fabsf(b2 * p2.z);
    it.p_error = smul(gamma_f(7), mk(xs, ys, zs));
    it.p = add(add(smul(b0, p0), smul(b1, p1)), smul(b2, p2));
    it.wo = neg(d);
    it.n = normalize(cross(dp02, dp12));
    it.shading_n = it.n;
    it.dpdu = dpdu;
}

// Sphere hit (sphere.rs:105-236 for a full sphere at center c): the object
// space ray is rebuilt exactly as in sphere_test, the hit point refined onto
// the sphere, and the object-to-world translation applied with its error
// (transform.rs:253-294, 449-479).
__device__ __forceinline__ void sphere_isect(f3 c, float radius, f3 ro, f3 rd, float t, Isect &it) {
    f3 o = mk(ro.x + (-c.x), ro.y + (-c.y), ro.z + (-c.z));
    f3 o_err = mul(mk(fabsf(ro.x) + fabsf(c.x), fabsf(ro.y) + fabsf(c.y), fabsf(ro.z) + fabsf(c.z)), gamma_f(3));
    float ls = length_squared(rd);
    if (ls > 0) {
        float dt = dot(fabs3(rd), o_err) / ls;
        o = add(o, mul(rd, dt));
    }
    f3 ph = add(o, mul(rd, t));
    float dist = length(sub(ph, mk(0.f, 0.f, 0.f)));
    ph = mul(ph, radius / dist);
    if (ph.x == 0 && ph.y == 0) ph.x = 1e-5f * radius;
    const float phi_max = (kPi / 180.f) * 360.f;
    const float theta_min = acosf(-1.f), theta_max = acosf(1.f);
    float theta = acosf(fminf(fmaxf(ph.z / radius, -1.f), 1.f));
    float z_radius = sqrtf(ph.x * ph.x + ph.y * ph.y);
    float inv_z_radius = 1.f / z_radius;
    float cos_phi = ph.x * inv_z_radius, sin_phi = ph.y * inv_z_radius;
    f3 dpdu = mk(-phi_max * ph.y, phi_max * ph.x, 0.f);
    f3 dpdv = smul(theta_max - theta_min, mk(ph.z * cos_phi, ph.z * sin_phi, -radius * sinf(theta)));
    f3 pe = smul(gamma_f(5), fabs3(ph));
    f3 n_obj = normalize(cross(dpdu, dpdv));
    // translation: p' = p + c, error (g3+1)*pe + g3*(|p| + |c|)
    const float g3 = gamma_f(3);
    it.p = mk((1.f * ph.x + 0.f * ph.y) + (0.f * ph.z + c.x), (0.f * ph.x + 1.f * ph.y) + (0.f * ph.z + c.y),
              (0.f * ph.x + 0.f * ph.y) + (1.f * ph.z + c.z));
    it.p_error = mk((g3 + 1.f) * (1.f * pe.x + 0.f * pe.y + 0.f * pe.z) +
                        g3 * (fabsf(1.f * ph.x) + fabsf(0.f * ph.y) + fabsf(0.f * ph.z) + fabsf(c.x)),
                    (g3 + 1.f) * (0.f * pe.x + 1.f * pe.y + 0.f * pe.z) +
                        g3 * (fabsf(0.f * ph.x) + fabsf(1.f * ph.y) + fabsf(0.f * ph.z) + fabsf(c.y)),
                    (g3 + 1.f) * (0.f * pe.x + 0.f * pe.y + 1.f * pe.z) +
                        g3 * (fabsf(0.f * ph.x) + fabsf(0.f * ph.y) + fabsf(1.f * ph.z) + fabsf(c.z)));
    it.n = normalize(n_obj);
    it.wo = normalize(neg(rd));
    it.shading_n = faceforward(normalize(n_obj), it.n);
    it.dpdu = dpdu;
}

// interaction.rs:33-43
__device__ __forceinline__ void spawn_ray(const Isect &it, f3 d, f3 &o) { o = offset_ray_origin(it.p, it.p_error, it.n, d); }
// interaction.rs:60-72: segment to another interaction, t_max = 1 - ShadowEpsilon
__device__ __forceinline__ void spawn_ray_to(const Isect &it, f3 ip, f3 ip_err, f3 in, f3 &o, f3 &d) {
    o = offset_ray_origin(it.p, it.p_error, it.n, sub(ip, it.p));
    f3 target = offset_ray_origin(ip, ip_err, in, sub(o, ip));
    d = sub(target, o);
}

// ---------------------------------------------------------------- sampling warps
// sampling.rs:210-230, 316-320
__device__ __forceinline__ f3 cosine_sample_hemisphere(float u0, float u1) {
    float ox = u0 * 2.0f - 1.0f, oy = u1 * 2.0f - 1.0f;
    float dx = 0.f, dy = 0.f;
    if (!(ox == 0 && oy == 0)) {
        float theta, r;
        if (fabsf(ox) > fabsf(oy)) {
            r = ox;
            theta = (kPi / 4) * (oy / ox);
        } else {
            r = oy;
            theta = (kPi / 2) - (kPi / 4) * (ox / oy);
        }
        dx = r * cosf(theta);
        dy = r * sinf(theta);
    }
    float z = sqrtf(fmaxf(0.f, 1 - dx * dx - dy * dy));
    return mk(dx, dy, z);
}

// sampling.rs:168-181 with y = r sin(phi) (the reference writes r cos(phi));
// pdf 1 / (2 pi)
__device__ __forceinline__ f3 uniform_sample_hemisphere(float u0, float u1) {
    float z = u0;
    float r = sqrtf(fmaxf(0.f, 1 - z * z));
    float phi = 2 * kPi * u1;
    return mk(r * cosf(phi), r * sinf(phi), z);
}
// sampling.rs:333-337
__device__ __forceinline__ float power_heuristic(float f_pdf, float g_pdf) {
    float f = 1 * f_pdf, g = 1 * g_pdf;
    return (f * f) / (f * f + g * g);
}

// ---------------------------------------------------------------- BSDFs
__device__ __forceinline__ float cos2_theta(f3 w) { return w.z * w.z; }
__device__ __forceinline__ float abs_cos_theta(f3 w) { return fabsf(w.z); }
__device__ __forceinline__ float sin2_theta(f3 w) { return fmaxf(0.f, 1 - cos2_theta(w)); }
__device__ __forceinline__ float sin_theta(f3 w) { return sqrtf(sin2_theta(w)); }
__device__ __forceinline__ float tan_theta(f3 w) { return sin_theta(w) / w.z; }
__device__ __forceinline__ float tan2_theta(f3 w) { return sin2_theta(w) / cos2_theta(w); }
__device__ __forceinline__ float cos_phi(f3 w) {
    float st = sin_theta(w);
    return (st == 0) ? 1.f : fminf(fmaxf(w.x / st, -1.f), 1.f);
}
__device__ __forceinline__ float sin_phi(f3 w) {
    float st = sin_theta(w);
    return (st == 0) ? 0.f : fminf(fmaxf(w.y / st, -1.f), 1.f);
}
__device__ __forceinline__ bool same_hemisphere(f3 w, f3 wp) { return w.z * wp.z > 0; }

// reflection/mod.rs:420-445
__device__ __forceinline__ float fr_dielectric(float cos_i, float eta_i, float eta_t) {
    cos_i = fminf(fmaxf(cos_i, -1.f), 1.f);
    if (!(cos_i > 0.f)) {
        float tmp = eta_i;
        eta_i = eta_t;
        eta_t = tmp;
        cos_i = fabsf(cos_i);
    }
    float sin_i = sqrtf(fmaxf(0.f, 1 - cos_i * cos_i));
    float sin_t = eta_i / eta_t * sin_i;
    if (sin_t >= 1) return 1.f;
    float cos_t = sqrtf(fmaxf(0.f, 1 - sin_t * sin_t));
    float r_parl = ((eta_t * cos_i) - (eta_i * cos_t)) / ((eta_t * cos_i) + (eta_i * cos_t));
    float r_perp = ((eta_i * cos_i) - (eta_t * cos_t)) / ((eta_i * cos_i) + (eta_t * cos_t));
    return (r_parl * r_parl + r_perp * r_perp) / 2;
}
// reflection/mod.rs:448-475 (eta_i = 1)
__device__ __forceinline__ float fr_conductor1(float cos_i, float eta_t, float k) {
    cos_i = fminf(fmaxf(cos_i, -1.f), 1.f);
    float eta = eta_t / 1.f, etak = k / 1.f;
    float cos2 = cos_i * cos_i, sin2 = 1.f - cos2;
    float eta2 = eta * eta, etak2 = etak * etak;
    float t0 = eta2 - etak2 - sin2;
    float a2plusb2 = sqrtf(t0 * t0 + 4 * eta2 * etak2);
    float t1 = a2plusb2 + cos2;
    float a = sqrtf(0.5f * (a2plusb2 + t0));
    float t2 = 2.f * cos_i * a;
    float rs = (t1 - t2) / (t1 + t2);
    float t3 = cos2 * a2plusb2 + sin2 * sin2;
    float t4 = t2 * sin2;
    float rp = rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (rp + rs);
}

// Isotropic Trowbridge-Reitz (trowbridge_reitz_distribution.rs, pbrt-v3
// semantics: |tan| in Lambda, visible-normal sampling).
__device__ __forceinline__ float tr_D(float a, f3 wh) {
    float t2 = tan2_theta(wh);
    if (__builtin_isinf(t2)) return 0.f;
    float cos4 = cos2_theta(wh) * cos2_theta(wh);
    float cp = cos_phi(wh), sp = sin_phi(wh);
    float e = (cp * cp / (a * a) + sp * sp / (a * a)) * t2;
    return 1 / (kPi * a * a * cos4 * (1 + e) * (1 + e));
}
__device__ __forceinline__ float tr_lambda(float a, f3 w) {
    float abs_tan = fabsf(tan_theta(w));
    if (__builtin_isinf(abs_tan)) return 0.f;
    float cp = cos_phi(w), sp = sin_phi(w);
    float alpha = sqrtf(cp * cp * a * a + sp * sp * a * a);
    float a2t2 = (alpha * abs_tan) * (alpha * abs_tan);
    return (-1 + sqrtf(1.f + a2t2)) / 2;
}
__device__ __forceinline__ float tr_pdf(float a, f3 wo, f3 wh) {
    return tr_D(a, wh) * (1 / (1 + tr_lambda(a, wo))) * fabsf(dot(wo, wh)) / abs_cos_theta(wo);
}
// Visible-normal sample of the distribution, exactly (Heitz 2018, "Sampling
// the GGX Distribution of Visible Normals", JCGT 7(4)): wo stretched to the
// unit-roughness configuration, a disk sample warped towards the visible half
// of the projected hemisphere, lifted onto it and unstretched. (The
// reference's sampler, trowbridge_reitz_distribution.rs:111-182, inverts the
// y slope with a rational fit whose tails miss the distribution the pdf
// describes.) The same operations in the same order as the oracle.
__device__ __forceinline__ f3 tr_sample_wh(float a, f3 wo, float u1, float u2) {
    bool flip = wo.z < 0;
    f3 w = flip ? neg(wo) : wo;
    f3 vh = normalize(mk(a * w.x, a * w.y, w.z));
    float lensq = vh.x * vh.x + vh.y * vh.y;
    f3 t1 = lensq > 0 ? mul(mk(-vh.y, vh.x, 0.f), 1 / sqrtf(lensq)) : mk(1.f, 0.f, 0.f);
    f3 t2 = cross(vh, t1);
    float r = sqrtf(u1);
    float phi = 2 * kPi * u2;
    float p1 = r * cosf(phi), p2 = r * sinf(phi);
    float s = 0.5f * (1 + vh.z);
    p2 = (1 - s) * sqrtf(fmaxf(0.f, 1 - p1 * p1)) + s * p2;
    f3 nh = add(add(smul(p1, t1), smul(p2, t2)), smul(sqrtf(fmaxf(0.f, 1 - p1 * p1 - p2 * p2)), vh));
    f3 wh = normalize(mk(a * nh.x, a * nh.y, fmaxf(0.f, nh.z)));
    return flip ? neg(wh) : wh;
}

enum { kRefl = 1, kTrans = 2, kDiffuse = 4, kGlossy = 8, kSpecular = 16, kAll = 31 };

// reflection/mod.rs:50-275 with one lobe per material.
struct Bsdf {
    f3 ng, ns, ss, ts;
    DevMaterial m;
    float eta;
    int type;
    __device__ __forceinline__ void init(const Isect &it, const DevMaterial &mat) {
        m = mat;
        ns = it.shading_n;
        ng = it.n;
        ss = normalize(it.dpdu);
        ts = cross(ns, ss);
        eta = 1.f;
        if (mat.type == kMatte)
            type = (mat.kd[0] == 0.f && mat.kd[1] == 0.f && mat.kd[2] == 0.f) ? 0 : (kRefl | kDiffuse);
        else if (mat.type == kMetal)
            type = kRefl | kGlossy;
        else {
            type = kRefl | kTrans | kSpecular;
            eta = mat.ior;
        }
    }
    __device__ __forceinline__ f3 w2l(f3 v) const { return mk(dot(v, ss), dot(v, ts), dot(v, ns)); }
    __device__ __forceinline__ f3 l2w(f3 v) const {
        return mk(ss.x * v.x + ts.x * v.y + ns.x * v.z, ss.y * v.x + ts.y * v.y + ns.y * v.z,
                  ss.z * v.x + ts.z * v.y + ns.z * v.z);
    }
    __device__ __forceinline__ bool matches(int flags) const { return type != 0 && (type & flags) == type; }
    __device__ __forceinline__ f3 bxdf_f(f3 wo, f3 wi) const {
        if (m.type == kMatte) return mk(m.kd[0] * kInvPi, m.kd[1] * kInvPi, m.kd[2] * kInvPi);
        if (m.type == kMetal) {
            float cto = abs_cos_theta(wo), cti = abs_cos_theta(wi);
            f3 wh = add(wi, wo);
            if (cti == 0 || cto == 0) return mk(0.f, 0.f, 0.f);
            if (wh.x == 0 && wh.y == 0 && wh.z == 0) return mk(0.f, 0.f, 0.f);
            wh = normalize(wh);
            float c = fabsf(dot(wi, faceforward(wh, mk(0.f, 0.f, 1.f))));
            f3 F = mk(fr_conductor1(c, m.eta[0], m.k[0]), fr_conductor1(c, m.eta[1], m.k[1]),
                      fr_conductor1(c, m.eta[2], m.k[2]));
            float dg = 1.f * tr_D(m.alpha, wh) * (1 / (1 + tr_lambda(m.alpha, wo) + tr_lambda(m.alpha, wi)));
            float den = 4 * cti * cto;
            return mk(dg * F.x / den, dg * F.y / den, dg * F.z / den);
        }
        return mk(0.f, 0.f, 0.f);
    }
    __device__ __forceinline__ float bxdf_pdf(f3 wo, f3 wi) const {
        if (m.type == kMatte) return same_hemisphere(wo, wi) ? abs_cos_theta(wi) * kInvPi : 0.f;
        if (m.type == kMetal) {
            if (!same_hemisphere(wo, wi)) return 0.f;
            f3 wh = normalize(add(wo, wi));
            return tr_pdf(m.alpha, wo, wh) / (4 * dot(wo, wh));
        }
        return 0.f;
    }
    __device__ __forceinline__ f3 bxdf_sample_f(f3 wo, f3 &wi, float u0, float u1, float &pdf, int &sampled) const {
        if (m.type == kMatte) {
            wi = cosine_sample_hemisphere(u0, u1);
            if (wo.z < 0) wi.z *= -1;
            pdf = bxdf_pdf(wo, wi);
            return bxdf_f(wo, wi);
        }
        if (m.type == kMetal) {
            if (wo.z == 0) return mk(0.f, 0.f, 0.f);
            f3 wh = tr_sample_wh(m.alpha, wo, u0, u1);
            if (dot(wo, wh) < 0) return mk(0.f, 0.f, 0.f);
            wi = add(neg(wo), smul(2 * dot(wo, wh), wh));
            if (!same_hemisphere(wo, wi)) return mk(0.f, 0.f, 0.f);
            pdf = tr_pdf(m.alpha, wo, wh) / (4 * dot(wo, wh));
            return bxdf_f(wo, wi);
        }
        float F = fr_dielectric(wo.z, 1.f, m.ior);
        if (u0 < F) {
            wi = mk(-wo.x, -wo.y, wo.z);
            sampled = kSpecular | kRefl;
            pdf = F;
            float v = F / abs_cos_theta(wi);
            return mk(v, v, v);
        }
        bool entering = wo.z > 0;
        float eta_i = entering ? 1.f : m.ior;
        float eta_t = entering ? m.ior : 1.f;
        f3 n = faceforward(mk(0.f, 0.f, 1.f), wo);
        float eta = eta_i / eta_t;
        float cos_i = dot(n, wo);
        float sin2_i = fmaxf(0.f, 1.f - cos_i * cos_i);
        float sin2_t = eta * eta * sin2_i;
        if (sin2_t >= 1) return mk(0.f, 0.f, 0.f);
        float cos_t = sqrtf(1 - sin2_t);
        wi = add(smul(eta, neg(wo)), smul(eta * cos_i - cos_t, n));
        float ft = 1 * (1 - F);
        ft *= (eta_i * eta_i) / (eta_t * eta_t);
        sampled = kSpecular | kTrans;
        pdf = 1 - F;
        float v = ft / abs_cos_theta(wi);
        return mk(v, v, v);
    }
    __device__ __forceinline__ f3 f(f3 wo_w, f3 wi_w, int flags) const {
        f3 wi = w2l(wi_w), wo = w2l(wo_w);
        if (wo.z == 0) return mk(0.f, 0.f, 0.f);
        bool refl = dot(wi_w, ng) * dot(wo_w, ng) > 0;
        if (matches(flags) && ((refl && (type & kRefl)) || (!refl && (type & kTrans)))) return bxdf_f(wo, wi);
        return mk(0.f, 0.f, 0.f);
    }
    __device__ __forceinline__ f3 sample_f(f3 wo_w, f3 &wi_w, float u0, float u1, float &pdf, int flags,
                                           int &sampled) const {
        int matching = matches(flags) ? 1 : 0;
        if (matching == 0) {
            pdf = 0;
            sampled = 0;
            return mk(0.f, 0.f, 0.f);
        }
        int comp = min((int)floorf(u0 * matching), matching - 1);
        float ur0 = fminf(u0 * matching - comp, kOneMinusEpsilon);
        f3 wi = mk(0.f, 0.f, 0.f), wo = w2l(wo_w);
        if (wo.z == 0) return mk(0.f, 0.f, 0.f);
        pdf = 0;
        sampled = type;
        f3 fv = bxdf_sample_f(wo, wi, ur0, u1, pdf, sampled);
        if (pdf == 0) {
            sampled = 0;
            return mk(0.f, 0.f, 0.f);
        }
        wi_w = l2w(wi);
        if (!(type & kSpecular)) {
            bool refl = dot(wi_w, ng) * dot(wo_w, ng) > 0;
            fv = mk(0.f, 0.f, 0.f);
            if ((refl && (type & kRefl)) || (!refl && (type & kTrans))) fv = bxdf_f(wo, wi);
        }
        return fv;
    }
    __device__ __forceinline__ float pdf(f3 wo_w, f3 wi_w, int flags) const {
        if (type == 0) return 0.f;
        f3 wo = w2l(wo_w), wi = w2l(wi_w);
        if (wo.z == 0) return 0.f;
        if (!matches(flags)) return 0.f;
        return bxdf_pdf(wo, wi) / 1;
    }
};

}  // namespace illume
