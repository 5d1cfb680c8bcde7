// Device geometry: pbrt-v3 numerics, the watertight ray-triangle test, the
// interval-arithmetic ray-sphere test and the two-box slab test.
//
// Every expression is evaluated in the order the reference writes it, with
// floating-point contraction off (this file and the whole library are built
// with -ffp-contract=off), IEEE division and square root, and denormals kept,
// so hit decisions and hit distances match the reference's FP32 arithmetic
// bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace illume {

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 mul(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 smul(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 fabs3(f3 a) { return f3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// vector3.rs:66-82: cross product in f64, rounded to f32.
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
    return f3{(float)((ay * bz) - (az * by)), (float)((az * bx) - (ax * bz)), (float)((ax * by) - (ay * bx))};
}
__device__ __forceinline__ float length_squared(f3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
__device__ __forceinline__ float length(f3 v) { return sqrtf(length_squared(v)); }
__device__ __forceinline__ f3 normalize(f3 v) {
    float l = length(v);
    return f3{v.x / l, v.y / l, v.z / l};
}
__device__ __forceinline__ float max3f(float a, float b, float c) { return fmaxf(a, fmaxf(b, c)); }

// core/pbrt.rs:172-174, evaluated in f32 exactly as the reference does.
__host__ __device__ constexpr float gamma_f(int n) {
    return ((float)n * (5.9604644775390625e-08f)) / (1.0f - (float)n * (5.9604644775390625e-08f));
}

// core/pbrt.rs:129-169
__device__ __forceinline__ float next_float_up(float v) {
    if (__builtin_isinf(v) && v > 0.f) return v;
    if (v == -0.f) v = 0.f;
    uint32_t ui = __float_as_uint(v);
    if (v >= 0) ++ui; else --ui;
    return __uint_as_float(ui);
}
__device__ __forceinline__ float next_float_down(float v) {
    if (__builtin_isinf(v) && v < 0.f) return v;
    if (v == 0.f) v = -0.f;
    uint32_t ui = __float_as_uint(v);
    if (v > 0) --ui; else ++ui;
    return __uint_as_float(ui);
}

// The per-ray part of the watertight test (triangle.rs:156-176): the axis
// permutation that makes the largest |d| component z, and the shear that
// maps the ray onto +z.
struct Shear {
    int k;  // kx | ky << 2 | kz << 4
    float sx, sy, sz;
};
__device__ __forceinline__ Shear make_shear(f3 d) {
    f3 ad = fabs3(d);
    int kz = (ad.x > ad.y) ? ((ad.x > ad.z) ? 0 : 2) : ((ad.y > ad.z) ? 1 : 2);
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    float dx = comp(d, kx), dy = comp(d, ky), dz = comp(d, kz);
    Shear s;
    s.k = kx | (ky << 2) | (kz << 4);
    s.sx = -dx / dz;
    s.sy = -dy / dz;
    s.sz = 1.f / dz;
    return s;
}

// Watertight ray-triangle test, shapes/triangle.rs:142-256 (pbrt-v3; the
// delta_t bound uses the true max component, divergence 6 of the reference),
// with the ray's shear computed once by the caller. On a hit returns true
// with t and barycentrics b0..b2. Apart from the accept/reject comparisons
// the results do not depend on t_max.
__device__ __forceinline__ bool triangle_test_sheared(const Shear &sh, f3 p0, f3 p1, f3 p2, f3 o, float t_max,
                                                      float &t_out, float &b0_out, float &b1_out, float &b2_out) {
    f3 p0t = sub(p0, o), p1t = sub(p1, o), p2t = sub(p2, o);
    const int kx = sh.k & 3, ky = (sh.k >> 2) & 3, kz = sh.k >> 4;
    f3 q0 = mk(comp(p0t, kx), comp(p0t, ky), comp(p0t, kz));
    f3 q1 = mk(comp(p1t, kx), comp(p1t, ky), comp(p1t, kz));
    f3 q2 = mk(comp(p2t, kx), comp(p2t, ky), comp(p2t, kz));
    const float sx = sh.sx, sy = sh.sy, sz = sh.sz;
    q0.x += sx * q0.z;
    q0.y += sy * q0.z;
    q1.x += sx * q1.z;
    q1.y += sy * q1.z;
    q2.x += sx * q2.z;
    q2.y += sy * q2.z;
    float e0 = q1.x * q2.y - q1.y * q2.x;
    float e1 = q2.x * q0.y - q2.y * q0.x;
    float e2 = q0.x * q1.y - q0.y * q1.x;
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {
        double p2txp1ty = (double)q2.x * (double)q1.y;
        double p2typ1tx = (double)q2.y * (double)q1.x;
        e0 = (float)(p2typ1tx - p2txp1ty);
        double p0txp2ty = (double)q0.x * (double)q2.y;
        double p0typ2tx = (double)q0.y * (double)q2.x;
        e1 = (float)(p0typ2tx - p0txp2ty);
        double p1txp0ty = (double)q1.x * (double)q0.y;
        double p1typ0tx = (double)q1.y * (double)q0.x;
        e2 = (float)(p1typ0tx - p1txp0ty);
    }
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return false;
    float det = e0 + e1 + e2;
    if (det == 0) return false;
    q0.z *= sz;
    q1.z *= sz;
    q2.z *= sz;
    float t_scaled = e0 * q0.z + e1 * q1.z + e2 * q2.z;
    if (det < 0 && (t_scaled >= 0 || t_scaled < t_max * det)) return false;
    if (det > 0 && (t_scaled <= 0 || t_scaled > t_max * det)) return false;
    float inv_det = 1.f / det;
    float b0 = e0 * inv_det;
    float b1 = e1 * inv_det;
    float b2 = e2 * inv_det;
    float t = t_scaled * inv_det;
    float max_zt = max3f(fabsf(q0.z), fabsf(q1.z), fabsf(q2.z));
    float delta_z = gamma_f(3) * max_zt;
    float max_xt = max3f(fabsf(q0.x), fabsf(q1.x), fabsf(q2.x));
    float max_yt = max3f(fabsf(q0.y), fabsf(q1.y), fabsf(q2.y));
    float delta_x = gamma_f(5) * (max_xt + max_zt);
    float delta_y = gamma_f(5) * (max_yt + max_zt);
    float delta_e = 2.f * (gamma_f(2) * max_xt * max_yt + delta_y * max_xt + delta_x * max_yt);
    float max_e = max3f(fabsf(e0), fabsf(e1), fabsf(e2));
    float delta_t = 3.f * (gamma_f(3) * max_e * max_zt + delta_e * max_zt + delta_z * max_e) * fabsf(inv_det);
    if (t <= delta_t) return false;
    t_out = t;
    b0_out = b0;
    b1_out = b1;
    b2_out = b2;
    return true;
}

__device__ __forceinline__ bool triangle_test(f3 p0, f3 p1, f3 p2, f3 o, f3 d, float t_max, float &t_out,
                                              float &b0_out, float &b1_out, float &b2_out) {
    return triangle_test_sheared(make_shear(d), p0, p1, p2, o, t_max, t_out, b0_out, b1_out, b2_out);
}

// triangle.rs:258-296 for the default uv parameterisation: a closest hit on a
// triangle whose partial-derivative cross product AND geometric normal both
// vanish in f32 is rejected. Only evaluated for accepted candidates.
__device__ __forceinline__ bool triangle_degenerate(f3 p0, f3 p1, f3 p2) {
    f3 dp02 = sub(p0, p2), dp12 = sub(p1, p2);
    // dpdu = (-1*dp02 - (-1)*dp12) * 1, dpdv = (-0*dp02 + (-1)*dp12) * 1
    f3 dpdu = mk((-1.f * dp02.x - (-1.f) * dp12.x) * 1.f, (-1.f * dp02.y - (-1.f) * dp12.y) * 1.f,
                 (-1.f * dp02.z - (-1.f) * dp12.z) * 1.f);
    f3 dpdv = mk((-0.f * dp02.x + (-1.f) * dp12.x) * 1.f, (-0.f * dp02.y + (-1.f) * dp12.y) * 1.f,
                 (-0.f * dp02.z + (-1.f) * dp12.z) * 1.f);
    if (length_squared(cross(dpdu, dpdv)) != 0.f) return false;
    f3 ng = cross(sub(p2, p0), sub(p1, p0));
    return length_squared(ng) == 0.f;
}

// ---------------------------------------------------------------- sphere
// core/efloat.rs with pbrt-v3 interval semantics (exact scalars, low-high
// subtraction).
struct EF {
    float v, lo, hi;
};
__device__ __forceinline__ EF ef(float v, float err) {
    EF r;
    r.v = v;
    if (err == 0.f) {
        r.lo = v;
        r.hi = v;
    } else {
        r.lo = next_float_down(v - err);
        r.hi = next_float_up(v + err);
    }
    return r;
}
__device__ __forceinline__ EF ef_exact(float v) { return EF{v, v, v}; }
__device__ __forceinline__ EF ef_add(EF a, EF b) {
    return EF{a.v + b.v, next_float_down(a.lo + b.lo), next_float_up(a.hi + b.hi)};
}
__device__ __forceinline__ EF ef_sub(EF a, EF b) {
    return EF{a.v - b.v, next_float_down(a.lo - b.hi), next_float_up(a.hi - b.lo)};
}
__device__ __forceinline__ EF ef_mul(EF a, EF b) {
    float p0 = a.lo * b.lo, p1 = a.hi * b.lo, p2 = a.lo * b.hi, p3 = a.hi * b.hi;
    return EF{a.v * b.v, next_float_down(fminf(fminf(p0, p1), fminf(p2, p3))),
              next_float_up(fmaxf(fmaxf(p0, p1), fmaxf(p2, p3)))};
}
__device__ __forceinline__ EF ef_div(EF a, EF b) {
    EF r;
    r.v = a.v / b.v;
    if (b.lo < 0 && b.hi > 0) {
        r.lo = -__builtin_inff();
        r.hi = __builtin_inff();
        return r;
    }
    float d0 = a.lo / b.lo, d1 = a.hi / b.lo, d2 = a.lo / b.hi, d3 = a.hi / b.hi;
    r.lo = next_float_down(fminf(fminf(d0, d1), fminf(d2, d3)));
    r.hi = next_float_up(fmaxf(fmaxf(d0, d1), fmaxf(d2, d3)));
    return r;
}

// Ray-sphere test for a full sphere (shapes/sphere.rs:105-204 with z_min=-r,
// z_max=r, phi_max=2pi): world-to-object is the translation by -center
// (transform.rs:225-310, 392-434 evaluated for an identity linear part), the
// EFloat quadratic (efloat.rs:226-257), then the conservative root choice.
// Returns true with the object-space hit distance of the chosen root.
__device__ __forceinline__ bool sphere_test(f3 c, float radius, f3 ro, f3 rd, float t_max, float &t_out) {
    // transform_point_with_error: x' = x + (-cx); error gamma(3)*(|x| + |cx|)
    f3 o = mk(ro.x + (-c.x), ro.y + (-c.y), ro.z + (-c.z));
    f3 o_err = mul(mk(fabsf(ro.x) + fabsf(c.x), fabsf(ro.y) + fabsf(c.y), fabsf(ro.z) + fabsf(c.z)), gamma_f(3));
    f3 d = rd;
    f3 d_err = mk(gamma_f(3) * fabsf(d.x), gamma_f(3) * fabsf(d.y), gamma_f(3) * fabsf(d.z));
    float ls = length_squared(d);
    if (ls > 0) {
        float dt = dot(fabs3(d), o_err) / ls;
        o = add(o, mul(d, dt));
    }
    EF ox = ef(o.x, o_err.x), oy = ef(o.y, o_err.y), oz = ef(o.z, o_err.z);
    EF dx = ef(d.x, d_err.x), dy = ef(d.y, d_err.y), dz = ef(d.z, d_err.z);
    EF a = ef_add(ef_add(ef_mul(dx, dx), ef_mul(dy, dy)), ef_mul(dz, dz));
    EF b = ef_mul(ef_add(ef_add(ef_mul(dx, ox), ef_mul(dy, oy)), ef_mul(dz, oz)), ef_exact(2.f));
    EF cc = ef_sub(ef_add(ef_add(ef_mul(ox, ox), ef_mul(oy, oy)), ef_mul(oz, oz)),
                   ef_mul(ef_exact(radius), ef_exact(radius)));
    double discrim = (double)b.v * (double)b.v - 4. * (double)a.v * (double)cc.v;
    if (discrim < 0.) return false;
    double root_discrim = sqrt(discrim);
    float rdsc = (float)root_discrim;
    EF frd = ef(rdsc, 5.9604644775390625e-08f * rdsc);
    EF q = (b.v < 0) ? ef_mul(ef_sub(b, frd), ef_exact(-0.5f)) : ef_mul(ef_add(b, frd), ef_exact(-0.5f));
    EF t0 = ef_div(q, a), t1 = ef_div(cc, q);
    if (t0.v > t1.v) {
        EF tmp = t0;
        t0 = t1;
        t1 = tmp;
    }
    if (t0.hi > t_max || t1.lo <= 0) return false;
    EF tsh = t0;
    if (tsh.lo <= 0) {
        tsh = t1;
        if (tsh.hi > t_max) return false;
    }
    // Full sphere: the z clips never apply and phi <= 2pi always holds in f32,
    // so the clipping branch of sphere.rs:166-195 cannot reject.
    t_out = tsh.v;
    return true;
}

// Two-child slab test. Per axis the entry/exit distances are the min/max of
// (lo-o)*inv and (hi-o)*inv — the same products the reference forms after
// choosing planes by dir_is_neg (bounds3.rs:284-318) — with exits scaled by
// 1+2*gamma(3). fminf/fmaxf drop NaNs (0*inf), so a box the reference's test
// enters is always entered here: culling is at least as conservative.
// Rounding is monotone, so scaling the smallest exit once gives the same
// value as scaling each axis's exit before taking the minimum.
__device__ __forceinline__ bool slab(float lo_x, float hi_x, float lo_y, float hi_y, float lo_z, float hi_z, f3 o,
                                     f3 inv, float t_max, float &t_enter) {
    constexpr float k = 1.0f + 2.0f * gamma_f(3);
    float ax = (lo_x - o.x) * inv.x, bx = (hi_x - o.x) * inv.x;
    float ay = (lo_y - o.y) * inv.y, by = (hi_y - o.y) * inv.y;
    float az = (lo_z - o.z) * inv.z, bz = (hi_z - o.z) * inv.z;
    float t0 = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
    float t1 = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz)) * k;
    t_enter = t0;
    return t0 <= t1 && t0 < t_max && t1 > 0.f;
}

}  // namespace illume
