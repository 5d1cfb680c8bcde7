// ORACLE — test infrastructure only. This is the CPU restatement of rs_illume's
// ray-tracing hot path that the HIP kernels are checked against. Nothing in the
// product (the rs_illume_amd package, its HIP library, bench.py's timed leg)
// may include, link or call it; tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it only as the checker.
//
// Core numerics and geometry. Every function cites the reference file:line it
// restates. Where the reference diverges from pbrt-v3 (the book it ports and
// cites), the pbrt-v3 semantics are taken and the divergence is named.
// Compiled with -ffp-contract=off so each expression rounds exactly as written.
#pragma once
#include <cstdint>
#include <cstring>
#include <cmath>
#include <algorithm>
#include <limits>

namespace orc {

typedef float Float;

// core/pbrt.rs:23 MACHINE_EPSILON = f32::EPSILON * 0.5
static const Float MachineEpsilon = std::numeric_limits<float>::epsilon() * 0.5f;
// core/pbrt.rs:33
static const Float ShadowEpsilon = 0.0001f;
static const Float Pi = 3.14159265358979323846f;
static const Float InvPi = 0.31830988618379067154f;
// core/pbrt.rs:35-36 compute 1/pi*2 and 1/pi*4 (divergence); pbrt-v3 values:
static const Float Inv2Pi = 0.15915494309189533577f;
static const Float Inv4Pi = 0.07957747154594766788f;
static const Float Infinity = std::numeric_limits<float>::infinity();

// core/pbrt.rs:172-174
inline Float gamma(int n) { return ((Float)n * MachineEpsilon) / (1 - (Float)n * MachineEpsilon); }

inline uint32_t float_to_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bits_to_float(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// core/pbrt.rs:129-147
inline float next_float_up(float v) {
    if (std::isinf(v) && v > 0.f) return v;
    if (v == -0.f) v = 0.f;
    uint32_t ui = float_to_bits(v);
    if (v >= 0) ++ui; else --ui;
    return bits_to_float(ui);
}
// core/pbrt.rs:151-169
inline float next_float_down(float v) {
    if (std::isinf(v) && v < 0.f) return v;
    if (v == 0.f) v = -0.f;
    uint32_t ui = float_to_bits(v);
    if (v > 0) --ui; else ++ui;
    return bits_to_float(ui);
}

template <typename T> inline T clamp(T v, T lo, T hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---------------------------------------------------------------- vectors
// core/geometry/vector3.rs, point3.rs, normal3.rs: one type serves all three.
struct V3 {
    Float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(Float a, Float b, Float c) : x(a), y(b), z(c) {}
    Float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    Float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    V3 operator+(const V3 &o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3 &o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator*(Float s) const { return V3(x * s, y * s, z * s); }
    V3 operator/(Float s) const { return V3(x / s, y / s, z / s); }
    V3 operator-() const { return V3(-x, -y, -z); }
    V3 &operator+=(const V3 &o) { x += o.x; y += o.y; z += o.z; return *this; }
    V3 &operator*=(Float s) { x *= s; y *= s; z *= s; return *this; }
};
inline V3 operator*(Float s, const V3 &v) { return V3(s * v.x, s * v.y, s * v.z); }
inline V3 vabs(const V3 &v) { return V3(std::fabs(v.x), std::fabs(v.y), std::fabs(v.z)); }
inline Float dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Float abs_dot(const V3 &a, const V3 &b) { return std::fabs(dot(a, b)); }
// vector3.rs:66-82: the cross product is evaluated in f64 and rounded to f32.
inline V3 cross(const V3 &a, const V3 &b) {
    double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
    return V3((Float)((ay * bz) - (az * by)), (Float)((az * bx) - (ax * bz)), (Float)((ax * by) - (ay * bx)));
}
inline Float length_squared(const V3 &v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
inline Float length(const V3 &v) { return std::sqrt(length_squared(v)); }
// vector3.rs:85-95 multiplies by 1/len (divergence 7); pbrt-v3 divides.
inline V3 normalize(const V3 &v) { return v / length(v); }
// vector3.rs:98-110 max_component is wrong (divergence 6); pbrt-v3: true max.
inline Float max_component(const V3 &v) { return std::max(v.x, std::max(v.y, v.z)); }
// vector3.rs:140-149
inline int max_dimension(const V3 &v) { return (v.x > v.y) ? ((v.x > v.z) ? 0 : 2) : ((v.y > v.z) ? 1 : 2); }
// vector3.rs:152-160
inline V3 permute(const V3 &v, int x, int y, int z) { return V3(v[x], v[y], v[z]); }
inline Float distance(const V3 &a, const V3 &b) { return length(a - b); }
inline V3 faceforward(const V3 &n, const V3 &v) { return (dot(n, v) < 0.f) ? -n : n; }
// vector3.rs:520-537
inline void coordinate_system(const V3 &v1, V3 *v2, V3 *v3) {
    if (std::fabs(v1.x) > std::fabs(v1.y))
        *v2 = V3(-v1.z, 0, v1.x) / std::sqrt(v1.x * v1.x + v1.z * v1.z);
    else
        *v2 = V3(0, v1.z, -v1.y) / std::sqrt(v1.y * v1.y + v1.z * v1.z);
    *v3 = cross(v1, *v2);
}

// ---------------------------------------------------------------- ray
// core/geometry/ray.rs:9-31 (medium and differentials are not on this path)
struct Ray {
    V3 o, d;
    Float t_max;
    Float time;
    Ray() : t_max(Infinity), time(0) {}
    Ray(const V3 &o_, const V3 &d_, Float tm = Infinity, Float t = 0) : o(o_), d(d_), t_max(tm), time(t) {}
    V3 at(Float t) const { return o + d * t; }
};

// core/geometry/ray.rs:79-102
inline V3 offset_ray_origin(const V3 &p, const V3 &p_error, const V3 &n, const V3 &w) {
    Float d = dot(vabs(n), p_error);
    V3 offset = d * n;
    if (dot(w, n) < 0) offset = -offset;
    V3 po = p + offset;
    for (int i = 0; i < 3; ++i) {
        if (offset[i] > 0) po[i] = next_float_up(po[i]);
        else if (offset[i] < 0) po[i] = next_float_down(po[i]);
    }
    return po;
}

// ---------------------------------------------------------------- bounds
// core/geometry/bounds3.rs
struct Bounds3 {
    V3 pmin, pmax;
    Bounds3() : pmin(Infinity, Infinity, Infinity), pmax(-Infinity, -Infinity, -Infinity) {}
    Bounds3(const V3 &p) : pmin(p), pmax(p) {}
    Bounds3(const V3 &a, const V3 &b)
        : pmin(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)),
          pmax(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)) {}
    const V3 &operator[](int i) const { return i == 0 ? pmin : pmax; }
    V3 diagonal() const { return pmax - pmin; }
    // bounds3.rs:59-62
    Float surface_area() const {
        V3 d = diagonal();
        return 2 * (d.x * d.y + d.x * d.z + d.y * d.z);
    }
    // bounds3.rs:53-56
    int maximum_extent() const {
        V3 d = diagonal();
        if (d.x > d.y && d.x > d.z) return 0;
        else if (d.y > d.z) return 1;
        else return 2;
    }
    // bounds3.rs:178-190
    V3 offset(const V3 &p) const {
        V3 o = p - pmin;
        if (pmax.x > pmin.x) o.x /= pmax.x - pmin.x;
        if (pmax.y > pmin.y) o.y /= pmax.y - pmin.y;
        if (pmax.z > pmin.z) o.z /= pmax.z - pmin.z;
        return o;
    }
    // bounds3.rs:238-242
    V3 centroid() const { return pmin * 0.5f + pmax * 0.5f; }
    // bounds3.rs:284-318: slab test with the far distances scaled by 1+2*gamma(3).
    bool intersect_p(const Ray &ray, const V3 &inv_dir, const int dir_is_neg[3]) const {
        const Bounds3 &b = *this;
        Float t_min = (b[dir_is_neg[0]].x - ray.o.x) * inv_dir.x;
        Float t_max = (b[1 - dir_is_neg[0]].x - ray.o.x) * inv_dir.x;
        Float ty_min = (b[dir_is_neg[1]].y - ray.o.y) * inv_dir.y;
        Float ty_max = (b[1 - dir_is_neg[1]].y - ray.o.y) * inv_dir.y;
        t_max *= 1 + 2 * gamma(3);
        ty_max *= 1 + 2 * gamma(3);
        if (t_min > ty_max || ty_min > t_max) return false;
        if (ty_min > t_min) t_min = ty_min;
        if (ty_max < t_max) t_max = ty_max;
        Float tz_min = (b[dir_is_neg[2]].z - ray.o.z) * inv_dir.z;
        Float tz_max = (b[1 - dir_is_neg[2]].z - ray.o.z) * inv_dir.z;
        tz_max *= 1 + 2 * gamma(3);
        if (t_min > tz_max || tz_min > t_max) return false;
        if (tz_min > t_min) t_min = tz_min;
        if (tz_max < t_max) t_max = tz_max;
        return (t_min < ray.t_max) && (t_max > 0);
    }
};
// bounds3.rs:37-46
inline Bounds3 bunion(const Bounds3 &a, const Bounds3 &b) {
    Bounds3 r;
    r.pmin = V3(std::min(a.pmin.x, b.pmin.x), std::min(a.pmin.y, b.pmin.y), std::min(a.pmin.z, b.pmin.z));
    r.pmax = V3(std::max(a.pmax.x, b.pmax.x), std::max(a.pmax.y, b.pmax.y), std::max(a.pmax.z, b.pmax.z));
    return r;
}
inline Bounds3 bunion(const Bounds3 &a, const V3 &p) { return bunion(a, Bounds3(p)); }

// ---------------------------------------------------------------- PCG32
// core/rng.rs:9-83 (pbrt-v3 RNG). advance() follows pbrt-v3 (divergence 25:
// the reference loops `while delta > 1` with non-wrapping arithmetic).
static const uint64_t PCG32_DEFAULT_STATE = 0x853c49e6748fea9bULL;
static const uint64_t PCG32_DEFAULT_STREAM = 0xda3e39cb94b95bdbULL;
static const uint64_t PCG32_MULT = 0x5851f42d4c957f2dULL;
static const float OneMinusEpsilon = 0x1.fffffep-1f;

struct Rng {
    uint64_t state, inc;
    Rng() : state(PCG32_DEFAULT_STATE), inc(PCG32_DEFAULT_STREAM) {}
    explicit Rng(uint64_t seq) { set_sequence(seq); }
    // rng.rs:30-36
    void set_sequence(uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        uniform_uint32();
        state += PCG32_DEFAULT_STATE;
        uniform_uint32();
    }
    // rng.rs:38-49
    uint32_t uniform_uint32() {
        uint64_t oldstate = state;
        state = oldstate * PCG32_MULT + inc;
        uint32_t xorshifted = (uint32_t)(((oldstate >> 18u) ^ oldstate) >> 27u);
        uint32_t rot = (uint32_t)(oldstate >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31));
    }
    // rng.rs:60-63
    Float uniform_float() { return std::min(OneMinusEpsilon, (Float)uniform_uint32() * 0x1p-32f); }
    // rng.rs:65-82
    void advance(int64_t idelta) {
        uint64_t cur_mult = PCG32_MULT, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
        uint64_t delta = (uint64_t)idelta;
        while (delta > 0) {
            if (delta & 1) {
                acc_mult *= cur_mult;
                acc_plus = acc_plus * cur_mult + cur_plus;
            }
            cur_plus = (cur_mult + 1) * cur_plus;
            cur_mult *= cur_mult;
            delta /= 2;
        }
        state = acc_mult * state + acc_plus;
    }
};

}  // namespace orc
