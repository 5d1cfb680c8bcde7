// ORACLE — test infrastructure only (see pbrt.h). Triangle and sphere
// intersection, EFloat, transforms with error, and the surface interaction the
// path tracer shades with.
#pragma once
#include "pbrt.h"

namespace orc {

// ---------------------------------------------------------------- EFloat
// core/efloat.rs:5-258 with pbrt-v3 semantics where the reference diverges:
// scalar operands are exact (divergence 13: the reference builds new(v, v)),
// and subtraction pairs low-high / high-low (the reference pairs low-low).
struct EFloat {
    float v, low, high;
    EFloat() : v(0), low(0), high(0) {}
    // efloat.rs:12-28
    EFloat(float v_, float err = 0.f) : v(v_) {
        if (err == 0.f) low = high = v;
        else {
            low = next_float_down(v - err);
            high = next_float_up(v + err);
        }
    }
    float lower_bound() const { return low; }
    float upper_bound() const { return high; }
    // efloat.rs:97-107
    EFloat operator+(const EFloat &e) const {
        EFloat r;
        r.v = v + e.v;
        r.low = next_float_down(low + e.low);
        r.high = next_float_up(high + e.high);
        return r;
    }
    // efloat.rs:117-127 (pbrt-v3 interval subtraction)
    EFloat operator-(const EFloat &e) const {
        EFloat r;
        r.v = v - e.v;
        r.low = next_float_down(low - e.high);
        r.high = next_float_up(high - e.low);
        return r;
    }
    // efloat.rs:147-163
    EFloat operator*(const EFloat &e) const {
        EFloat r;
        r.v = v * e.v;
        float p0 = low * e.low, p1 = high * e.low, p2 = low * e.high, p3 = high * e.high;
        r.low = next_float_down(std::min(std::min(p0, p1), std::min(p2, p3)));
        r.high = next_float_up(std::max(std::max(p0, p1), std::max(p2, p3)));
        return r;
    }
    // efloat.rs:181-206
    EFloat operator/(const EFloat &e) const {
        EFloat r;
        r.v = v / e.v;
        if (e.low < 0 && e.high > 0) {
            r.low = -Infinity;
            r.high = Infinity;
            return r;
        }
        float d0 = low / e.low, d1 = high / e.low, d2 = low / e.high, d3 = high / e.high;
        r.low = next_float_down(std::min(std::min(d0, d1), std::min(d2, d3)));
        r.high = next_float_up(std::max(std::max(d0, d1), std::max(d2, d3)));
        return r;
    }
};

// efloat.rs:226-257
bool quadratic(EFloat a, EFloat b, EFloat c, EFloat *t0, EFloat *t1);

// ---------------------------------------------------------------- transform
// core/transform.rs (pbrt-v3 Transform). Only the 4x4 forward matrix is needed
// on this path; the inverse is kept beside it by the caller.
struct Mat4 {
    float m[4][4];
};
Mat4 mat_identity();
Mat4 mat_mul(const Mat4 &a, const Mat4 &b);  // matrix.rs:51-79 (divergence 18 fixed)
bool mat_inverse(const Mat4 &a, Mat4 *out);   // matrix.rs:84-162

struct Transform {
    Mat4 m, m_inv;
};
Transform transform_translate(const V3 &d);        // transform.rs:63-77
Transform transform_scale(float x, float y, float z);  // transform.rs:79-92
Transform transform_look_at(const V3 &pos, const V3 &look, const V3 &up);  // transform.rs:181-208
Transform transform_perspective(float fov, float n, float f);  // transform.rs:539-547 (divergence 17 fixed)
Transform transform_compose(const Transform &a, const Transform &b);  // a * b
Transform transform_inverse(const Transform &t);    // transform.rs:47-52 (divergence 15 fixed)
V3 xform_point(const Transform &t, const V3 &p);
V3 xform_vector(const Transform &t, const V3 &v);
V3 xform_normal(const Transform &t, const V3 &n);
// transform.rs:225-251
V3 xform_point_with_error(const Transform &t, const V3 &p, V3 *p_error);
// transform.rs:253-294 (point with incoming absolute error)
V3 xform_point_with_abs_error(const Transform &t, const V3 &p, const V3 &pe_in, V3 *pe_out);
// transform.rs:296-310
V3 xform_vector_with_error(const Transform &t, const V3 &v, V3 *v_error);
// transform.rs:392-434
Ray xform_ray_with_error(const Transform &t, const Ray &r, V3 *o_error, V3 *d_error);

// ---------------------------------------------------------------- interaction
// core/interaction.rs:233-316 SurfaceInteraction (the parts the path uses).
struct SurfaceInteraction {
    V3 p, p_error, n, wo;
    float u = 0, v = 0;
    V3 dpdu, dpdv;
    V3 shading_n, shading_dpdu;
    float time = 0;
    float t = Infinity;
    float b1 = 0, b2 = 0;  // triangle barycentrics of the hit (0 for spheres)
    uint32_t prim = 0xffffffffu;
    // interaction.rs:33-43
    Ray spawn_ray(const V3 &d) const {
        V3 o = offset_ray_origin(p, p_error, n, d);
        return Ray(o, d, Infinity, time);
    }
    // interaction.rs:46-57
    Ray spawn_ray_to_point(const V3 &p2) const {
        V3 d = p2 - p;
        V3 o = offset_ray_origin(p, p_error, n, d);
        return Ray(o, d, 1 - ShadowEpsilon, time);
    }
    // interaction.rs:60-72
    Ray spawn_ray_to(const V3 &ip, const V3 &ip_error, const V3 &in) const {
        V3 o = offset_ray_origin(p, p_error, n, ip - p);
        V3 target = offset_ray_origin(ip, ip_error, in, o - ip);
        V3 d = target - o;
        return Ray(o, d, 1 - ShadowEpsilon, time);
    }
};

// ---------------------------------------------------------------- triangle
// shapes/triangle.rs:142-419 (intersect) and 421-607 (intersect_p), with mesh
// vertex positions already in world space (triangle.rs:16-74). Meshes here
// carry neither uv, normals nor tangents, so get_uvs (96-111) yields the
// default (0,0),(1,0),(1,1) parameterisation.
struct TriHit {
    float t, b0, b1, b2;
};
// Watertight test only: returns true and the hit when the ray hits in (0, t_max].
bool triangle_test(const V3 &p0, const V3 &p1, const V3 &p2, const Ray &ray, TriHit *h);
// Full closest-hit intersect: the test above plus the degenerate-triangle
// rejection and surface interaction (triangle.rs:258-345).
bool triangle_intersect(const V3 &p0, const V3 &p1, const V3 &p2, bool reverse_orientation, const Ray &ray,
                        SurfaceInteraction *si, float *t_hit);
// triangle.rs:421-607 (alpha-mask branch out of scope)
bool triangle_intersect_p(const V3 &p0, const V3 &p1, const V3 &p2, const Ray &ray);
// triangle.rs:609-614
float triangle_area(const V3 &p0, const V3 &p1, const V3 &p2);
// triangle.rs:616-642 (divergence 9 fixed): sampled point, normal and p_error.
void triangle_sample(const V3 &p0, const V3 &p1, const V3 &p2, bool reverse_orientation, float u0, float u1,
                     V3 *p, V3 *n, V3 *p_error);

// ---------------------------------------------------------------- sphere
// shapes/sphere.rs: a full sphere of `radius` at `center` (object_to_world =
// Translate(center)); z_min = -r, z_max = r, phi_max = 360 degrees.
struct SphereShape {
    V3 center;
    float radius;
    bool reverse_orientation = false;
    Transform object_to_world, world_to_object;
    float z_min, z_max, theta_min, theta_max, phi_max;
    void init(const V3 &c, float r);  // sphere.rs:50-77
    Bounds3 world_bound() const;      // sphere.rs:81-99 + transform.rs:482
};
bool sphere_intersect(const SphereShape &s, const Ray &r, SurfaceInteraction *si, float *t_hit);  // sphere.rs:105-236
bool sphere_intersect_p(const SphereShape &s, const Ray &r);  // sphere.rs:238-323

}  // namespace orc
