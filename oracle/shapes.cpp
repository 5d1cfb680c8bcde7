// ORACLE — test infrastructure only (see pbrt.h).
#include "shapes.h"

namespace orc {

// efloat.rs:226-257. The root-discriminant error is MACHINE_EPSILON *
// root_discrim in f32, as the reference computes it (efloat.rs:234-237).
bool quadratic(EFloat a, EFloat b, EFloat c, EFloat *t0, EFloat *t1) {
    double discrim = (double)b.v * (double)b.v - 4. * (double)a.v * (double)c.v;
    if (discrim < 0.) return false;
    double root_discrim = std::sqrt(discrim);
    float rd = (float)root_discrim;
    EFloat float_root_discrim(rd, MachineEpsilon * rd);
    EFloat q;
    if (b.v < 0) q = (b - float_root_discrim) * EFloat(-0.5f);
    else q = (b + float_root_discrim) * EFloat(-0.5f);
    *t0 = q / a;
    *t1 = c / q;
    if (t0->v > t1->v) std::swap(*t0, *t1);
    return true;
}

// ---------------------------------------------------------------- matrices
Mat4 mat_identity() {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = (i == j) ? 1.f : 0.f;
    return r;
}
// matrix.rs:51-79; the reference mixes rows of m2 (divergence 18), this is the
// standard product with pbrt-v3's summation order.
Mat4 mat_mul(const Mat4 &a, const Mat4 &b) {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] + a.m[i][3] * b.m[3][j];
    return r;
}
// matrix.rs:84-162: Gauss-Jordan with full pivoting (numerically stable
// inversion as pbrt-v3 does it).
bool mat_inverse(const Mat4 &a, Mat4 *out) {
    int indxc[4], indxr[4];
    int ipiv[4] = {0, 0, 0, 0};
    float minv[4][4];
    std::memcpy(minv, a.m, sizeof(minv));
    for (int i = 0; i < 4; i++) {
        int irow = 0, icol = 0;
        float big = 0.f;
        for (int j = 0; j < 4; j++) {
            if (ipiv[j] != 1) {
                for (int k = 0; k < 4; k++) {
                    if (ipiv[k] == 0) {
                        if (std::fabs(minv[j][k]) >= big) {
                            big = std::fabs(minv[j][k]);
                            irow = j;
                            icol = k;
                        }
                    } else if (ipiv[k] > 1)
                        return false;
                }
            }
        }
        ++ipiv[icol];
        if (irow != icol)
            for (int k = 0; k < 4; ++k) std::swap(minv[irow][k], minv[icol][k]);
        indxr[i] = irow;
        indxc[i] = icol;
        if (minv[icol][icol] == 0.f) return false;
        float pivinv = 1.f / minv[icol][icol];
        minv[icol][icol] = 1.f;
        for (int j = 0; j < 4; j++) minv[icol][j] *= pivinv;
        for (int j = 0; j < 4; j++) {
            if (j != icol) {
                float save = minv[j][icol];
                minv[j][icol] = 0;
                for (int k = 0; k < 4; k++) minv[j][k] -= minv[icol][k] * save;
            }
        }
    }
    for (int j = 3; j >= 0; j--) {
        if (indxr[j] != indxc[j]) {
            for (int k = 0; k < 4; k++) std::swap(minv[k][indxr[j]], minv[k][indxc[j]]);
        }
    }
    std::memcpy(out->m, minv, sizeof(minv));
    return true;
}

static Mat4 mat_transpose(const Mat4 &a) {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = a.m[j][i];
    return r;
}

Transform transform_translate(const V3 &d) {
    Transform t;
    t.m = mat_identity();
    t.m_inv = mat_identity();
    t.m.m[0][3] = d.x; t.m.m[1][3] = d.y; t.m.m[2][3] = d.z;
    t.m_inv.m[0][3] = -d.x; t.m_inv.m[1][3] = -d.y; t.m_inv.m[2][3] = -d.z;
    return t;
}
Transform transform_scale(float x, float y, float z) {
    Transform t;
    t.m = mat_identity();
    t.m_inv = mat_identity();
    t.m.m[0][0] = x; t.m.m[1][1] = y; t.m.m[2][2] = z;
    t.m_inv.m[0][0] = 1.f / x; t.m_inv.m[1][1] = 1.f / y; t.m_inv.m[2][2] = 1.f / z;
    return t;
}
// transform.rs:181-208
Transform transform_look_at(const V3 &pos, const V3 &look, const V3 &up) {
    Mat4 c2w;
    std::memset(&c2w, 0, sizeof(c2w));
    c2w.m[0][3] = pos.x; c2w.m[1][3] = pos.y; c2w.m[2][3] = pos.z; c2w.m[3][3] = 1;
    V3 dir = normalize(look - pos);
    V3 right = normalize(cross(normalize(up), dir));
    V3 new_up = cross(dir, right);
    c2w.m[0][0] = right.x; c2w.m[1][0] = right.y; c2w.m[2][0] = right.z; c2w.m[3][0] = 0.;
    c2w.m[0][1] = new_up.x; c2w.m[1][1] = new_up.y; c2w.m[2][1] = new_up.z; c2w.m[3][1] = 0.;
    c2w.m[0][2] = dir.x; c2w.m[1][2] = dir.y; c2w.m[2][2] = dir.z; c2w.m[3][2] = 0.;
    Transform t;
    mat_inverse(c2w, &t.m);
    t.m_inv = c2w;
    return t;
}
// transform.rs:539-547; the reference drops the projective matrix (divergence
// 17). pbrt-v3: Scale(invTan, invTan, 1) * Transform(persp).
Transform transform_perspective(float fov, float n, float f) {
    Mat4 persp = mat_identity();
    persp.m[2][2] = f / (f - n);
    persp.m[2][3] = -f * n / (f - n);
    persp.m[3][2] = 1;
    persp.m[3][3] = 0;
    Transform p;
    p.m = persp;
    mat_inverse(persp, &p.m_inv);
    float inv_tan_ang = 1.f / std::tan((Pi / 180.f * fov) / 2.f);
    return transform_compose(transform_scale(inv_tan_ang, inv_tan_ang, 1), p);
}
Transform transform_compose(const Transform &a, const Transform &b) {
    Transform r;
    r.m = mat_mul(a.m, b.m);
    r.m_inv = mat_mul(b.m_inv, a.m_inv);
    return r;
}
// transform.rs:47-52 (divergence 15 fixed)
Transform transform_inverse(const Transform &t) {
    Transform r;
    r.m = t.m_inv;
    r.m_inv = t.m;
    return r;
}
V3 xform_point(const Transform &t, const V3 &p) {
    const float(*m)[4] = t.m.m;
    float x = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
    float y = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
    float z = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
    float w = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
    if (w == 1) return V3(x, y, z);
    return V3(x, y, z) / w;
}
V3 xform_vector(const Transform &t, const V3 &v) {
    const float(*m)[4] = t.m.m;
    return V3(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
              m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
}
V3 xform_normal(const Transform &t, const V3 &n) {
    const float(*mi)[4] = t.m_inv.m;
    return V3(mi[0][0] * n.x + mi[1][0] * n.y + mi[2][0] * n.z, mi[0][1] * n.x + mi[1][1] * n.y + mi[2][1] * n.z,
              mi[0][2] * n.x + mi[1][2] * n.y + mi[2][2] * n.z);
}
V3 xform_point_with_error(const Transform &t, const V3 &p, V3 *p_error) {
    const float(*m)[4] = t.m.m;
    float x = p.x, y = p.y, z = p.z;
    float xp = (m[0][0] * x + m[0][1] * y) + (m[0][2] * z + m[0][3]);
    float yp = (m[1][0] * x + m[1][1] * y) + (m[1][2] * z + m[1][3]);
    float zp = (m[2][0] * x + m[2][1] * y) + (m[2][2] * z + m[2][3]);
    float wp = (m[3][0] * x + m[3][1] * y) + (m[3][2] * z + m[3][3]);
    float xs = std::fabs(m[0][0] * x) + std::fabs(m[0][1] * y) + std::fabs(m[0][2] * z) + std::fabs(m[0][3]);
    float ys = std::fabs(m[1][0] * x) + std::fabs(m[1][1] * y) + std::fabs(m[1][2] * z) + std::fabs(m[1][3]);
    float zs = std::fabs(m[2][0] * x) + std::fabs(m[2][1] * y) + std::fabs(m[2][2] * z) + std::fabs(m[2][3]);
    *p_error = V3(xs, ys, zs) * gamma(3);
    if (wp == 1) return V3(xp, yp, zp);
    return V3(xp, yp, zp) / wp;
}
// transform.rs:253-294; the z error uses m[0][3] in the reference
// (transform.rs:279, a typo); pbrt-v3 uses m[2][3].
V3 xform_point_with_abs_error(const Transform &t, const V3 &p, const V3 &pe, V3 *abs_error) {
    const float(*m)[4] = t.m.m;
    float x = p.x, y = p.y, z = p.z;
    float xp = (m[0][0] * x + m[0][1] * y) + (m[0][2] * z + m[0][3]);
    float yp = (m[1][0] * x + m[1][1] * y) + (m[1][2] * z + m[1][3]);
    float zp = (m[2][0] * x + m[2][1] * y) + (m[2][2] * z + m[2][3]);
    float wp = (m[3][0] * x + m[3][1] * y) + (m[3][2] * z + m[3][3]);
    float g3 = gamma(3);
    for (int i = 0; i < 3; ++i) {
        float v = (g3 + (float)1) * (std::fabs(m[i][0]) * pe.x + std::fabs(m[i][1]) * pe.y + std::fabs(m[i][2]) * pe.z) +
                  g3 * (std::fabs(m[i][0] * x) + std::fabs(m[i][1] * y) + std::fabs(m[i][2] * z) + std::fabs(m[i][3]));
        (*abs_error)[i] = v;
    }
    if (wp == 1) return V3(xp, yp, zp);
    return V3(xp, yp, zp) / wp;
}
V3 xform_vector_with_error(const Transform &t, const V3 &v, V3 *e) {
    const float(*m)[4] = t.m.m;
    float x = v.x, y = v.y, z = v.z;
    e->x = gamma(3) * (std::fabs(m[0][0] * x) + std::fabs(m[0][1] * y) + std::fabs(m[0][2] * z));
    e->y = gamma(3) * (std::fabs(m[1][0] * x) + std::fabs(m[1][1] * y) + std::fabs(m[1][2] * z));
    e->z = gamma(3) * (std::fabs(m[2][0] * x) + std::fabs(m[2][1] * y) + std::fabs(m[2][2] * z));
    return V3(m[0][0] * x + m[0][1] * y + m[0][2] * z, m[1][0] * x + m[1][1] * y + m[1][2] * z,
              m[2][0] * x + m[2][1] * y + m[2][2] * z);
}
// transform.rs:392-434 (t_max is not shortened, as in the reference).
Ray xform_ray_with_error(const Transform &t, const Ray &r, V3 *o_error, V3 *d_error) {
    V3 o = xform_point_with_error(t, r.o, o_error);
    V3 d = xform_vector_with_error(t, r.d, d_error);
    float ls = length_squared(d);
    if (ls > 0) {
        float dt = dot(vabs(d), *o_error) / ls;
        o += d * dt;
    }
    return Ray(o, d, r.t_max, r.time);
}

// ---------------------------------------------------------------- triangle
// triangle.rs:142-256: translate, permute, shear, edge functions (f64 retry
// when one is exactly 0), sign and determinant tests, scaled t range test,
// conservative t > delta_t rejection.
bool triangle_test(const V3 &p0, const V3 &p1, const V3 &p2, const Ray &ray, TriHit *h) {
    V3 p0t = p0 - ray.o;
    V3 p1t = p1 - ray.o;
    V3 p2t = p2 - ray.o;
    int kz = max_dimension(vabs(ray.d));
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    V3 d = permute(ray.d, kx, ky, kz);
    p0t = permute(p0t, kx, ky, kz);
    p1t = permute(p1t, kx, ky, kz);
    p2t = permute(p2t, kx, ky, kz);
    float sx = -d.x / d.z;
    float sy = -d.y / d.z;
    float sz = 1.f / d.z;
    p0t.x += sx * p0t.z;
    p0t.y += sy * p0t.z;
    p1t.x += sx * p1t.z;
    p1t.y += sy * p1t.z;
    p2t.x += sx * p2t.z;
    p2t.y += sy * p2t.z;
    float e0 = p1t.x * p2t.y - p1t.y * p2t.x;
    float e1 = p2t.x * p0t.y - p2t.y * p0t.x;
    float e2 = p0t.x * p1t.y - p0t.y * p1t.x;
    // triangle.rs:192-204
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {
        double p2txp1ty = (double)p2t.x * (double)p1t.y;
        double p2typ1tx = (double)p2t.y * (double)p1t.x;
        e0 = (float)(p2typ1tx - p2txp1ty);
        double p0txp2ty = (double)p0t.x * (double)p2t.y;
        double p0typ2tx = (double)p0t.y * (double)p2t.x;
        e1 = (float)(p0typ2tx - p0txp2ty);
        double p1txp0ty = (double)p1t.x * (double)p0t.y;
        double p1typ0tx = (double)p1t.y * (double)p0t.x;
        e2 = (float)(p1typ0tx - p1txp0ty);
    }
    if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return false;
    float det = e0 + e1 + e2;
    if (det == 0) return false;
    p0t.z *= sz;
    p1t.z *= sz;
    p2t.z *= sz;
    float t_scaled = e0 * p0t.z + e1 * p1t.z + e2 * p2t.z;
    if (det < 0 && (t_scaled >= 0 || t_scaled < ray.t_max * det)) return false;
    else if (det > 0 && (t_scaled <= 0 || t_scaled > ray.t_max * det)) return false;
    float inv_det = 1 / det;
    float b0 = e0 * inv_det;
    float b1 = e1 * inv_det;
    float b2 = e2 * inv_det;
    float t = t_scaled * inv_det;
    // triangle.rs:227-256 (max_component per pbrt-v3, divergence 6)
    float max_zt = max_component(vabs(V3(p0t.z, p1t.z, p2t.z)));
    float delta_z = gamma(3) * max_zt;
    float max_xt = max_component(vabs(V3(p0t.x, p1t.x, p2t.x)));
    float max_yt = max_component(vabs(V3(p0t.y, p1t.y, p2t.y)));
    float delta_x = gamma(5) * (max_xt + max_zt);
    float delta_y = gamma(5) * (max_yt + max_zt);
    float delta_e = 2 * (gamma(2) * max_xt * max_yt + delta_y * max_xt + delta_x * max_yt);
    float max_e = max_component(vabs(V3(e0, e1, e2)));
    float delta_t = 3 * (gamma(3) * max_e * max_zt + delta_e * max_zt + delta_z * max_e) * std::fabs(inv_det);
    if (t <= delta_t) return false;
    h->t = t;
    h->b0 = b0;
    h->b1 = b1;
    h->b2 = b2;
    return true;
}

// triangle.rs:258-345 with the default uv parameterisation. The reference's uv
// determinant is identically zero (divergence 8); pbrt-v3's is used.
bool triangle_intersect(const V3 &p0, const V3 &p1, const V3 &p2, bool reverse_orientation, const Ray &ray,
                        SurfaceInteraction *si, float *t_hit) {
    TriHit h;
    if (!triangle_test(p0, p1, p2, ray, &h)) return false;
    // default uvs (triangle.rs:96-111)
    const float uv[3][2] = {{0, 0}, {1, 0}, {1, 1}};
    float duv02x = uv[0][0] - uv[2][0], duv02y = uv[0][1] - uv[2][1];
    float duv12x = uv[1][0] - uv[2][0], duv12y = uv[1][1] - uv[2][1];
    V3 dp02 = p0 - p2, dp12 = p1 - p2;
    float determinant = duv02x * duv12y - duv02y * duv12x;
    bool degenerate_uv = std::fabs(determinant) < 1e-8f;
    V3 dpdu, dpdv;
    if (!degenerate_uv) {
        float invdet = 1 / determinant;
        dpdu = (duv12y * dp02 - duv02y * dp12) * invdet;
        dpdv = (-duv12x * dp02 + duv02x * dp12) * invdet;
    }
    if (degenerate_uv || length_squared(cross(dpdu, dpdv)) == 0) {
        V3 ng = cross(p2 - p0, p1 - p0);
        if (length_squared(ng) == 0) return false;
        coordinate_system(normalize(ng), &dpdu, &dpdv);
    }
    float b0 = h.b0, b1 = h.b1, b2 = h.b2;
    float x_abs_sum = std::fabs(b0 * p0.x) + std::fabs(b1 * p1.x) + std::fabs(b2 * p2.x);
    float y_abs_sum = std::fabs(b0 * p0.y) + std::fabs(b1 * p1.y) + std::fabs(b2 * p2.y);
    float z_abs_sum = std::fabs(b0 * p0.z) + std::fabs(b1 * p1.z) + std::fabs(b2 * p2.z);
    V3 p_error = gamma(7) * V3(x_abs_sum, y_abs_sum, z_abs_sum);
    V3 p_hit = b0 * p0 + b1 * p1 + b2 * p2;
    float u_hit = b0 * uv[0][0] + b1 * uv[1][0] + b2 * uv[2][0];
    float v_hit = b0 * uv[0][1] + b1 * uv[1][1] + b2 * uv[2][1];
    si->p = p_hit;
    si->p_error = p_error;
    si->u = u_hit;
    si->v = v_hit;
    si->wo = -ray.d;
    si->dpdu = dpdu;
    si->dpdv = dpdv;
    si->time = ray.time;
    // triangle.rs:332-340: geometric normal from the vertex order
    V3 n = normalize(cross(dp02, dp12));
    if (reverse_orientation) n = -n;
    si->n = n;
    si->shading_n = n;
    si->shading_dpdu = dpdu;
    si->t = h.t;
    si->b1 = h.b1;
    si->b2 = h.b2;
    *t_hit = h.t;
    return true;
}

bool triangle_intersect_p(const V3 &p0, const V3 &p1, const V3 &p2, const Ray &ray) {
    TriHit h;
    return triangle_test(p0, p1, p2, ray, &h);
}

float triangle_area(const V3 &p0, const V3 &p1, const V3 &p2) { return 0.5f * length(cross(p1 - p0, p2 - p0)); }

// sampling.rs:264-267 uniform_sample_triangle
static void uniform_sample_triangle(float u0, float u1, float *b0, float *b1) {
    float su0 = std::sqrt(u0);
    *b0 = 1 - su0;
    *b1 = u1 * su0;
}

// triangle.rs:616-642
void triangle_sample(const V3 &p0, const V3 &p1, const V3 &p2, bool reverse_orientation, float u0, float u1,
                     V3 *p, V3 *n, V3 *p_error) {
    float b0, b1;
    uniform_sample_triangle(u0, u1, &b0, &b1);
    *p = b0 * p0 + b1 * p1 + (1 - b0 - b1) * p2;
    *n = normalize(cross(p1 - p0, p2 - p0));
    if (reverse_orientation) *n = -*n;
    V3 pa = vabs(b0 * p0) + vabs(b1 * p1) + vabs((1 - b0 - b1) * p2);
    *p_error = gamma(6) * pa;
}

// ---------------------------------------------------------------- sphere
void SphereShape::init(const V3 &c, float r) {
    center = c;
    radius = r;
    object_to_world = transform_translate(c);
    world_to_object = transform_inverse(object_to_world);
    // sphere.rs:50-77 with z0 = -r, z1 = r, phi = 360
    float zmin = -r, zmax = r;
    z_min = clamp(std::min(zmin, zmax), -radius, radius);
    z_max = clamp(std::max(zmin, zmax), -radius, radius);
    theta_min = std::acos(clamp(std::min(zmin, zmax) / radius, -1.f, 1.f));
    theta_max = std::acos(clamp(std::max(zmin, zmax) / radius, -1.f, 1.f));
    phi_max = (Pi / 180.f) * clamp(360.f, 0.f, 360.f);
}
Bounds3 SphereShape::world_bound() const {
    // transform of the 8 corners of the object bound (bounds3.rs transform)
    V3 lo(-radius, -radius, z_min), hi(radius, radius, z_max);
    Bounds3 b;
    for (int i = 0; i < 8; ++i) {
        V3 c((i & 1) ? hi.x : lo.x, (i & 2) ? hi.y : lo.y, (i & 4) ? hi.z : lo.z);
        b = bunion(b, xform_point(object_to_world, c));
    }
    return b;
}

// sphere.rs:105-236. The hit side: t0/t1 bounds tests, refinement, phi clip.
static bool sphere_hit_core(const SphereShape &s, const Ray &r, Ray *oray, EFloat *t_hit, V3 *p_hit_out, float *phi_out) {
    V3 o_err, d_err;
    Ray ray = xform_ray_with_error(s.world_to_object, r, &o_err, &d_err);
    EFloat ox(ray.o.x, o_err.x), oy(ray.o.y, o_err.y), oz(ray.o.z, o_err.z);
    EFloat dx(ray.d.x, d_err.x), dy(ray.d.y, d_err.y), dz(ray.d.z, d_err.z);
    EFloat a = dx * dx + dy * dy + dz * dz;
    EFloat b = (dx * ox + dy * oy + dz * oz) * EFloat(2.f);
    EFloat c = ox * ox + oy * oy + oz * oz - EFloat(s.radius) * EFloat(s.radius);
    EFloat t0, t1;
    if (!quadratic(a, b, c, &t0, &t1)) return false;
    if (t0.upper_bound() > ray.t_max || t1.lower_bound() <= 0) return false;
    EFloat t_shape_hit = t0;
    if (t_shape_hit.lower_bound() <= 0) {
        t_shape_hit = t1;
        if (t_shape_hit.upper_bound() > ray.t_max) return false;
    }
    V3 p_hit = ray.at(t_shape_hit.v);
    p_hit *= s.radius / distance(p_hit, V3(0, 0, 0));
    if (p_hit.x == 0 && p_hit.y == 0) p_hit.x = 1e-5f * s.radius;
    float phi = std::atan2(p_hit.y, p_hit.x);
    if (phi < 0) phi += 2 * Pi;
    if ((s.z_min > -s.radius && p_hit.z < s.z_min) || (s.z_max < s.radius && p_hit.z > s.z_max) || phi > s.phi_max) {
        if (t_shape_hit.v == t1.v) return false;
        if (t1.upper_bound() > ray.t_max) return false;
        t_shape_hit = t1;
        p_hit = ray.at(t_shape_hit.v);
        p_hit *= s.radius / distance(p_hit, V3(0, 0, 0));
        if (p_hit.x == 0 && p_hit.y == 0) p_hit.x = 1e-5f * s.radius;
        phi = std::atan2(p_hit.y, p_hit.x);
        if (phi < 0) phi += 2 * Pi;
        if ((s.z_min > -s.radius && p_hit.z < s.z_min) || (s.z_max < s.radius && p_hit.z > s.z_max) || phi > s.phi_max)
            return false;
    }
    *oray = ray;
    *t_hit = t_shape_hit;
    *p_hit_out = p_hit;
    *phi_out = phi;
    return true;
}

bool sphere_intersect(const SphereShape &s, const Ray &r, SurfaceInteraction *si, float *t_hit) {
    Ray ray;
    EFloat tsh;
    V3 p_hit;
    float phi;
    if (!sphere_hit_core(s, r, &ray, &tsh, &p_hit, &phi)) return false;
    float u = phi / s.phi_max;
    float theta = std::acos(clamp(p_hit.z / s.radius, -1.f, 1.f));
    float v = (theta - s.theta_min) / (s.theta_max - s.theta_min);
    float z_radius = std::sqrt(p_hit.x * p_hit.x + p_hit.y * p_hit.y);
    float inv_z_radius = 1 / z_radius;
    float cos_phi = p_hit.x * inv_z_radius;
    float sin_phi = p_hit.y * inv_z_radius;
    V3 dpdu(-s.phi_max * p_hit.y, s.phi_max * p_hit.x, 0);
    V3 dpdv = (s.theta_max - s.theta_min) * V3(p_hit.z * cos_phi, p_hit.z * sin_phi, -s.radius * std::sin(theta));
    V3 p_error = gamma(5) * vabs(p_hit);
    // Object-space interaction (interaction.rs:263-316; divergence 12: the
    // reference negates the partials) followed by the object-to-world
    // transform (transform.rs:449-479).
    V3 n_obj = normalize(cross(dpdu, dpdv));
    if (s.reverse_orientation) n_obj = -n_obj;
    V3 pe_w;
    si->p = xform_point_with_abs_error(s.object_to_world, p_hit, p_error, &pe_w);
    si->p_error = pe_w;
    si->n = normalize(xform_normal(s.object_to_world, n_obj));
    si->wo = normalize(xform_vector(s.object_to_world, -ray.d));
    si->u = u;
    si->v = v;
    si->dpdu = xform_vector(s.object_to_world, dpdu);
    si->dpdv = xform_vector(s.object_to_world, dpdv);
    si->shading_n = faceforward(normalize(xform_normal(s.object_to_world, n_obj)), si->n);
    si->shading_dpdu = si->dpdu;
    si->time = r.time;
    si->t = tsh.v;
    si->b1 = 0;
    si->b2 = 0;
    *t_hit = tsh.v;
    return true;
}

bool sphere_intersect_p(const SphereShape &s, const Ray &r) {
    Ray ray;
    EFloat tsh;
    V3 p_hit;
    float phi;
    return sphere_hit_core(s, r, &ray, &tsh, &p_hit, &phi);
}

}  // namespace orc
