// ORACLE — test infrastructure only (see pbrt.h).
#include "render.h"

namespace orc {

// ---------------------------------------------------------------- sampling
// sampling.rs:8-93 (discrete sampling with pbrt-v3's FindInterval)
void Distribution1D::init(const std::vector<float> &f) {
    func = f;
    size_t n = f.size();
    cdf.assign(n + 1, 0.f);
    for (size_t i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + func[i - 1] / (float)n;
    func_int = cdf[n];
    if (func_int == 0)
        for (size_t i = 1; i < n + 1; ++i) cdf[i] = (float)i / (float)n;
    else
        for (size_t i = 1; i < n + 1; ++i) cdf[i] /= func_int;
}
int Distribution1D::sample_discrete(float u, float *pdf) const {
    int size = (int)cdf.size(), first = 0, len = size;
    while (len > 0) {
        int half = len >> 1, middle = first + half;
        if (cdf[middle] <= u) {
            first = middle + 1;
            len -= half + 1;
        } else
            len = half;
    }
    int offset = clamp(first - 1, 0, size - 2);
    if (pdf) *pdf = (func_int > 0) ? func[offset] / (func_int * (float)func.size()) : 0;
    return offset;
}

// splitmix64 finaliser (Steele, Lea, Flood 2014): decorrelates the stream
// selector so neighbouring pixels/samples get unrelated PCG32 increments.
static uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
uint64_t sampler_stream(uint64_t seed, uint64_t pixel, uint64_t sample) {
    return splitmix64(seed ^ splitmix64((pixel << 24) | (sample & 0xffffff)));
}
void Sampler::start(uint64_t seed, uint64_t pixel, uint64_t sample) {
    halton = nullptr;
    sobol = nullptr;
    rng.set_sequence(sampler_stream(seed, pixel, sample));
}

static float hashed_dimension(int64_t index, int d) {
    const uint64_t h = splitmix64((uint64_t)index * 0x9E3779B97F4A7C15ull + (uint64_t)d);
    return std::min(OneMinusEpsilon, (float)(uint32_t)(h >> 32) * 0x1p-32f);
}

float Sampler::halton_next() {
    const int d = dim++;
    if (d < kHaltonDims) return halton->sample_dimension(index, d);
    return hashed_dimension(index, d);
}

float Sampler::sobol_next() {
    const int d = dim++;
    if (d < kSobolDims) return sobol->sample_dimension(index, d, px, py);
    return hashed_dimension(index, d);
}

// sampling.rs:210-230
static void concentric_sample_disk(const float *u, float *dx, float *dy) {
    float ox = u[0] * 2.0f - 1.0f, oy = u[1] * 2.0f - 1.0f;
    if (ox == 0 && oy == 0) {
        *dx = *dy = 0;
        return;
    }
    float theta, r;
    if (std::fabs(ox) > std::fabs(oy)) {
        r = ox;
        theta = (Pi / 4) * (oy / ox);
    } else {
        r = oy;
        theta = (Pi / 2) - (Pi / 4) * (ox / oy);
    }
    *dx = r * std::cos(theta);
    *dy = r * std::sin(theta);
}
// sampling.rs:316-320
static V3 cosine_sample_hemisphere(const float *u) {
    float dx, dy;
    concentric_sample_disk(u, &dx, &dy);
    float z = std::sqrt(std::max(0.f, 1 - dx * dx - dy * dy));
    return V3(dx, dy, z);
}
// sampling.rs:168-181 with pbrt-v3's y = r sin(phi) (divergence 22) and
// pdf 1 / (2 pi) (divergence 23)
static V3 uniform_sample_hemisphere(const float *u) {
    float z = u[0];
    float r = std::sqrt(std::max(0.f, 1 - z * z));
    float phi = 2 * Pi * u[1];
    return V3(r * std::cos(phi), r * std::sin(phi), z);
}
// sampling.rs:333-337
static float power_heuristic(int nf, float f_pdf, int ng, float g_pdf) {
    float f = nf * f_pdf, g = ng * g_pdf;
    return (f * f) / (f * f + g * g);
}

// ---------------------------------------------------------------- BSDF
// reflection/mod.rs:479-568 local-frame helpers
static inline float cos_theta(const V3 &w) { return w.z; }
static inline float cos2_theta(const V3 &w) { return w.z * w.z; }
static inline float abs_cos_theta(const V3 &w) { return std::fabs(w.z); }
static inline float sin2_theta(const V3 &w) { return std::max(0.f, 1 - cos2_theta(w)); }
static inline float sin_theta(const V3 &w) { return std::sqrt(sin2_theta(w)); }
static inline float tan_theta(const V3 &w) { return sin_theta(w) / cos_theta(w); }
static inline float tan2_theta(const V3 &w) { return sin2_theta(w) / cos2_theta(w); }
static inline float cos_phi(const V3 &w) {
    float st = sin_theta(w);
    return (st == 0) ? 1 : clamp(w.x / st, -1.f, 1.f);
}
static inline float sin_phi(const V3 &w) {
    float st = sin_theta(w);
    return (st == 0) ? 0 : clamp(w.y / st, -1.f, 1.f);
}
static inline float cos2_phi(const V3 &w) { return cos_phi(w) * cos_phi(w); }
static inline float sin2_phi(const V3 &w) { return sin_phi(w) * sin_phi(w); }
static inline bool same_hemisphere(const V3 &w, const V3 &wp) { return w.z * wp.z > 0; }
static inline V3 reflect(const V3 &wo, const V3 &n) { return -wo + 2 * dot(wo, n) * n; }
// reflection/mod.rs:547-560 (pbrt-v3 semantics)
static bool refract(const V3 &wi, const V3 &n, float eta, V3 *wt) {
    float cos_theta_i = dot(n, wi);
    float sin2_theta_i = std::max(0.f, 1.f - cos_theta_i * cos_theta_i);
    float sin2_theta_t = eta * eta * sin2_theta_i;
    if (sin2_theta_t >= 1) return false;
    float cos_theta_t = std::sqrt(1 - sin2_theta_t);
    *wt = eta * -wi + (eta * cos_theta_i - cos_theta_t) * n;
    return true;
}
// reflection/mod.rs:420-445
static float fr_dielectric(float cos_theta_i, float eta_i, float eta_t) {
    cos_theta_i = clamp(cos_theta_i, -1.f, 1.f);
    bool entering = cos_theta_i > 0.f;
    if (!entering) {
        std::swap(eta_i, eta_t);
        cos_theta_i = std::fabs(cos_theta_i);
    }
    float sin_theta_i = std::sqrt(std::max(0.f, 1 - cos_theta_i * cos_theta_i));
    float sin_theta_t = eta_i / eta_t * sin_theta_i;
    if (sin_theta_t >= 1) return 1;
    float cos_theta_t = std::sqrt(std::max(0.f, 1 - sin_theta_t * sin_theta_t));
    float r_parl = ((eta_t * cos_theta_i) - (eta_i * cos_theta_t)) / ((eta_t * cos_theta_i) + (eta_i * cos_theta_t));
    float r_perp = ((eta_i * cos_theta_i) - (eta_t * cos_theta_t)) / ((eta_i * cos_theta_i) + (eta_t * cos_theta_t));
    return (r_parl * r_parl + r_perp * r_perp) / 2;
}
// reflection/mod.rs:448-475, per channel
static float fr_conductor1(float cos_theta_i, float eta_i, float eta_t, float k) {
    cos_theta_i = clamp(cos_theta_i, -1.f, 1.f);
    float eta = eta_t / eta_i, etak = k / eta_i;
    float cos2 = cos_theta_i * cos_theta_i, sin2 = 1.f - cos2;
    float eta2 = eta * eta, etak2 = etak * etak;
    float t0 = eta2 - etak2 - sin2;
    float a2plusb2 = std::sqrt(t0 * t0 + 4 * eta2 * etak2);
    float t1 = a2plusb2 + cos2;
    float a = std::sqrt(0.5f * (a2plusb2 + t0));
    float t2 = 2.f * cos_theta_i * a;
    float rs = (t1 - t2) / (t1 + t2);
    float t3 = cos2 * a2plusb2 + sin2 * sin2;
    float t4 = t2 * sin2;
    float rp = rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (rp + rs);
}
// fresnel.rs:25-41 FresnelConductor::evaluate
static RGB fr_conductor(float cos_theta_i, const RGB &eta, const RGB &k) {
    float c = std::fabs(cos_theta_i);
    return RGB(fr_conductor1(c, 1.f, eta.x, k.x), fr_conductor1(c, 1.f, eta.y, k.y), fr_conductor1(c, 1.f, eta.z, k.z));
}

// trowbridge_reitz_distribution.rs:8-182 (isotropic alpha), pbrt-v3
// semantics: Lambda uses |tan| (divergence 31), sample_wh samples visible
// normals (divergence 32).
struct TR {
    float ax, ay;
    float D(const V3 &wh) const {
        float t2 = tan2_theta(wh);
        if (std::isinf(t2)) return 0.;
        float cos4 = cos2_theta(wh) * cos2_theta(wh);
        float e = (cos2_phi(wh) / (ax * ax) + sin2_phi(wh) / (ay * ay)) * t2;
        return 1 / (Pi * ax * ay * cos4 * (1 + e) * (1 + e));
    }
    float lambda(const V3 &w) const {
        float abs_tan = std::fabs(tan_theta(w));
        if (std::isinf(abs_tan)) return 0.;
        float alpha = std::sqrt(cos2_phi(w) * ax * ax + sin2_phi(w) * ay * ay);
        float a2t2 = (alpha * abs_tan) * (alpha * abs_tan);
        return (-1 + std::sqrt(1.f + a2t2)) / 2;
    }
    float G1(const V3 &w) const { return 1 / (1 + lambda(w)); }
    float G(const V3 &wo, const V3 &wi) const { return 1 / (1 + lambda(wo) + lambda(wi)); }
    float pdf(const V3 &wo, const V3 &wh) const { return D(wh) * G1(wo) * std::fabs(dot(wo, wh)) / abs_cos_theta(wo); }
    // Visible-normal sample (divergence 32: the reference's sample_visible_area
    // branch runs the classic sampler). The reference's visible sampler
    // (trowbridge_reitz_distribution.rs:111-182, pbrt-v3's sample11) inverts
    // the y-slope CDF with a rational fit that never reaches |slope| > 7.3 and
    // misplaces tail mass while the estimator divides by the exact pdf; this
    // samples the distribution of visible normals exactly (Heitz 2018,
    // "Sampling the GGX Distribution of Visible Normals", JCGT 7(4)): stretch
    // wo to the unit-roughness configuration, sample the projected area of the
    // hemisphere as a disk warped towards the visible half, lift it onto the
    // hemisphere and unstretch.
    V3 sample_wh(const V3 &wo, const float *u) const {
        bool flip = wo.z < 0;
        V3 w = flip ? -wo : wo;
        V3 vh = normalize(V3(ax * w.x, ay * w.y, w.z));
        float lensq = vh.x * vh.x + vh.y * vh.y;
        V3 t1 = lensq > 0 ? V3(-vh.y, vh.x, 0) * (1 / std::sqrt(lensq)) : V3(1, 0, 0);
        V3 t2 = cross(vh, t1);
        float r = std::sqrt(u[0]);
        float phi = 2 * Pi * u[1];
        float p1 = r * std::cos(phi), p2 = r * std::sin(phi);
        float s = 0.5f * (1 + vh.z);
        p2 = (1 - s) * std::sqrt(std::max(0.f, 1 - p1 * p1)) + s * p2;
        V3 nh = p1 * t1 + p2 * t2 + std::sqrt(std::max(0.f, 1 - p1 * p1 - p2 * p2)) * vh;
        V3 wh = normalize(V3(ax * nh.x, ay * nh.y, std::max(0.f, nh.z)));
        return flip ? -wh : wh;
    }
};

enum { BSDF_REFLECTION = 1, BSDF_TRANSMISSION = 2, BSDF_DIFFUSE = 4, BSDF_GLOSSY = 8, BSDF_SPECULAR = 16,
       BSDF_ALL = 31 };

// reflection/mod.rs:50-275 with one BxDF per material (Matte -> Lambertian,
// Metal -> MicrofacetReflection(TR, FresnelConductor), Glass ->
// FresnelSpecular); lobe selection per pbrt-v3 (divergence 28).
struct BSDF {
    const Material *m = nullptr;
    V3 ng, ns, ss, ts;
    float eta = 1;
    int type = 0;  // BxDF type bits, 0 = no lobe
    void init(const SurfaceInteraction &si, const Material &mat) {
        m = &mat;
        ns = si.shading_n;
        ng = si.n;
        ss = normalize(si.shading_dpdu);
        ts = cross(ns, ss);
        if (mat.type == MAT_MATTE) type = is_black(mat.kd) ? 0 : (BSDF_REFLECTION | BSDF_DIFFUSE);
        else if (mat.type == MAT_METAL) type = BSDF_REFLECTION | BSDF_GLOSSY;
        else {
            type = BSDF_REFLECTION | BSDF_TRANSMISSION | BSDF_SPECULAR;
            eta = mat.ior;
        }
    }
    V3 w2l(const V3 &v) const { return V3(dot(v, ss), dot(v, ts), dot(v, ns)); }
    V3 l2w(const V3 &v) const {
        return V3(ss.x * v.x + ts.x * v.y + ns.x * v.z, ss.y * v.x + ts.y * v.y + ns.y * v.z,
                  ss.z * v.x + ts.z * v.y + ns.z * v.z);
    }
    bool matches(int flags) const { return type != 0 && (type & flags) == type; }
    int num_components(int flags) const { return matches(flags) ? 1 : 0; }
    TR tr() const { return TR{m->alpha, m->alpha}; }
    // BxDF::f in the local frame
    RGB bxdf_f(const V3 &wo, const V3 &wi) const {
        if (m->type == MAT_MATTE) return m->kd * InvPi;  // lambertian_reflection.rs:25-31
        if (m->type == MAT_METAL) {                       // microfacet_reflection.rs:29-51
            float cto = abs_cos_theta(wo), cti = abs_cos_theta(wi);
            V3 wh = wi + wo;
            if (cti == 0 || cto == 0) return RGB();
            if (wh.x == 0 && wh.y == 0 && wh.z == 0) return RGB();
            wh = normalize(wh);
            RGB F = fr_conductor(dot(wi, faceforward(wh, V3(0, 0, 1))), m->eta, m->k);
            TR d = tr();
            return cmul(RGB(1, 1, 1) * d.D(wh) * d.G(wo, wi), F) / (4 * cti * cto);
        }
        return RGB();  // specular: f = 0
    }
    float bxdf_pdf(const V3 &wo, const V3 &wi) const {
        if (m->type == MAT_MATTE) return same_hemisphere(wo, wi) ? abs_cos_theta(wi) * InvPi : 0;  // mod.rs:357-359
        if (m->type == MAT_METAL) {  // microfacet_reflection.rs:79-86 (divergence 29 fixed)
            if (!same_hemisphere(wo, wi)) return 0;
            V3 wh = normalize(wo + wi);
            return tr().pdf(wo, wh) / (4 * dot(wo, wh));
        }
        return 0;
    }
    RGB bxdf_sample_f(const V3 &wo, V3 *wi, const float *u, float *pdf, int *sampled) const {
        if (m->type == MAT_MATTE) {  // mod.rs:300-315
            *wi = cosine_sample_hemisphere(u);
            if (wo.z < 0) wi->z *= -1;
            *pdf = bxdf_pdf(wo, *wi);
            return bxdf_f(wo, *wi);
        }
        if (m->type == MAT_METAL) {  // microfacet_reflection.rs:54-77
            if (wo.z == 0) return RGB();
            V3 wh = tr().sample_wh(wo, u);
            if (dot(wo, wh) < 0) return RGB();
            *wi = reflect(wo, wh);
            if (!same_hemisphere(wo, *wi)) return RGB();
            *pdf = tr().pdf(wo, wh) / (4 * dot(wo, wh));
            return bxdf_f(wo, *wi);
        }
        // fresnel_specular.rs:42-81 (divergence 33 fixed), Kr = Kt = 1, radiance transport
        float F = fr_dielectric(cos_theta(wo), 1.f, m->ior);
        if (u[0] < F) {
            *wi = V3(-wo.x, -wo.y, wo.z);
            *sampled = BSDF_SPECULAR | BSDF_REFLECTION;
            *pdf = F;
            return RGB(F, F, F) / abs_cos_theta(*wi);
        }
        bool entering = cos_theta(wo) > 0;
        float eta_i = entering ? 1.f : m->ior;
        float eta_t = entering ? m->ior : 1.f;
        if (!refract(wo, faceforward(V3(0, 0, 1), wo), eta_i / eta_t, wi)) return RGB();
        float ft = 1 * (1 - F);
        ft *= (eta_i * eta_i) / (eta_t * eta_t);
        *sampled = BSDF_SPECULAR | BSDF_TRANSMISSION;
        *pdf = 1 - F;
        return RGB(ft, ft, ft) / abs_cos_theta(*wi);
    }
    // reflection/mod.rs:109-131
    RGB f(const V3 &wo_w, const V3 &wi_w, int flags) const {
        V3 wi = w2l(wi_w), wo = w2l(wo_w);
        if (wo.z == 0) return RGB();
        bool refl = dot(wi_w, ng) * dot(wo_w, ng) > 0;
        RGB r;
        if (matches(flags) && ((refl && (type & BSDF_REFLECTION)) || (!refl && (type & BSDF_TRANSMISSION))))
            r += bxdf_f(wo, wi);
        return r;
    }
    // reflection/mod.rs:164-244
    RGB sample_f(const V3 &wo_w, V3 *wi_w, const float *u, float *pdf, int flags, int *sampled) const {
        int matching = num_components(flags);
        if (matching == 0) {
            *pdf = 0;
            *sampled = 0;
            return RGB();
        }
        int comp = std::min((int)std::floor(u[0] * matching), matching - 1);
        float ur[2] = {std::min(u[0] * matching - comp, OneMinusEpsilon), u[1]};
        V3 wi, wo = w2l(wo_w);
        if (wo.z == 0) return RGB();
        *pdf = 0;
        *sampled = type;
        RGB fv = bxdf_sample_f(wo, &wi, ur, pdf, sampled);
        if (*pdf == 0) {
            *sampled = 0;
            return RGB();
        }
        *wi_w = l2w(wi);
        if (!(type & BSDF_SPECULAR)) {
            bool refl = dot(*wi_w, ng) * dot(wo_w, ng) > 0;
            fv = RGB();
            if ((refl && (type & BSDF_REFLECTION)) || (!refl && (type & BSDF_TRANSMISSION))) fv += bxdf_f(wo, wi);
        }
        return fv;
    }
    // reflection/mod.rs:246-274
    float pdf(const V3 &wo_w, const V3 &wi_w, int flags) const {
        if (type == 0) return 0;
        V3 wo = w2l(wo_w), wi = w2l(wi_w);
        if (wo.z == 0) return 0.;
        if (!matches(flags)) return 0;
        return bxdf_pdf(wo, wi) / 1;
    }
};

// ---------------------------------------------------------------- camera
void Camera::init(const V3 &pos, const V3 &look, const V3 &up, float fov, int w, int h) {
    width = w;
    height = h;
    camera_to_world = transform_inverse(transform_look_at(pos, look, up));
    float aspect = (float)w / (float)h;
    float sx0, sx1, sy0, sy1;
    if (aspect > 1.f) {
        sx0 = -aspect; sx1 = aspect; sy0 = -1.f; sy1 = 1.f;
    } else {
        sx0 = -1.f; sx1 = 1.f; sy0 = -1.f / aspect; sy1 = 1.f / aspect;
    }
    // perspective.rs:46-53
    Transform screen_to_raster = transform_compose(
        transform_compose(transform_scale((float)w, (float)h, 1), transform_scale(1 / (sx1 - sx0), 1 / (sy0 - sy1), 1)),
        transform_translate(V3(-sx0, -sy1, 0)));
    Transform raster_to_screen = transform_inverse(screen_to_raster);
    Transform camera_to_screen = transform_perspective(fov, 1e-2f, 1000.f);
    raster_to_camera = transform_compose(transform_inverse(camera_to_screen), raster_to_screen);
}

// perspective.rs:103-126 (lens radius 0) and Transform applied to a ray
// (transform.rs:365-389: origin advanced past its error bound, t_max reduced).
Ray Camera::generate_ray(float fx, float fy, float time) const {
    V3 p_camera = xform_point(raster_to_camera, V3(fx, fy, 0));
    Ray r(V3(0, 0, 0), normalize(p_camera));
    r.time = time;
    V3 o_err;
    V3 o = xform_point_with_error(camera_to_world, r.o, &o_err);
    V3 d = xform_vector(camera_to_world, r.d);
    float ls = length_squared(d);
    float t_max = r.t_max;
    if (ls > 0) {
        float dt = dot(vabs(d), o_err) / ls;
        o += d * dt;
        t_max -= dt;
    }
    return Ray(o, d, t_max, r.time);
}

// ---------------------------------------------------------------- lights
// DiffuseAreaLight::L (PBR book 3e 12.5): one-sided emission.
static RGB area_L(const AreaLight &l, const V3 &n, const V3 &w) { return dot(n, w) > 0 ? l.L : RGB(); }

// Light::sample_li (light.rs:29-36) for a DiffuseAreaLight on a triangle:
// Shape::sample_with_point (shape.rs:43-57) converts the area pdf to solid
// angle.
static RGB sample_li(const AreaLight &l, const SurfaceInteraction &ref, const float *u, V3 *wi, float *pdf, V3 *lp,
                     V3 *lp_err, V3 *ln) {
    triangle_sample(l.p0, l.p1, l.p2, false, u[0], u[1], lp, ln, lp_err);
    *pdf = 1 / l.area;
    V3 w = *lp - ref.p;
    if (length_squared(w) == 0)
        *pdf = 0;
    else {
        w = normalize(w);
        *pdf *= length_squared(ref.p - *lp) / abs_dot(*ln, -w);
        if (std::isinf(*pdf)) *pdf = 0.f;
    }
    if (*pdf == 0 || length_squared(*lp - ref.p) == 0) {
        *pdf = 0;
        return RGB();
    }
    *wi = normalize(*lp - ref.p);
    return area_L(l, *ln, -*wi);
}

// Shape::pdf_with_point (shape.rs:60-81, divergence 36 fixed): intersect the
// light's own shape only.
static float pdf_li(const AreaLight &l, const SurfaceInteraction &ref, const V3 &wi) {
    Ray ray = ref.spawn_ray(wi);
    SurfaceInteraction isl;
    float t_hit;
    if (!triangle_intersect(l.p0, l.p1, l.p2, false, ray, &isl, &t_hit)) return 0;
    float pdf = length_squared(ref.p - isl.p) / (abs_dot(isl.n, -wi) * l.area);
    if (std::isinf(pdf)) pdf = 0.f;
    return pdf;
}

RGB Sky::L(const V3 &w_in) const {
    const V3 w = normalize(w_in);
    if (!(w.y > 0)) return RGB();
    const float t = std::sqrt(w.y);
    return zenith * t + horizon * (1 - t);
}

// Sky sample: cosine-weighted direction about +y (the local frame's z is the
// world's y); pdf = w.y / pi.
static RGB sky_sample_li(const Sky &sky, const float *u, V3 *wi, float *pdf) {
    const V3 d = cosine_sample_hemisphere(u);
    *wi = V3(d.x, d.z, d.y);
    *pdf = wi->y > 0 ? wi->y * InvPi : 0.f;
    return *pdf > 0 ? sky.L(*wi) : RGB();
}
static float sky_pdf_li(const V3 &wi) { return wi.y > 0 ? wi.y * InvPi : 0.f; }

// Scene::intersect with the surface interaction of the closest hit.
static bool scene_intersect(const RenderScene &sc, Ray &ray, SurfaceInteraction *si, RenderStats *st) {
    if (st) st->closest_rays++;
    uint32_t p = sc.bvh.intersect(ray, si);
    return p != 0xffffffffu;
}
static bool scene_intersect_p(const RenderScene &sc, const Ray &ray, RenderStats *st) {
    if (st) st->any_rays++;
    return sc.bvh.intersect_p(ray);
}

// integrator.rs:353-494 estimate_direct with pbrt-v3's BSDF-sampled MIS term
// (divergence 34).
static RGB estimate_direct_sky(const RenderScene &sc, const SurfaceInteraction &it, const BSDF &bsdf,
                               const float *u_scat, const float *u_light, RenderStats *st);

static RGB estimate_direct(const RenderScene &sc, const SurfaceInteraction &it, const BSDF &bsdf, const float *u_scat,
                           int light_idx, const float *u_light, RenderStats *st) {
    if (light_idx == (int)sc.lights.size()) return estimate_direct_sky(sc, it, bsdf, u_scat, u_light, st);
    const AreaLight &light = sc.lights[light_idx];
    int bsdf_flags = BSDF_ALL & ~BSDF_SPECULAR;
    RGB ld;
    V3 wi, lp, lp_err, ln;
    float light_pdf = 0, scattering_pdf = 0;
    RGB li = sample_li(light, it, u_light, &wi, &light_pdf, &lp, &lp_err, &ln);
    if (light_pdf > 0 && !is_black(li)) {
        RGB f = bsdf.f(it.wo, wi, bsdf_flags) * abs_dot(wi, it.shading_n);
        scattering_pdf = bsdf.pdf(it.wo, wi, bsdf_flags);
        if (!is_black(f)) {
            // VisibilityTester::unoccluded (light.rs:90-92)
            if (scene_intersect_p(sc, it.spawn_ray_to(lp, lp_err, ln), st)) li = RGB();
            if (!is_black(li)) {
                float weight = power_heuristic(1, light_pdf, 1, scattering_pdf);
                ld += cmul(f, li) * weight / light_pdf;
            }
        }
    }
    {
        int sampled = 0;
        RGB f = bsdf.sample_f(it.wo, &wi, u_scat, &scattering_pdf, bsdf_flags, &sampled);
        f *= abs_dot(wi, it.shading_n);
        bool sampled_specular = (sampled & BSDF_SPECULAR) != 0;
        if (!is_black(f) && scattering_pdf > 0) {
            float weight = 1;
            if (!sampled_specular) {
                light_pdf = pdf_li(light, it, wi);
                if (light_pdf == 0) return ld;
                weight = power_heuristic(1, scattering_pdf, 1, light_pdf);
            }
            SurfaceInteraction lisect;
            Ray ray = it.spawn_ray(wi);
            bool found = scene_intersect(sc, ray, &lisect, st);
            RGB li2;
            if (found && sc.prim_light[lisect.prim] == light_idx) li2 = area_L(light, lisect.n, -wi);
            if (!is_black(li2)) ld += cmul(f, li2) * weight / scattering_pdf;
        }
    }
    return ld;
}

// estimate_direct for the sky: the light sample's shadow ray leaves the scene
// (any hit over (0, inf)); the BSDF sample counts only if its ray escapes.
static RGB estimate_direct_sky(const RenderScene &sc, const SurfaceInteraction &it, const BSDF &bsdf,
                               const float *u_scat, const float *u_light, RenderStats *st) {
    int bsdf_flags = BSDF_ALL & ~BSDF_SPECULAR;
    RGB ld;
    V3 wi;
    float light_pdf = 0, scattering_pdf = 0;
    RGB li = sky_sample_li(sc.sky, u_light, &wi, &light_pdf);
    if (light_pdf > 0 && !is_black(li)) {
        RGB f = bsdf.f(it.wo, wi, bsdf_flags) * abs_dot(wi, it.shading_n);
        scattering_pdf = bsdf.pdf(it.wo, wi, bsdf_flags);
        if (!is_black(f)) {
            if (scene_intersect_p(sc, it.spawn_ray(wi), st)) li = RGB();
            if (!is_black(li)) {
                float weight = power_heuristic(1, light_pdf, 1, scattering_pdf);
                ld += cmul(f, li) * weight / light_pdf;
            }
        }
    }
    {
        int sampled = 0;
        RGB f = bsdf.sample_f(it.wo, &wi, u_scat, &scattering_pdf, bsdf_flags, &sampled);
        f *= abs_dot(wi, it.shading_n);
        bool sampled_specular = (sampled & BSDF_SPECULAR) != 0;
        if (!is_black(f) && scattering_pdf > 0) {
            float weight = 1;
            if (!sampled_specular) {
                light_pdf = sky_pdf_li(wi);
                if (light_pdf == 0) return ld;
                weight = power_heuristic(1, scattering_pdf, 1, light_pdf);
            }
            SurfaceInteraction lisect;
            Ray ray = it.spawn_ray(wi);
            bool found = scene_intersect(sc, ray, &lisect, st);
            if (!found) {
                RGB li2 = sc.sky.L(wi);
                if (!is_black(li2)) ld += cmul(f, li2) * weight / scattering_pdf;
            }
        }
    }
    return ld;
}

// integrator.rs:321-350 uniform_sample_one_light with the power distribution.
static RGB sample_one_light(const RenderScene &sc, const SurfaceInteraction &it, const BSDF &bsdf, Sampler &sampler,
                            RenderStats *st) {
    int n = sc.n_lights();
    if (n == 0) return RGB();
    float light_pdf;
    int light_num = sc.light_distrib.sample_discrete(sampler.get_1d(), &light_pdf);
    if (light_pdf == 0) return RGB();
    float u_light[2], u_scat[2];
    sampler.get_2d(u_light);
    sampler.get_2d(u_scat);
    return estimate_direct(sc, it, bsdf, u_scat, light_num, u_light, st) / light_pdf;
}

// PathIntegrator::li (PBR book 3e 14.5.4; the reference has only the trait,
// integrator.rs:31-35). Russian roulette after 3 bounces with threshold 1.
RGB path_li(const RenderScene &sc, Ray ray, Sampler &sampler, int max_depth, RenderStats *st) {
    RGB L, beta(1, 1, 1);
    bool specular_bounce = false;
    float eta_scale = 1;
    for (int bounces = 0;; ++bounces) {
        SurfaceInteraction isect;
        bool found = scene_intersect(sc, ray, &isect, st);
        if (bounces == 0 || specular_bounce) {
            if (found) {
                int li = sc.prim_light[isect.prim];
                if (li >= 0) L += cmul(beta, area_L(sc.lights[li], isect.n, -ray.d));
            } else if (sc.sky.on) {
                L += cmul(beta, sc.sky.L(ray.d));  // InfiniteAreaLight::Le of the escaped ray
            }
        }
        if (!found || bounces >= max_depth) break;
        const Material &mat = sc.materials[sc.prim_material[isect.prim]];
        BSDF bsdf;  // every primitive has a material, so the BSDF always exists
        bsdf.init(isect, mat);
        if (bsdf.num_components(BSDF_ALL & ~BSDF_SPECULAR) > 0) L += cmul(beta, sample_one_light(sc, isect, bsdf, sampler, st));
        V3 wo = -ray.d, wi;
        float pdf;
        int flags = 0;
        float u[2];
        sampler.get_2d(u);
        RGB f = bsdf.sample_f(wo, &wi, u, &pdf, BSDF_ALL, &flags);
        if (is_black(f) || pdf == 0.f) break;
        beta = cmul(beta, f) * abs_dot(wi, isect.shading_n) / pdf;
        specular_bounce = (flags & BSDF_SPECULAR) != 0;
        if ((flags & BSDF_SPECULAR) && (flags & BSDF_TRANSMISSION)) {
            float eta = bsdf.eta;
            eta_scale *= (dot(wo, isect.n) > 0) ? (eta * eta) : 1 / (eta * eta);
        }
        ray = isect.spawn_ray(wi);
        RGB rr_beta = beta * eta_scale;
        if (max_comp(rr_beta) < 1.f && bounces > 3) {
            float q = std::max(0.05f, 1 - max_comp(rr_beta));
            if (sampler.get_1d() < q) break;
            beta = beta / (1 - q);
        }
    }
    return L;
}

// AOIntegrator::li (integrators/ao.rs:45-93): at the camera ray's closest
// hit, n = the geometric normal facing the ray's origin, s = normalize(dpdu),
// t = n_geometric x s; each of n_samples hemisphere directions (cosine- or
// uniform-distributed, 2 sampler dimensions each, in order) adds
// (wi . n) / (pdf * n_samples) when its shadow ray is unoccluded. Every
// primitive here has a material, so the skip of BSDF-less surfaces (58-62)
// never applies.
RGB ao_li(const RenderScene &sc, Ray ray, Sampler &sampler, int n_samples, bool cos_sample, RenderStats *st) {
    SurfaceInteraction isect;
    if (!scene_intersect(sc, ray, &isect, st)) return RGB();
    const V3 n = faceforward(isect.n, -ray.d);
    const V3 s = normalize(isect.dpdu);
    const V3 t = cross(isect.n, s);
    RGB L;
    for (int i = 0; i < n_samples; ++i) {
        float u[2];
        sampler.get_2d(u);
        V3 wi;
        float pdf;
        if (cos_sample) {
            wi = cosine_sample_hemisphere(u);
            pdf = std::fabs(wi.z) * InvPi;  // cosine_hemisphere_pdf
        } else {
            wi = uniform_sample_hemisphere(u);
            pdf = Inv2Pi;  // uniform_hemisphere_pdf
        }
        wi = V3(s.x * wi.x + t.x * wi.y + n.x * wi.z, s.y * wi.x + t.y * wi.y + n.y * wi.z,
                s.z * wi.x + t.z * wi.y + n.z * wi.z);
        if (!scene_intersect_p(sc, isect.spawn_ray(wi), st)) {
            const float v = dot(wi, n) / (pdf * (float)n_samples);
            L += RGB(v, v, v);
        }
    }
    return L;
}

}  // namespace orc

// ---------------------------------------------------------------- BSDF-level entry points
// Local-frame (shading normal = +z) evaluations of the BxDFs above, one call
// per batch of directions (3 floats each), for the closed-form and
// statistical tests of the lobes. Material record: type, kd[3], eta[3], k[3],
// alpha, ior (the layout of illume_material without its padding).
namespace {
struct MatIn {
    int32_t type;
    float kd[3], eta[3], k[3], alpha, ior;
};
orc::Material mat_of(const MatIn *in) {
    orc::Material m;
    m.type = in->type;
    m.kd = orc::RGB(in->kd[0], in->kd[1], in->kd[2]);
    m.eta = orc::RGB(in->eta[0], in->eta[1], in->eta[2]);
    m.k = orc::RGB(in->k[0], in->k[1], in->k[2]);
    m.alpha = in->alpha;
    m.ior = in->ior;
    return m;
}
orc::V3 v3(const float *p) { return orc::V3(p[0], p[1], p[2]); }
void put(float *p, const orc::V3 &v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
orc::BSDF local_bsdf(const orc::Material *m) {
    orc::BSDF b;
    b.m = m;
    if (m->type == orc::MAT_MATTE) b.type = orc::BSDF_REFLECTION | orc::BSDF_DIFFUSE;
    else if (m->type == orc::MAT_METAL) b.type = orc::BSDF_REFLECTION | orc::BSDF_GLOSSY;
    else b.type = orc::BSDF_REFLECTION | orc::BSDF_TRANSMISSION | orc::BSDF_SPECULAR;
    return b;
}
}  // namespace

extern "C" {

// Trowbridge-Reitz (isotropic alpha): D(wh), Lambda(w), G1(w), and the
// visible-normal pdf of wh given wo (D G1 |wo.wh| / |cos wo|).
void orc_tr_eval(float alpha, const float *wo, const float *wh, uint32_t n, float *d, float *lambda_wo,
                 float *g1_wo, float *pdf) {
    orc::TR tr{alpha, alpha};
    for (uint32_t i = 0; i < n; ++i) {
        const orc::V3 o = v3(wo + 3 * i), h = v3(wh + 3 * i);
        d[i] = tr.D(h);
        lambda_wo[i] = tr.lambda(o);
        g1_wo[i] = tr.G1(o);
        pdf[i] = tr.pdf(o, h);
    }
}

// TR height-correlated masking-shadowing G(wo, wi).
void orc_tr_g(float alpha, const float *wo, const float *wi, uint32_t n, float *g) {
    orc::TR tr{alpha, alpha};
    for (uint32_t i = 0; i < n; ++i) g[i] = tr.G(v3(wo + 3 * i), v3(wi + 3 * i));
}

// Visible-normal sample of the TR distribution: wh for wo and u (2 floats each).
void orc_tr_sample_wh(float alpha, const float *wo, const float *u, uint32_t n, float *wh) {
    orc::TR tr{alpha, alpha};
    for (uint32_t i = 0; i < n; ++i) put(wh + 3 * i, tr.sample_wh(v3(wo + 3 * i), u + 2 * i));
}

// BxDF::f (rgb), BxDF::pdf and BxDF::sample_f of one material in the local frame.
void orc_bxdf_f(const void *mat, const float *wo, const float *wi, uint32_t n, float *rgb) {
    const orc::Material m = mat_of((const MatIn *)mat);
    const orc::BSDF b = local_bsdf(&m);
    for (uint32_t i = 0; i < n; ++i) put(rgb + 3 * i, b.bxdf_f(v3(wo + 3 * i), v3(wi + 3 * i)));
}
void orc_bxdf_pdf(const void *mat, const float *wo, const float *wi, uint32_t n, float *pdf) {
    const orc::Material m = mat_of((const MatIn *)mat);
    const orc::BSDF b = local_bsdf(&m);
    for (uint32_t i = 0; i < n; ++i) pdf[i] = b.bxdf_pdf(v3(wo + 3 * i), v3(wi + 3 * i));
}
// wi (3), f (3), pdf and the sampled lobe's flags per sample; pdf 0 = no sample.
void orc_bxdf_sample_f(const void *mat, const float *wo, const float *u, uint32_t n, float *wi, float *rgb,
                       float *pdf, int32_t *sampled) {
    const orc::Material m = mat_of((const MatIn *)mat);
    const orc::BSDF b = local_bsdf(&m);
    for (uint32_t i = 0; i < n; ++i) {
        orc::V3 w;
        float p = 0.f;
        int s = b.type;
        const orc::RGB f = b.bxdf_sample_f(v3(wo + 3 * i), &w, u + 2 * i, &p, &s);
        put(wi + 3 * i, w);
        put(rgb + 3 * i, f);
        pdf[i] = p;
        sampled[i] = p > 0.f ? s : 0;
    }
}

// Fresnel terms and refraction (reflection/mod.rs:420-475, 547-560).
void orc_fr_dielectric(const float *cos_i, uint32_t n, float eta_i, float eta_t, float *out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = orc::fr_dielectric(cos_i[i], eta_i, eta_t);
}
void orc_fr_conductor(const float *cos_i, uint32_t n, const float *eta, const float *k, float *rgb) {
    for (uint32_t i = 0; i < n; ++i)
        put(rgb + 3 * i, orc::fr_conductor(cos_i[i], orc::RGB(eta[0], eta[1], eta[2]), orc::RGB(k[0], k[1], k[2])));
}
// wt and a success flag for each (wi, normal) pair at relative index eta = eta_i / eta_t.
void orc_refract(const float *wi, const float *nrm, uint32_t n, float eta, float *wt, int32_t *ok) {
    for (uint32_t i = 0; i < n; ++i) {
        orc::V3 t;
        ok[i] = orc::refract(v3(wi + 3 * i), v3(nrm + 3 * i), eta, &t) ? 1 : 0;
        put(wt + 3 * i, ok[i] ? t : orc::V3());
    }
}

}  // extern "C"
