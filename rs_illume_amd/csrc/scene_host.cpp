// Scene, light table and camera set-up on the host for the path tracer, and
// the illume_scene_* / illume_render entry points.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include "../../include/illume_hip.h"
#include "kernels.h"
#include "scene_types.h"

namespace illume {
int abi_fail(int status, const std::string &msg);
int abi_hip_fail(hipError_t e, const char *what);
void bvh_device_arrays(const illume_bvh *bvh, const GpuNode4Q **nodes, const PrimRecord **prims);
hipError_t scene_tri_verts(const PrimRecord *prims, uint32_t n_records, uint32_t n_tris, float4 *tv, hipStream_t st);
}  // namespace illume

using namespace illume;

namespace {

// 4x4 row-major matrices for the camera transforms (core/matrix.rs,
// core/transform.rs), composed and inverted as pbrt-v3 does.
struct M4 {
    float m[4][4];
};
M4 m_identity() {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = i == j ? 1.f : 0.f;
    return r;
}
M4 m_mul(const M4 &a, const M4 &b) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] + a.m[i][3] * b.m[3][j];
    return r;
}
// Gauss-Jordan elimination with full pivoting (matrix.rs:84-162).
bool m_inverse(const M4 &a, M4 *out) {
    int indxc[4], indxr[4], ipiv[4] = {0, 0, 0, 0};
    float minv[4][4];
    std::memcpy(minv, a.m, sizeof(minv));
    for (int i = 0; i < 4; i++) {
        int irow = 0, icol = 0;
        float big = 0.f;
        for (int j = 0; j < 4; j++) {
            if (ipiv[j] == 1) continue;
            for (int k = 0; k < 4; k++) {
                if (ipiv[k] == 0) {
                    if (std::fabs(minv[j][k]) >= big) {
                        big = std::fabs(minv[j][k]);
                        irow = j;
                        icol = k;
                    }
                } else if (ipiv[k] > 1) {
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
            if (j == icol) continue;
            float save = minv[j][icol];
            minv[j][icol] = 0;
            for (int k = 0; k < 4; k++) minv[j][k] -= minv[icol][k] * save;
        }
    }
    for (int j = 3; j >= 0; j--)
        if (indxr[j] != indxc[j])
            for (int k = 0; k < 4; k++) std::swap(minv[k][indxr[j]], minv[k][indxc[j]]);
    std::memcpy(out->m, minv, sizeof(minv));
    return true;
}
struct Xf {
    M4 m, mi;
};
Xf xf_compose(const Xf &a, const Xf &b) { return Xf{m_mul(a.m, b.m), m_mul(b.mi, a.mi)}; }
Xf xf_inverse(const Xf &a) { return Xf{a.mi, a.m}; }
Xf xf_scale(float x, float y, float z) {
    Xf t{m_identity(), m_identity()};
    t.m.m[0][0] = x; t.m.m[1][1] = y; t.m.m[2][2] = z;
    t.mi.m[0][0] = 1.f / x; t.mi.m[1][1] = 1.f / y; t.mi.m[2][2] = 1.f / z;
    return t;
}
Xf xf_translate(float x, float y, float z) {
    Xf t{m_identity(), m_identity()};
    t.m.m[0][3] = x; t.m.m[1][3] = y; t.m.m[2][3] = z;
    t.mi.m[0][3] = -x; t.mi.m[1][3] = -y; t.mi.m[2][3] = -z;
    return t;
}
struct V {
    float x, y, z;
};
V vsub(V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; }
V vnorm(V a) {
    float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return V{a.x / l, a.y / l, a.z / l};
}
V vcross(V a, V b) {  // evaluated in f64 (vector3.rs:66-82)
    double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
    return V{(float)((ay * bz) - (az * by)), (float)((az * bx) - (ax * bz)), (float)((ax * by) - (ay * bx))};
}
// transform.rs:181-208: world-to-camera of a look-at frame
Xf xf_look_at(V pos, V look, V up) {
    M4 c2w;
    std::memset(&c2w, 0, sizeof(c2w));
    c2w.m[0][3] = pos.x; c2w.m[1][3] = pos.y; c2w.m[2][3] = pos.z; c2w.m[3][3] = 1;
    V dir = vnorm(vsub(look, pos));
    V right = vnorm(vcross(vnorm(up), dir));
    V new_up = vcross(dir, right);
    c2w.m[0][0] = right.x; c2w.m[1][0] = right.y; c2w.m[2][0] = right.z;
    c2w.m[0][1] = new_up.x; c2w.m[1][1] = new_up.y; c2w.m[2][1] = new_up.z;
    c2w.m[0][2] = dir.x; c2w.m[1][2] = dir.y; c2w.m[2][2] = dir.z;
    Xf t;
    m_inverse(c2w, &t.m);
    t.mi = c2w;
    return t;
}
// pbrt-v3 Perspective(fov, n, f) = Scale(1/tan(fov/2), 1/tan(fov/2), 1) * persp
Xf xf_perspective(float fov, float n, float f) {
    M4 p = m_identity();
    p.m[2][2] = f / (f - n);
    p.m[2][3] = -f * n / (f - n);
    p.m[3][2] = 1;
    p.m[3][3] = 0;
    Xf pt;
    pt.m = p;
    m_inverse(p, &pt.mi);
    const float pi = 3.14159265358979323846f;
    float inv_tan = 1.f / std::tan((pi / 180.f * fov) / 2.f);
    return xf_compose(xf_scale(inv_tan, inv_tan, 1), pt);
}

bool make_camera(const illume_camera &c, DevCamera *out) {
    if (c.width == 0 || c.height == 0 || !(c.fov > 0.f && c.fov < 180.f)) return false;
    Xf c2w = xf_inverse(xf_look_at(V{c.pos[0], c.pos[1], c.pos[2]}, V{c.look[0], c.look[1], c.look[2]},
                                   V{c.up[0], c.up[1], c.up[2]}));
    float aspect = (float)c.width / (float)c.height;
    float sx0, sx1, sy0, sy1;
    if (aspect > 1.f) {
        sx0 = -aspect; sx1 = aspect; sy0 = -1.f; sy1 = 1.f;
    } else {
        sx0 = -1.f; sx1 = 1.f; sy0 = -1.f / aspect; sy1 = 1.f / aspect;
    }
    // perspective.rs:46-53
    Xf screen_to_raster = xf_compose(
        xf_compose(xf_scale((float)c.width, (float)c.height, 1), xf_scale(1 / (sx1 - sx0), 1 / (sy0 - sy1), 1)),
        xf_translate(-sx0, -sy1, 0));
    Xf raster_to_screen = xf_inverse(screen_to_raster);
    Xf camera_to_screen = xf_perspective(c.fov, 1e-2f, 1000.f);
    Xf raster_to_camera = xf_compose(xf_inverse(camera_to_screen), raster_to_screen);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out->raster_to_camera[4 * i + j] = raster_to_camera.m.m[i][j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j) out->camera_to_world[4 * i + j] = c2w.m.m[i][j];
    out->width = c.width;
    out->height = c.height;
    return true;
}

// Queues the copy on st: src must outlive it (the caller synchronises st).
template <typename T>
hipError_t upload(T **dst, const std::vector<T> &src, hipStream_t st) {
    *dst = nullptr;
    if (src.empty()) return hipSuccess;
    hipError_t e = hipMalloc(dst, src.size() * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, st);
}

// A stream of the caller's own for one scene's set-up (default priority, not
// synchronising with the null stream), so that scenes created from several
// host threads build their trees side by side.
struct SetupStream {
    hipStream_t st = nullptr;
    SetupStream() {
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = nullptr;
    }
    ~SetupStream() {
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    }
};

}  // namespace

struct illume_scene {
    illume_bvh *bvh = nullptr;
    DevScene dev;
    DevScene *d_dev = nullptr;  // the same record in device memory (job tables point at it)
    std::vector<void *> allocs;
    uint32_t n_prims = 0;
};

namespace illume {
// Scenes with at least this many primitives render through the batch kernel.
constexpr uint32_t kLargeScenePrims = 1u << 20;
const DevScene *scene_dev(const illume_scene *s) { return &s->dev; }
const DevScene *scene_dev_device(const illume_scene *s) { return s->d_dev; }
uint32_t scene_prims(const illume_scene *s) { return s->n_prims; }
void scene_bounds(const illume_scene *s, float *b) {
    illume_bvh_info_t info;
    illume_bvh_info(s->bvh, &info);
    std::memcpy(b, info.bounds, sizeof(info.bounds));
}
bool scene_camera(const illume_camera &c, DevCamera *out);
void filter_splat(const illume_filter *f, uint32_t width, uint32_t height, FilmSplat *out);
uint32_t sampler_kind(uint32_t flags) {
    if (flags == 0) return 0;
    if (flags == ILLUME_RENDER_HALTON) return kSamplerHalton;
    if (flags == ILLUME_RENDER_SOBOL) return kSamplerSobol;
    return ~0u;
}
}  // namespace illume

extern "C" {

int illume_scene_destroy(illume_scene *s) {
    if (!s) return ILLUME_OK;
    for (void *p : s->allocs)
        if (p) hipFree(p);
    illume_bvh_destroy(s->bvh);
    delete s;
    return ILLUME_OK;
}

int illume_scene_create(const float *verts, uint32_t n_verts, const uint32_t *tri_idx, uint32_t n_tris,
                        const float *spheres, uint32_t n_spheres, const uint32_t *prim_material,
                        const float *prim_emission, const illume_material *materials, uint32_t n_materials,
                        illume_scene **out) {
    return illume_scene_create_ex(verts, n_verts, tri_idx, n_tris, spheres, n_spheres, prim_material, prim_emission,
                                  materials, n_materials, nullptr, out);
}

int illume_scene_create_ex(const float *verts, uint32_t n_verts, const uint32_t *tri_idx, uint32_t n_tris,
                           const float *spheres, uint32_t n_spheres, const uint32_t *prim_material,
                           const float *prim_emission, const illume_material *materials, uint32_t n_materials,
                           const illume_scene_opts *opts, illume_scene **out) {
    if (!out) return abi_fail(ILLUME_ERR_INVALID_ARG, "out is null");
    const illume_sky *sky = opts ? opts->sky : nullptr;
    const int builder = opts ? opts->bvh_builder : 0;
    if (builder < 0 || builder > 2) return abi_fail(ILLUME_ERR_UNSUPPORTED, "bvh_builder must be 0, 1 or 2");
    if (sky)
        for (int c = 0; c < 3; ++c)
            if (!(sky->zenith[c] >= 0.f && sky->horizon[c] >= 0.f && std::isfinite(sky->zenith[c]) &&
                  std::isfinite(sky->horizon[c])))
                return abi_fail(ILLUME_ERR_INVALID_ARG, "sky radiance must be finite and >= 0");
    *out = nullptr;
    uint64_t np = (uint64_t)n_tris + n_spheres;
    if (np && (!prim_material || !prim_emission)) return abi_fail(ILLUME_ERR_INVALID_ARG, "null per-primitive array");
    if (n_materials == 0 || !materials) return abi_fail(ILLUME_ERR_INVALID_ARG, "no materials");
    for (uint64_t p = 0; p < np; ++p)
        if (prim_material[p] >= n_materials)
            return abi_fail(ILLUME_ERR_INVALID_ARG, "primitive " + std::to_string(p) + " has no valid material");
    for (uint32_t i = 0; i < n_materials; ++i) {
        const illume_material &m = materials[i];
        if (m.type < ILLUME_MATTE || m.type > ILLUME_GLASS)
            return abi_fail(ILLUME_ERR_INVALID_ARG, "material " + std::to_string(i) + " has an unknown type");
        if (m.type == ILLUME_METAL && !(m.alpha > 0.f)) return abi_fail(ILLUME_ERR_INVALID_ARG, "metal alpha must be > 0");
        if (m.type == ILLUME_GLASS && !(m.ior > 0.f)) return abi_fail(ILLUME_ERR_INVALID_ARG, "glass ior must be > 0");
    }
    for (uint32_t s = 0; s < n_spheres; ++s) {
        const float *e = prim_emission + 3 * ((size_t)n_tris + s);
        if (e[0] != 0.f || e[1] != 0.f || e[2] != 0.f)
            return abi_fail(ILLUME_ERR_UNSUPPORTED, "only triangles may emit");
    }
    SetupStream ss;
    if (!ss.st) return abi_fail(ILLUME_ERR_HIP, "stream creation failed");
    illume_scene *s = new (std::nothrow) illume_scene();
    if (!s) return abi_fail(ILLUME_ERR_OUT_OF_MEMORY, "host allocation failed");
    std::memset(&s->dev, 0, sizeof(s->dev));
    int st = builder == 1   ? illume_bvh_create_gpu(verts, n_verts, tri_idx, n_tris, spheres, n_spheres, 4, 4, ss.st,
                                                    &s->bvh)
             : builder == 2 ? illume_bvh_create_gpu_sah(verts, n_verts, tri_idx, n_tris, spheres, n_spheres, 4, 4,
                                                        ss.st, &s->bvh)
                            : illume_bvh_create_ex(verts, n_verts, tri_idx, n_tris, spheres, n_spheres, 4, 4, &s->bvh);
    if (st != ILLUME_OK) {
        delete s;
        return st;
    }
    s->n_prims = (uint32_t)np;
    bvh_device_arrays(s->bvh, &s->dev.nodes, &s->dev.prims);

    std::vector<float4> sv(n_spheres);
    for (uint32_t i = 0; i < n_spheres; ++i)
        sv[i] = make_float4(spheres[4 * i], spheres[4 * i + 1], spheres[4 * i + 2], spheres[4 * i + 3]);
    std::vector<uint32_t> pm(prim_material, prim_material + np);
    std::vector<int32_t> pl(np, -1);
    std::vector<DevMaterial> mats(n_materials);
    std::memcpy(mats.data(), materials, n_materials * sizeof(DevMaterial));
    std::vector<DevLight> lights;
    std::vector<float> power;
    for (uint32_t t = 0; t < n_tris; ++t) {
        const float *e = prim_emission + 3 * (size_t)t;
        if (e[0] == 0.f && e[1] == 0.f && e[2] == 0.f) continue;
        DevLight l;
        std::memset(&l, 0, sizeof(l));
        const float *pa = verts + 3 * (size_t)tri_idx[3 * (size_t)t], *pb = verts + 3 * (size_t)tri_idx[3 * (size_t)t + 1],
                    *pc = verts + 3 * (size_t)tri_idx[3 * (size_t)t + 2];
        const float4 a = make_float4(pa[0], pa[1], pa[2], 0.f), b = make_float4(pb[0], pb[1], pb[2], 0.f),
                     c = make_float4(pc[0], pc[1], pc[2], 0.f);
        l.p0[0] = a.x; l.p0[1] = a.y; l.p0[2] = a.z;
        l.p1[0] = b.x; l.p1[1] = b.y; l.p1[2] = b.z;
        l.p2[0] = c.x; l.p2[1] = c.y; l.p2[2] = c.z;
        l.L[0] = e[0]; l.L[1] = e[1]; l.L[2] = e[2];
        // Triangle::area (triangle.rs:609-614)
        V cr = vcross(V{b.x - a.x, b.y - a.y, b.z - a.z}, V{c.x - a.x, c.y - a.y, c.z - a.z});
        l.area = 0.5f * std::sqrt(cr.x * cr.x + cr.y * cr.y + cr.z * cr.z);
        l.prim = t;
        pl[t] = (int32_t)lights.size();
        lights.push_back(l);
        // DiffuseAreaLight power L * area * pi, by luminance (spectrum.rs:498-501)
        const float pi = 3.14159265358979323846f;
        float r = e[0] * l.area * pi, g = e[1] * l.area * pi, bb = e[2] * l.area * pi;
        power.push_back(0.212671f * r + 0.715160f * g + 0.072169f * bb);
    }
    int32_t sky_index = -1;
    if (sky) {
        // the sky's power: pi r^2 * sphere-averaged radiance ((2/3 zenith +
        // 1/3 horizon) / 2), r the radius of the scene's bounding sphere
        // bounds of the triangles' vertices and the spheres' centre +- radius
        float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (uint32_t t = 0; t < n_tris; ++t)
            for (int k = 0; k < 3; ++k) {
                const float *v = verts + 3 * (size_t)tri_idx[3 * (size_t)t + k];
                for (int a = 0; a < 3; ++a) b[a] = std::min(b[a], v[a]), b[3 + a] = std::max(b[3 + a], v[a]);
            }
        for (uint32_t i = 0; i < n_spheres; ++i)
            for (int a = 0; a < 3; ++a) {
                b[a] = std::min(b[a], spheres[4 * i + a] - spheres[4 * i + 3]);
                b[3 + a] = std::max(b[3 + a], spheres[4 * i + a] + spheres[4 * i + 3]);
            }
        const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
        const float r = np > 0 ? 0.5f * std::sqrt(dx * dx + dy * dy + dz * dz) : 0.f;
        const float pi = 3.14159265358979323846f;
        float avg[3];
        for (int c = 0; c < 3; ++c) avg[c] = (sky->zenith[c] * (2.f / 3.f) + sky->horizon[c] * (1.f / 3.f)) * 0.5f;
        sky_index = (int32_t)lights.size();
        power.push_back(0.212671f * (avg[0] * (pi * r * r)) + 0.715160f * (avg[1] * (pi * r * r)) +
                        0.072169f * (avg[2] * (pi * r * r)));
    }
    // Distribution1D (sampling.rs:20-42)
    std::vector<float> cdf(power.size() + 1, 0.f);
    float func_int = 0.f;
    if (!power.empty()) {
        size_t n = power.size();
        for (size_t i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + power[i - 1] / (float)n;
        func_int = cdf[n];
        if (func_int == 0)
            for (size_t i = 1; i < n + 1; ++i) cdf[i] = (float)i / (float)n;
        else
            for (size_t i = 1; i < n + 1; ++i) cdf[i] /= func_int;
    }
    hipError_t e = hipSuccess;
    float4 *d_tv = nullptr, *d_sv = nullptr;
    uint32_t *d_pm = nullptr;
    int32_t *d_pl = nullptr;
    DevMaterial *d_m = nullptr;
    DevLight *d_l = nullptr;
    float *d_f = nullptr, *d_c = nullptr;
    // the triangles' vertex table, built on the device from the tree's
    // primitive records (the same vertices, in mesh index order)
    if (n_tris) {
        if ((e = hipMalloc(&d_tv, 3 * (size_t)n_tris * sizeof(float4))) == hipSuccess) {
            s->allocs.push_back(d_tv);
            e = scene_tri_verts(s->dev.prims, (uint32_t)np, n_tris, d_tv, ss.st);
        }
    }
    if (e == hipSuccess && (e = upload(&d_sv, sv, ss.st)) == hipSuccess) s->allocs.push_back(d_sv);
    if (e == hipSuccess && (e = upload(&d_pm, pm, ss.st)) == hipSuccess) s->allocs.push_back(d_pm);
    if (e == hipSuccess && (e = upload(&d_pl, pl, ss.st)) == hipSuccess) s->allocs.push_back(d_pl);
    if (e == hipSuccess && (e = upload(&d_m, mats, ss.st)) == hipSuccess) s->allocs.push_back(d_m);
    if (e == hipSuccess && (e = upload(&d_l, lights, ss.st)) == hipSuccess) s->allocs.push_back(d_l);
    if (e == hipSuccess && (e = upload(&d_f, power, ss.st)) == hipSuccess) s->allocs.push_back(d_f);
    if (e == hipSuccess && (e = upload(&d_c, cdf, ss.st)) == hipSuccess) s->allocs.push_back(d_c);
    // the uploads read the host vectors above: wait for them here
    const hipError_t es = hipStreamSynchronize(ss.st);
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) {
        illume_scene_destroy(s);
        return abi_hip_fail(e, "scene upload");
    }
    s->dev.tri_verts = d_tv;
    s->dev.spheres = d_sv;
    s->dev.prim_material = d_pm;
    s->dev.prim_light = d_pl;
    s->dev.materials = d_m;
    s->dev.lights = d_l;
    s->dev.light_func = d_f;
    s->dev.light_cdf = d_c;
    s->dev.light_func_int = func_int;
    s->dev.n_lights = (uint32_t)power.size();
    s->dev.n_tris = n_tris;
    s->dev.sky_index = sky_index;
    if (sky)
        for (int c = 0; c < 3; ++c) {
            s->dev.sky_zenith[c] = sky->zenith[c];
            s->dev.sky_horizon[c] = sky->horizon[c];
        }
    e = hipMalloc(&s->d_dev, sizeof(DevScene));
    if (e == hipSuccess) {
        s->allocs.push_back(s->d_dev);
        e = hipMemcpyAsync(s->d_dev, &s->dev, sizeof(DevScene), hipMemcpyHostToDevice, ss.st);
        if (e == hipSuccess) e = hipStreamSynchronize(ss.st);
    }
    if (e != hipSuccess) {
        illume_scene_destroy(s);
        return abi_hip_fail(e, "scene upload");
    }
    *out = s;
    return ILLUME_OK;
}

}  // extern "C"

bool illume::scene_camera(const illume_camera &c, DevCamera *out) { return make_camera(c, out); }

extern "C" {

int illume_scene_bvh(const illume_scene *scene, const illume_bvh **out) {
    if (!scene || !out) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    *out = scene->bvh;
    return ILLUME_OK;
}

int illume_render(const illume_scene *scene, const illume_camera *camera, const illume_render_cfg *cfg,
                  float *rgb_sum, uint64_t *ray_counts, void *stream) {
    if (!scene || !camera || !cfg || !rgb_sum) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    DevCamera cam;
    if (!make_camera(*camera, &cam)) return abi_fail(ILLUME_ERR_INVALID_ARG, "invalid camera");
    if (cfg->spp == 0) return ILLUME_OK;
    if ((uint64_t)cfg->sample_offset + cfg->spp > (1u << 24))
        return abi_fail(ILLUME_ERR_UNSUPPORTED, "sample index beyond 2^24");
    if ((uint64_t)camera->width * camera->height >= (1ull << 40))
        return abi_fail(ILLUME_ERR_UNSUPPORTED, "image too large");
    const uint32_t kind = sampler_kind(cfg->flags);
    if (kind == ~0u) return abi_fail(ILLUME_ERR_INVALID_ARG, "unknown render flags");
    if (scene->n_prims == 0) return ILLUME_OK;  // nothing to hit, no lights: black
    SamplerLaunch hl{};
    if (kind && !sampler_launch(kind, 0, 0, (int)camera->width, (int)camera->height, &hl))
        return abi_fail(ILLUME_ERR_OUT_OF_MEMORY, "low-discrepancy tables");
    if (scene->n_prims >= kLargeScenePrims) {
        // large trees: the batch kernel's 5 waves per SIMD hide more of the node
        // fetches than the single-scene kernel's 3 (2M-triangle terrain view
        // 4.83 -> 4.46 ms); the paths and their sums are the same
        illume_job job;
        std::memset(&job, 0, sizeof(job));
        job.scene = scene;
        job.camera = *camera;
        job.cfg = *cfg;
        job.film_offset = 0;
        return illume_render_batch(&job, 1, rgb_sum, ray_counts, stream);
    }
    hipError_t e = launch_render(scene->dev, cam, cfg->spp, cfg->max_depth, cfg->seed, cfg->sample_offset, rgb_sum,
                                 (unsigned long long *)ray_counts, hl, (hipStream_t)stream);
    if (e != hipSuccess) return abi_hip_fail(e, "render launch");
    return ILLUME_OK;
}

int illume_render_ao(const illume_scene *scene, const illume_camera *camera, const illume_render_cfg *cfg,
                     uint32_t n_ao_samples, int32_t cos_sample, float *rgb_sum, uint64_t *ray_counts, void *stream) {
    if (!scene || !camera || !cfg || !rgb_sum) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    DevCamera cam;
    if (!make_camera(*camera, &cam)) return abi_fail(ILLUME_ERR_INVALID_ARG, "invalid camera");
    if (cfg->spp == 0) return ILLUME_OK;
    if ((uint64_t)cfg->sample_offset + cfg->spp > (1u << 24))
        return abi_fail(ILLUME_ERR_UNSUPPORTED, "sample index beyond 2^24");
    if ((uint64_t)camera->width * camera->height >= (1ull << 40))
        return abi_fail(ILLUME_ERR_UNSUPPORTED, "image too large");
    if (n_ao_samples > (1u << 16)) return abi_fail(ILLUME_ERR_UNSUPPORTED, "more than 65536 AO samples");
    const uint32_t kind = sampler_kind(cfg->flags);
    if (kind == ~0u) return abi_fail(ILLUME_ERR_INVALID_ARG, "unknown render flags");
    if (scene->n_prims == 0 || n_ao_samples == 0) return ILLUME_OK;  // nothing hit, or no directions: 0
    SamplerLaunch hl{};
    if (kind && !sampler_launch(kind, 0, 0, (int)camera->width, (int)camera->height, &hl))
        return abi_fail(ILLUME_ERR_OUT_OF_MEMORY, "low-discrepancy tables");
    hipError_t e = launch_render_ao(scene->dev, cam, cfg->spp, n_ao_samples, cos_sample != 0, cfg->seed,
                                    cfg->sample_offset, rgb_sum, (unsigned long long *)ray_counts, hl,
                                    (hipStream_t)stream);
    if (e != hipSuccess) return abi_hip_fail(e, "AO render launch");
    return ILLUME_OK;
}

int illume_li_batch(const illume_scene *scene, const illume_ray *rays, uint32_t n, const illume_render_cfg *cfg,
                    float *L_out, uint64_t *ray_counts, void *stream) {
    if (!scene || !cfg) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    if (cfg->flags != 0)
        return abi_fail(ILLUME_ERR_UNSUPPORTED, "illume_li_batch takes the random sampler streams only (flags 0)");
    if (cfg->sample_offset >= (1u << 24)) return abi_fail(ILLUME_ERR_UNSUPPORTED, "sample index beyond 2^24");
    if (n > (1u << 31)) return abi_fail(ILLUME_ERR_UNSUPPORTED, "at most 2^31 rays per call");
    if (n == 0) return ILLUME_OK;
    if (!rays || !L_out) return abi_fail(ILLUME_ERR_INVALID_ARG, "rays or L_out is null");
    DevScene dev = scene->dev;
    if (scene->n_prims == 0) dev.nodes = nullptr;  // the launcher writes black
    hipError_t e = launch_li_batch(dev, rays, n, cfg->max_depth, cfg->seed, cfg->sample_offset, L_out,
                                   (unsigned long long *)ray_counts, (hipStream_t)stream);
    if (e != hipSuccess) return abi_hip_fail(e, "li batch launch");
    return ILLUME_OK;
}

int illume_render_batch(const illume_job *jobs, uint32_t n_jobs, float *film, uint64_t *ray_counts, void *stream) {
    if (!film || (n_jobs && !jobs)) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    std::vector<RenderJob> hj;
    unsigned long long total = 0;
    for (uint32_t j = 0; j < n_jobs; ++j) {
        const illume_job &in = jobs[j];
        if (!in.scene) return abi_fail(ILLUME_ERR_INVALID_ARG, "job without a scene");
        const uint32_t kind = sampler_kind(in.cfg.flags);
        if (kind == ~0u) return abi_fail(ILLUME_ERR_INVALID_ARG, "unknown render flags");
        if ((uint64_t)in.cfg.sample_offset + in.cfg.spp > (1u << 24))
            return abi_fail(ILLUME_ERR_UNSUPPORTED, "sample index beyond 2^24");
        RenderJob job;
        std::memset(&job, 0, sizeof(job));
        if (!make_camera(in.camera, &job.cam)) return abi_fail(ILLUME_ERR_INVALID_ARG, "invalid camera");
        const unsigned long long n_pix = (unsigned long long)in.camera.width * in.camera.height;
        if (n_pix >= (1ull << 40)) return abi_fail(ILLUME_ERR_UNSUPPORTED, "image too large");
        const unsigned long long n = in.scene->n_prims ? n_pix * in.cfg.spp : 0ull;
        if (n == 0) continue;  // no samples, or nothing to hit and no lights: black
        if (kind && !sampler_launch(kind, 0, 0, (int)in.camera.width, (int)in.camera.height, &job.sl))
            return abi_fail(ILLUME_ERR_OUT_OF_MEMORY, "low-discrepancy tables");
        job.scene = in.scene->d_dev;
        job.path_begin = total;
        job.film_offset = in.film_offset;
        job.seed = in.cfg.seed;
        job.sample_offset = in.cfg.sample_offset;
        job.max_depth = in.cfg.max_depth;
        hj.push_back(job);
        total += n;
    }
    if (total == 0) return ILLUME_OK;
    hipStream_t st = (hipStream_t)stream;
    // The job table is copied from a host buffer that must outlive the copy.
    // A captured graph would replay the copy from a buffer released after the
    // first replay, so capture is refused.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    hipError_t e = hipStreamIsCapturing(st, &cap);
    if (e != hipSuccess) return abi_hip_fail(e, "stream capture query");
    if (cap != hipStreamCaptureStatusNone)
        return abi_fail(ILLUME_ERR_UNSUPPORTED, "illume_render_batch (and illume_render of scenes with 2^20 or more "
                                                "primitives) cannot be captured in a graph");
    const size_t bytes = hj.size() * sizeof(RenderJob);
    // host staging owned by the stream: released by a host callback once the
    // stream has passed the copy, whatever the driver does with pageable memory
    RenderJob *staging = new (std::nothrow) RenderJob[hj.size()];
    if (!staging) return abi_fail(ILLUME_ERR_OUT_OF_MEMORY, "job table staging");
    std::memcpy(staging, hj.data(), bytes);
    RenderJob *d_jobs = nullptr;
    e = hipMallocAsync((void **)&d_jobs, bytes, st);
    if (e != hipSuccess) {
        delete[] staging;
        return abi_hip_fail(e, "job table allocation");
    }
    e = hipMemcpyAsync(d_jobs, staging, bytes, hipMemcpyHostToDevice, st);
    const hipError_t eh = e == hipSuccess
                              ? hipLaunchHostFunc(st, [](void *p) { delete[] static_cast<RenderJob *>(p); }, staging)
                              : e;
    if (eh != hipSuccess) {
        // no callback queued: wait for the copy (if it was queued) before the staging goes
        (void)hipStreamSynchronize(st);
        delete[] staging;
    }
    if (e == hipSuccess) e = eh;
    if (e == hipSuccess)
        e = launch_render_batch(d_jobs, (uint32_t)hj.size(), total, film, (unsigned long long *)ray_counts, st);
    const hipError_t ef = hipFreeAsync(d_jobs, st);
    if (e != hipSuccess) return abi_hip_fail(e, "batched render launch");
    if (ef != hipSuccess) return abi_hip_fail(ef, "job table release");
    return ILLUME_OK;
}

int illume_render_filtered(const illume_scene *scene, const illume_camera *camera, const illume_render_cfg *cfg,
                           const illume_filter *filter, float *rgbw_sum, uint64_t *ray_counts, void *stream) {
    if (!scene || !camera || !cfg || !filter || !rgbw_sum) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    DevCamera cam;
    if (!make_camera(*camera, &cam)) return abi_fail(ILLUME_ERR_INVALID_ARG, "invalid camera");
    if (cfg->spp == 0) return ILLUME_OK;
    if ((uint64_t)cfg->sample_offset + cfg->spp > (1u << 24))
        return abi_fail(ILLUME_ERR_UNSUPPORTED, "sample index beyond 2^24");
    if ((uint64_t)camera->width + 64 >= (1u << 24) || (uint64_t)camera->height + 64 >= (1u << 24))
        return abi_fail(ILLUME_ERR_UNSUPPORTED, "image side beyond 2^24 - 64");
    FilmSplat fs;
    filter_splat(filter, camera->width, camera->height, &fs);
    if ((uint64_t)fs.wext * fs.hext >= (1ull << 40)) return abi_fail(ILLUME_ERR_UNSUPPORTED, "image too large");
    const uint32_t kind = sampler_kind(cfg->flags);
    if (kind == ~0u) return abi_fail(ILLUME_ERR_INVALID_ARG, "unknown render flags");
    if (scene->n_prims == 0) return ILLUME_OK;
    SamplerLaunch hl{};
    if (kind && !sampler_launch(kind, fs.x0, fs.y0, (int)fs.wext, (int)fs.hext, &hl))
        return abi_fail(ILLUME_ERR_OUT_OF_MEMORY, "low-discrepancy tables");
    hipError_t e = launch_render_filtered(scene->dev, cam, fs, cfg->spp, cfg->max_depth, cfg->seed, cfg->sample_offset,
                                          rgbw_sum, (unsigned long long *)ray_counts, hl, (hipStream_t)stream);
    if (e != hipSuccess) return abi_hip_fail(e, "filtered render launch");
    return ILLUME_OK;
}

}  // extern "C"
