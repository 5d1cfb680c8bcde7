// ORACLE — test infrastructure only (see pbrt.h). The path tracer: pbrt-v3's
// PathIntegrator (absent from the reference; restated from the PBR book 3e,
// which the reference ports) on the reference's BSDF, light, camera, sampler
// and film pieces.
#pragma once
#include <vector>
#include "bvh.h"
#include "halton.h"
#include "sobol.h"

namespace orc {

typedef V3 RGB;
inline RGB cmul(const RGB &a, const RGB &b) { return RGB(a.x * b.x, a.y * b.y, a.z * b.z); }
inline RGB cdiv(const RGB &a, const RGB &b) { return RGB(a.x / b.x, a.y / b.y, a.z / b.z); }
inline bool is_black(const RGB &a) { return a.x == 0 && a.y == 0 && a.z == 0; }
// spectrum.rs:498-501
inline float luminance(const RGB &c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }
inline float max_comp(const RGB &c) { return std::max(c.x, std::max(c.y, c.z)); }

enum MaterialType { MAT_MATTE = 0, MAT_METAL = 1, MAT_GLASS = 2 };
struct Material {
    int type = MAT_MATTE;
    RGB kd;          // matte: diffuse reflectance
    RGB eta, k;      // metal: conductor eta, k (RGB)
    float alpha = 0.1f;  // metal: Trowbridge-Reitz alpha (isotropic)
    float ior = 1.5f;    // glass: index of refraction (Kr = Kt = 1)
};

// DiffuseAreaLight (absent from the reference; PBR book 3e 12.5) on one
// triangle, one-sided.
struct AreaLight {
    uint32_t prim;
    V3 p0, p1, p2;
    RGB L;
    float area;
};

// Analytic sky (an infinite area light; pbrt-v3 InfiniteAreaLight with an
// analytic radiance instead of an image map): L(w) = zenith * t +
// horizon * (1 - t), t = sqrt(w.y), for w.y > 0 and 0 below the horizon.
// Sampled cosine-weighted about +y (pdf w.y / pi).
struct Sky {
    bool on = false;
    RGB zenith, horizon;
    RGB L(const V3 &w) const;
};

// sampling.rs:8-93 Distribution1D (discrete part)
struct Distribution1D {
    std::vector<float> func, cdf;
    float func_int = 0;
    void init(const std::vector<float> &f);
    int sample_discrete(float u, float *pdf) const;
};

struct RenderScene {
    PrimSet ps;
    BVH bvh;
    std::vector<uint32_t> prim_material;
    std::vector<int> prim_light;  // light index or -1
    std::vector<Material> materials;
    std::vector<AreaLight> lights;
    Sky sky;                       // light index lights.size() in the distribution when on
    Distribution1D light_distrib;  // power distribution (integrator.rs:496-505)
    int n_lights() const { return (int)lights.size() + (sky.on ? 1 : 0); }
};

// PerspectiveCamera (cameras/perspective.rs:35-126) with pbrt-v3's
// perspective projection (divergence 17) and look-at camera_to_world.
struct Camera {
    Transform camera_to_world, raster_to_camera;
    int width = 0, height = 0;
    void init(const V3 &pos, const V3 &look, const V3 &up, float fov, int w, int h);
    Ray generate_ray(float fx, float fy, float time) const;  // perspective.rs:103-126
};

// The sampler of a camera path. Random (default): one PCG32 stream per
// (pixel, sample index), consumed sequentially by get_1d/get_2d in pbrt's
// RandomSampler order. Halton (start_halton) and Sobol' (start_sobol):
// dimension d of the path's global index, d counting up from 0 (pbrt-v3
// GlobalSampler); past the 1000 prime bases or 1024 Sobol' dimensions, a hash
// of (index, d).
struct Sampler {
    Rng rng;
    const HaltonPixelMap *halton = nullptr;
    const SobolPixelMap *sobol = nullptr;
    int64_t index = 0;
    int dim = 0, px = 0, py = 0;
    void start(uint64_t seed, uint64_t pixel, uint64_t sample);
    void start_halton(const HaltonPixelMap *h, int x, int y, int64_t sample) {
        halton = h;
        sobol = nullptr;
        index = h->index_for_sample(x, y, sample);
        dim = 0;
    }
    void start_sobol(const SobolPixelMap *s, int x, int y, int64_t sample) {
        halton = nullptr;
        sobol = s;
        px = x;
        py = y;
        index = s->index_for_sample(x, y, sample);
        dim = 0;
    }
    float get_1d() { return halton ? halton_next() : (sobol ? sobol_next() : rng.uniform_float()); }
    void get_2d(float *u) {
        u[0] = get_1d();
        u[1] = get_1d();
    }
    float halton_next();
    float sobol_next();
};
uint64_t sampler_stream(uint64_t seed, uint64_t pixel, uint64_t sample);

struct RenderStats {
    uint64_t closest_rays = 0, any_rays = 0;
};

// PathIntegrator::li (PBR book 3e 14.5.4)
RGB path_li(const RenderScene &sc, Ray ray, Sampler &sampler, int max_depth, RenderStats *st);
RGB ao_li(const RenderScene &sc, Ray ray, Sampler &sampler, int n_samples, bool cos_sample, RenderStats *st);

}  // namespace orc
