// ORACLE — test infrastructure only (see pbrt.h). C entry points of the path
// tracer for the tests and the CPU baseline.
#include <atomic>
#include <thread>
#include "film.h"
#include "render.h"

using namespace orc;

namespace {
// 64-byte material record: type, kd[3], eta[3], k[3], alpha, ior, pad[4]
struct OrcMaterial {
    int32_t type;
    float kd[3], eta[3], k[3];
    float alpha, ior;
    float pad[4];
};
// camera record: pos[3], look[3], up[3], fov, width, height
struct OrcCamera {
    float pos[3], look[3], up[3], fov;
    uint32_t width, height;
};
}  // namespace

namespace {
// The pixel / sample loop of orc_render and orc_render_ao; li(ray, sampler,
// stats) is the integrator's Li.
template <class Li>
void render_loop(const void *camera, uint32_t spp, uint64_t seed, uint32_t sample_offset, uint32_t flags,
                 float *rgb_sum, uint64_t *stats, int n_threads, const Li &li) {
    const OrcCamera *c = (const OrcCamera *)camera;
    Camera cam;
    cam.init(V3(c->pos[0], c->pos[1], c->pos[2]), V3(c->look[0], c->look[1], c->look[2]),
             V3(c->up[0], c->up[1], c->up[2]), c->fov, (int)c->width, (int)c->height);
    uint64_t n_pix = (uint64_t)c->width * c->height;
    HaltonPixelMap hmap;
    hmap.init((int)c->width, (int)c->height);
    SobolPixelMap smap;
    smap.init(0, 0, (int)c->width, (int)c->height);
    const bool halton = (flags & 1u) != 0, sobol = (flags & 2u) != 0;
    int nt = std::max(1, n_threads);
    std::vector<RenderStats> st(nt);
    std::atomic<uint64_t> next(0);
    auto work = [&](int t) {
        while (true) {
            uint64_t b = next.fetch_add(64);
            if (b >= n_pix) break;
            uint64_t e = std::min(n_pix, b + 64);
            for (uint64_t pix = b; pix < e; ++pix) {
                int x = (int)(pix % c->width), y = (int)(pix / c->width);
                RGB sum;
                for (uint32_t si = sample_offset; si < sample_offset + spp; ++si) {
                    Sampler sampler;
                    if (halton) sampler.start_halton(&hmap, x, y, si);
                    else if (sobol) sampler.start_sobol(&smap, x, y, si);
                    else sampler.start(seed, pix, si);
                    // sampler.rs:78-84 get_camera_sample: p_film, time, p_lens
                    float uf[2], ul[2];
                    sampler.get_2d(uf);
                    float time = sampler.get_1d();
                    sampler.get_2d(ul);
                    Ray ray = cam.generate_ray((float)x + uf[0], (float)y + uf[1], time);
                    RGB L = li(ray, sampler, &st[t]);
                    if (std::isnan(L.x) || std::isnan(L.y) || std::isnan(L.z)) L = RGB();
                    else if (luminance(L) < -1e-5f) L = RGB();
                    else if (std::isinf(luminance(L))) L = RGB();
                    sum += L;
                }
                rgb_sum[3 * pix] += sum.x;
                rgb_sum[3 * pix + 1] += sum.y;
                rgb_sum[3 * pix + 2] += sum.z;
            }
        }
    };
    if (nt == 1) work(0);
    else {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) th.emplace_back(work, t);
        for (auto &x : th) x.join();
    }
    if (stats) {
        stats[0] = stats[1] = 0;
        for (auto &x : st) {
            stats[0] += x.closest_rays;
            stats[1] += x.any_rays;
        }
    }
}
}  // namespace

extern "C" {

// prim_emission: 3 floats per primitive (zero = not emissive; only triangles
// may emit). Every emissive triangle becomes one one-sided DiffuseAreaLight.
// sky: zenith RGB, horizon RGB (6 floats) or null for no sky.
void *orc_render_scene_create(const float *verts, uint32_t n_verts, const uint32_t *idx, uint32_t n_tris,
                              const float *spheres, uint32_t n_spheres, const uint32_t *prim_material,
                              const float *prim_emission, const void *materials, uint32_t n_materials,
                              const float *sky) {
    RenderScene *s = new RenderScene();
    s->ps.verts.resize(n_verts);
    for (uint32_t i = 0; i < n_verts; ++i) s->ps.verts[i] = V3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    s->ps.idx.assign(idx, idx + 3 * (size_t)n_tris);
    s->ps.spheres.resize(n_spheres);
    for (uint32_t i = 0; i < n_spheres; ++i)
        s->ps.spheres[i].init(V3(spheres[4 * i], spheres[4 * i + 1], spheres[4 * i + 2]), spheres[4 * i + 3]);
    s->bvh.build(&s->ps, 4);
    uint32_t np = n_tris + n_spheres;
    s->prim_material.assign(prim_material, prim_material + np);
    const OrcMaterial *m = (const OrcMaterial *)materials;
    for (uint32_t i = 0; i < n_materials; ++i) {
        Material x;
        x.type = m[i].type;
        x.kd = RGB(m[i].kd[0], m[i].kd[1], m[i].kd[2]);
        x.eta = RGB(m[i].eta[0], m[i].eta[1], m[i].eta[2]);
        x.k = RGB(m[i].k[0], m[i].k[1], m[i].k[2]);
        x.alpha = m[i].alpha;
        x.ior = m[i].ior;
        s->materials.push_back(x);
    }
    s->prim_light.assign(np, -1);
    std::vector<float> power;
    for (uint32_t p = 0; p < n_tris; ++p) {
        RGB L(prim_emission[3 * p], prim_emission[3 * p + 1], prim_emission[3 * p + 2]);
        if (is_black(L)) continue;
        AreaLight l;
        l.prim = p;
        l.p0 = s->ps.verts[idx[3 * p]];
        l.p1 = s->ps.verts[idx[3 * p + 1]];
        l.p2 = s->ps.verts[idx[3 * p + 2]];
        l.L = L;
        l.area = triangle_area(l.p0, l.p1, l.p2);
        s->prim_light[p] = (int)s->lights.size();
        s->lights.push_back(l);
        // DiffuseAreaLight::power = L * area * pi (one-sided), by luminance
        power.push_back(luminance(L * l.area * Pi));
    }
    if (sky) {
        // InfiniteAreaLight::power = pi r^2 * (average radiance over the
        // sphere), r the radius of the scene's bounding sphere; the analytic
        // sky's average is (2/3 zenith + 1/3 horizon) / 2.
        s->sky.on = true;
        s->sky.zenith = RGB(sky[0], sky[1], sky[2]);
        s->sky.horizon = RGB(sky[3], sky[4], sky[5]);
        // bounds of the triangles' vertices and the spheres' centre +- radius
        float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (uint32_t t = 0; t < n_tris; ++t)
            for (int k = 0; k < 3; ++k) {
                const float *v = verts + 3 * (size_t)idx[3 * (size_t)t + k];
                for (int a = 0; a < 3; ++a) b[a] = std::min(b[a], v[a]), b[3 + a] = std::max(b[3 + a], v[a]);
            }
        for (uint32_t i = 0; i < n_spheres; ++i)
            for (int a = 0; a < 3; ++a) {
                b[a] = std::min(b[a], spheres[4 * i + a] - spheres[4 * i + 3]);
                b[3 + a] = std::max(b[3 + a], spheres[4 * i + a] + spheres[4 * i + 3]);
            }
        const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
        const float r = np > 0 ? 0.5f * std::sqrt(dx * dx + dy * dy + dz * dz) : 0.f;
        const RGB avg = (s->sky.zenith * (2.f / 3.f) + s->sky.horizon * (1.f / 3.f)) * 0.5f;
        power.push_back(luminance(avg * (Pi * r * r)));
    }
    if (!power.empty()) s->light_distrib.init(power);
    return s;
}

void orc_render_scene_destroy(void *h) { delete (RenderScene *)h; }

// SamplerIntegratorBase::render (integrator.rs:48-145): for every pixel and
// sample index in [sample_offset, sample_offset + spp) trace one camera path,
// apply the NaN/negative/inf radiance guard (113-122) and add L to
// rgb_sum[pixel] (box filter of radius 0.5: weight 1 for the pixel the
// sample lands in). flags bit 0: the Halton sampler (pbrt-v3 HaltonSampler over
// the image), bit 1: the Sobol' sampler (pbrt-v3 SobolSampler), instead of the
// random streams. stats: [0] closest-hit rays, [1]
// any-hit rays.
void orc_render(void *h, const void *camera, uint32_t spp, uint32_t max_depth, uint64_t seed, uint32_t sample_offset,
                uint32_t flags, float *rgb_sum, uint64_t *stats, int n_threads) {
    RenderScene *s = (RenderScene *)h;
    render_loop(camera, spp, seed, sample_offset, flags, rgb_sum, stats, n_threads,
                [&](const Ray &ray, Sampler &sampler, RenderStats *st) {
                    return path_li(*s, ray, sampler, (int)max_depth, st);
                });
}

// SamplerIntegrator::li (integrator.rs:31-35) for caller-given rays (32 bytes
// each: o, d, t_max, time): ray i takes the sampler stream (seed, i, sample),
// skips the five dimensions of a camera sample and runs PathIntegrator::li;
// L_out gets the guarded radiance, 3 floats per ray.
void orc_li_batch(void *h, const float *rays, uint32_t n, uint32_t max_depth, uint64_t seed, uint32_t sample,
                  float *L_out, uint64_t *stats, int n_threads) {
    RenderScene *s = (RenderScene *)h;
    int nt = std::max(1, n_threads);
    std::vector<RenderStats> st(nt);
    std::atomic<uint64_t> next(0);
    auto work = [&](int t) {
        while (true) {
            uint64_t b = next.fetch_add(64);
            if (b >= n) break;
            uint64_t e = std::min<uint64_t>(n, b + 64);
            for (uint64_t i = b; i < e; ++i) {
                const float *r = rays + 8 * i;
                Sampler sampler;
                sampler.start(seed, i, sample);
                float u[2];
                sampler.get_2d(u);
                sampler.get_1d();
                sampler.get_2d(u);
                Ray ray(V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5]), r[6], r[7]);
                RGB L = path_li(*s, ray, sampler, (int)max_depth, &st[t]);
                if (std::isnan(L.x) || std::isnan(L.y) || std::isnan(L.z)) L = RGB();
                else if (luminance(L) < -1e-5f) L = RGB();
                else if (std::isinf(luminance(L))) L = RGB();
                L_out[3 * i] = L.x;
                L_out[3 * i + 1] = L.y;
                L_out[3 * i + 2] = L.z;
            }
        }
    };
    if (nt == 1) work(0);
    else {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) th.emplace_back(work, t);
        for (auto &x : th) x.join();
    }
    if (stats) {
        stats[0] = stats[1] = 0;
        for (auto &x : st) {
            stats[0] += x.closest_rays;
            stats[1] += x.any_rays;
        }
    }
}

// SamplerIntegratorBase::render with AOIntegrator::li (integrators/ao.rs:16-112):
// n_ao hemisphere samples per camera sample, cosine-distributed if cos_sample.
void orc_render_ao(void *h, const void *camera, uint32_t spp, uint32_t n_ao, int cos_sample, uint64_t seed,
                   uint32_t sample_offset, uint32_t flags, float *rgb_sum, uint64_t *stats, int n_threads) {
    RenderScene *s = (RenderScene *)h;
    render_loop(camera, spp, seed, sample_offset, flags, rgb_sum, stats, n_threads,
                [&](const Ray &ray, Sampler &sampler, RenderStats *st) {
                    return ao_li(*s, ray, sampler, (int)n_ao, cos_sample != 0, st);
                });
}

// SamplerIntegratorBase::render through a reconstruction filter: camera
// samples for every pixel of Film::get_sample_bounds (film.rs:97-103), the
// sampler stream of sample pixel (x, y) numbered (y - y0) * (x1 - x0) + (x - x0)
// (so a box filter of radius 0.5 uses orc_render's streams), each sample
// splatted by film_add_sample into rgbw (4 floats per image pixel). Threads
// take sample rows t, t + n, ... into private films summed in thread order, so
// the result does not depend on timing.
int orc_render_filtered(void *h, const void *camera, const void *filter_desc, uint32_t spp, uint32_t max_depth,
                        uint64_t seed, uint32_t sample_offset, uint32_t flags, float *rgbw, uint64_t *stats,
                        int n_threads) {
    RenderScene *s = (RenderScene *)h;
    const OrcCamera *c = (const OrcCamera *)camera;
    Filter f;
    if (!f.init(*(const FilterDesc *)filter_desc)) return -1;
    FilmFilter ff;
    ff.init(f);
    Camera cam;
    cam.init(V3(c->pos[0], c->pos[1], c->pos[2]), V3(c->look[0], c->look[1], c->look[2]),
             V3(c->up[0], c->up[1], c->up[2]), c->fov, (int)c->width, (int)c->height);
    const int W = (int)c->width, H = (int)c->height;
    int x0, y0, x1, y1;
    film_sample_bounds(ff, W, H, &x0, &y0, &x1, &y1);
    const uint64_t wext = (uint64_t)(x1 - x0);
    HaltonPixelMap hmap;
    hmap.init(x1 - x0, y1 - y0);
    SobolPixelMap smap;
    smap.init(x0, y0, x1 - x0, y1 - y0);
    const bool halton = (flags & 1u) != 0, sobol = (flags & 2u) != 0;
    const int nt = std::max(1, n_threads);
    std::vector<RenderStats> st(nt);
    std::vector<std::vector<float>> films(nt, std::vector<float>((size_t)W * H * 4, 0.f));
    auto work = [&](int t) {
        for (int y = y0 + t; y < y1; y += nt)
            for (int x = x0; x < x1; ++x) {
                const uint64_t pix = (uint64_t)(y - y0) * wext + (uint64_t)(x - x0);
                for (uint32_t si = sample_offset; si < sample_offset + spp; ++si) {
                    Sampler sampler;
                    if (halton) sampler.start_halton(&hmap, x, y, si);
                    else if (sobol) sampler.start_sobol(&smap, x, y, si);
                    else sampler.start(seed, pix, si);
                    float uf[2], ul[2];
                    sampler.get_2d(uf);
                    float time = sampler.get_1d();
                    sampler.get_2d(ul);
                    const float fx = (float)x + uf[0], fy = (float)y + uf[1];
                    Ray ray = cam.generate_ray(fx, fy, time);
                    RGB L = path_li(*s, ray, sampler, (int)max_depth, &st[t]);
                    if (std::isnan(L.x) || std::isnan(L.y) || std::isnan(L.z)) L = RGB();
                    else if (luminance(L) < -1e-5f) L = RGB();
                    else if (std::isinf(luminance(L))) L = RGB();
                    const float l3[3] = {L.x, L.y, L.z};
                    film_add_sample(ff, films[t].data(), W, H, fx, fy, l3);
                }
            }
    };
    if (nt == 1) work(0);
    else {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) th.emplace_back(work, t);
        for (auto &x : th) x.join();
    }
    for (int t = 0; t < nt; ++t)
        for (size_t i = 0; i < films[t].size(); ++i) rgbw[i] += films[t][i];
    if (stats) {
        stats[0] = stats[1] = 0;
        for (auto &x : st) {
            stats[0] += x.closest_rays;
            stats[1] += x.any_rays;
        }
    }
    return 0;
}

// The 16 x 16 filter table of a filter description (0 on success).
int orc_filter_table(const void *filter_desc, float *table) {
    Filter f;
    if (!f.init(*(const FilterDesc *)filter_desc)) return -1;
    FilmFilter ff;
    ff.init(f);
    for (int i = 0; i < 256; ++i) table[i] = ff.table[i];
    return 0;
}

// Filter::evaluate at n points (xy pairs).
int orc_filter_eval(const void *filter_desc, const float *xy, uint32_t n, float *out) {
    Filter f;
    if (!f.init(*(const FilterDesc *)filter_desc)) return -1;
    for (uint32_t i = 0; i < n; ++i) out[i] = f.eval(xy[2 * i], xy[2 * i + 1]);
    return 0;
}

// Film::get_sample_bounds: out = x0, y0, x1, y1.
int orc_film_sample_bounds(const void *filter_desc, int width, int height, int *out) {
    Filter f;
    if (!f.init(*(const FilterDesc *)filter_desc)) return -1;
    FilmFilter ff;
    ff.init(f);
    film_sample_bounds(ff, width, height, out, out + 1, out + 2, out + 3);
    return 0;
}

// n samples (film x, film y, r, g, b) splatted into rgbw (width x height x 4).
int orc_film_add_samples(const void *filter_desc, const float *samples, uint32_t n, int width, int height,
                         float *rgbw) {
    Filter f;
    if (!f.init(*(const FilterDesc *)filter_desc)) return -1;
    FilmFilter ff;
    ff.init(f);
    for (uint32_t i = 0; i < n; ++i)
        film_add_sample(ff, rgbw, width, height, samples[5 * i], samples[5 * i + 1], samples + 5 * i + 2);
    return 0;
}

// Camera ray probe: rays (8 floats: o, d, t_max, time) for film points.
void orc_camera_rays(const void *camera, const float *film_xy_time, uint32_t n, float *rays) {
    const OrcCamera *c = (const OrcCamera *)camera;
    Camera cam;
    cam.init(V3(c->pos[0], c->pos[1], c->pos[2]), V3(c->look[0], c->look[1], c->look[2]),
             V3(c->up[0], c->up[1], c->up[2]), c->fov, (int)c->width, (int)c->height);
    for (uint32_t i = 0; i < n; ++i) {
        Ray r = cam.generate_ray(film_xy_time[3 * i], film_xy_time[3 * i + 1], film_xy_time[3 * i + 2]);
        float *o = rays + 8 * i;
        o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z; o[3] = r.d.x; o[4] = r.d.y; o[5] = r.d.z;
        o[6] = r.t_max; o[7] = r.time;
    }
}

// Halton probes: radical inverse (scrambled with base_index's permutation when
// scrambled != 0) of n indices; the permutation of a base; the global index
// and first n dimensions of sample `sample` of pixel (px, py) for a sample
// bounds of res_x x res_y.
void orc_radical_inverse(int base_index, const uint64_t *a, uint32_t n, int scrambled, float *out) {
    const HaltonTables &t = halton_tables();
    for (uint32_t i = 0; i < n; ++i)
        out[i] = scrambled ? scrambled_radical_inverse(base_index, a[i], t.perms.data() + t.prime_sums[base_index])
                           : radical_inverse(base_index, a[i]);
}
void orc_halton_perm(int base_index, uint16_t *out, uint32_t *prime) {
    const HaltonTables &t = halton_tables();
    *prime = t.primes[base_index];
    for (uint32_t j = 0; j < t.primes[base_index]; ++j) out[j] = t.perms[t.prime_sums[base_index] + j];
}
uint64_t orc_inverse_radical_inverse(uint32_t base, uint64_t inverse, int n_digits) {
    return inverse_radical_inverse(base, inverse, n_digits);
}
int64_t orc_halton_sample(int res_x, int res_y, int px, int py, int64_t sample, uint32_t n, float *out) {
    HaltonPixelMap m;
    m.init(res_x, res_y);
    Sampler s;
    s.start_halton(&m, px, py, sample);
    for (uint32_t i = 0; i < n; ++i) out[i] = s.get_1d();
    return m.index_for_sample(px, py, sample);
}
void orc_halton_map(int res_x, int res_y, int64_t *out) {
    HaltonPixelMap m;
    m.init(res_x, res_y);
    out[0] = m.exp[0];
    out[1] = m.exp[1];
    out[2] = m.scale[0];
    out[3] = m.scale[1];
    out[4] = m.stride;
    out[5] = m.mult_inv[0];
    out[6] = m.mult_inv[1];
}

// Sobol' probes: the generator matrices and interval-to-index tables; one
// sample's global index and first n dimensions for a sample bounds.
void orc_sobol_tables(uint32_t *c, uint64_t *vdc, uint64_t *vdc_inv) {
    const SobolTables &t = sobol_tables();
    for (int i = 0; i < kSobolDims * kSobolMatrixSize; ++i) c[i] = t.c[i];
    for (int m = 0; m < 25; ++m)
        for (int k = 0; k < kSobolMatrixSize; ++k) vdc[m * kSobolMatrixSize + k] = t.vdc[m][k];
    for (int m = 0; m < 26; ++m)
        for (int k = 0; k < kSobolMatrixSize; ++k) vdc_inv[m * kSobolMatrixSize + k] = t.vdc_inv[m][k];
}
int64_t orc_sobol_sample(int x0, int y0, int w, int h, int px, int py, int64_t sample, uint32_t n, float *out) {
    SobolPixelMap m;
    m.init(x0, y0, w, h);
    Sampler s;
    s.start_sobol(&m, px, py, sample);
    for (uint32_t i = 0; i < n; ++i) out[i] = s.get_1d();
    return m.index_for_sample(px, py, sample);
}
float orc_sobol_value(int64_t index, int dim) { return sobol_sample(index, dim, 0); }

// Stream probe: n uniform floats of the sampler stream for (seed, pixel, sample).
void orc_sampler_floats(uint64_t seed, uint64_t pixel, uint64_t sample, uint32_t n, float *out) {
    Sampler s;
    s.start(seed, pixel, sample);
    for (uint32_t i = 0; i < n; ++i) out[i] = s.get_1d();
}

}  // extern "C"
