/*
 * illume_hip.h — C ABI of the MI355X ray-tracing library (librs_illume_amd.so).
 *
 * This is the drop-in boundary for rs_illume's hot path. Each entry point names
 * the reference interface it replaces (file:line in mlunnay/rs_illume). No
 * C++ or torch types cross it: plain pointers, sizes and an opaque hipStream_t
 * passed as void*. Every call returns an illume_status; nothing unwinds across
 * the ABI. Ray/hit buffers are caller-owned DEVICE pointers.
 *
 * Threading (the reference's traits are Send + Sync and are queried from rayon
 * workers, core/primitive/mod.rs:6, core/integrator.rs:59): BVH, scene, filter
 * handles are immutable after creation and own their device memory, so any
 * number of host threads may query or render one handle at once, each on its
 * own stream. A guide handle is read-only for rendering and inference but
 * written by its training calls; a wavefront handle holds the per-render path
 * pool and is used by one call at a time (one stream, one host thread). The
 * last-error message is kept per host thread. Scenes may be created from
 * several host threads at once: each scene's set-up (its device tree build,
 * vertex table and uploads) runs on a stream of its own, so builds of
 * different scenes overlap on the device.
 */
#ifndef ILLUME_HIP_H
#define ILLUME_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (0 = success). core/primitive/mod.rs:6-17 returns Option/bool and
 * panics on broken invariants; here a miss is data (ILLUME_MISS), errors are codes. */
typedef enum {
    ILLUME_OK = 0,
    ILLUME_ERR_INVALID_ARG = 1,   /* null pointer, index out of range, bad size */
    ILLUME_ERR_HIP = 2,           /* a HIP runtime call failed */
    ILLUME_ERR_OUT_OF_MEMORY = 3, /* device allocation failed */
    ILLUME_ERR_UNSUPPORTED = 4,   /* parameter outside what the library implements */
    ILLUME_ERR_QUEUE_OVERFLOW = 5 /* a work queue was too small (renderer) */
} illume_status;

#define ILLUME_MISS 0xFFFFFFFFu

/* core/geometry/ray.rs:9-31 (origin, direction, t_max, time); 32 bytes.
 * Medium and differentials are not part of the hot path. */
typedef struct {
    float o[3];
    float d[3];
    float t_max;
    float time;
} illume_ray;

/* Closest hit record; 16 bytes. prim_id indexes the caller's primitive order
 * (triangles 0..n_tris-1, then spheres), not the BVH's internal order
 * (accelerators/bvh.rs:58). ILLUME_MISS on a miss. b1, b2 are the triangle
 * barycentrics of vertices 1 and 2 (0 for spheres). */
typedef struct {
    uint32_t prim_id;
    float t;
    float b1;
    float b2;
} illume_hit;

typedef struct illume_bvh illume_bvh;

typedef struct {
    uint64_t n_prims;
    uint64_t n_inner_nodes;  /* 64-byte nodes: two-child, or quantized four-child (width 4) */
    uint64_t n_leaves;
    uint64_t max_depth;
    uint64_t device_bytes;   /* nodes + primitive records */
    double build_seconds;
    float bounds[6];         /* scene bounds: min xyz, max xyz */
} illume_bvh_info_t;

/* Build a BVH over triangles (world-space vertex array + index triples) and
 * full spheres (cx, cy, cz, radius) and upload it to the current device.
 * Replaces BVHAccel::new (accelerators/bvh.rs:32-93) with
 * SplitMethod::SAH (bvh.rs:17-22) over GeometricPrimitive(Triangle|Sphere)
 * (core/primitive/geometric_primitive.rs:8-30). max_prims_in_node: 1..15.
 * Host pointers; verts/tri_idx may be null when n_tris == 0, spheres when
 * n_spheres == 0. */
int illume_bvh_create(const float *verts, uint32_t n_verts, const uint32_t *tri_idx, uint32_t n_tris,
                      const float *spheres, uint32_t n_spheres, int max_prims_in_node, illume_bvh **out);

/* Same build with a chosen node width: 2 = 64-byte two-child nodes (float
 * boxes), 4 = 64-byte four-child nodes (the binary SAH tree collapsed to half
 * the depth, child boxes quantized to one byte per plane and rounded outwards,
 * so every box the float test enters is entered). Hit results do not depend
 * on the width. illume_bvh_create uses width 2. */
int illume_bvh_create_ex(const float *verts, uint32_t n_verts, const uint32_t *tri_idx, uint32_t n_tris,
                         const float *spheres, uint32_t n_spheres, int max_prims_in_node, int width,
                         illume_bvh **out);

/* The same aggregate built on the device: a linear BVH over Morton codes of
 * the primitive centroids (the role of BVHAccel::new's HLBVH branch,
 * accelerators/bvh.rs:249-502), subtrees of at most max_prims_in_node
 * primitives as leaves; width 4 collapses it on the host. Builds in
 * milliseconds where the SAH build takes seconds; traversal is somewhat
 * slower (a lower-quality tree), hit results are identical. Host input
 * arrays; the build runs on `stream` and is complete on return. */
int illume_bvh_create_gpu(const float *verts, uint32_t n_verts, const uint32_t *tri_idx, uint32_t n_tris,
                          const float *spheres, uint32_t n_spheres, int max_prims_in_node, int width, void *stream,
                          illume_bvh **out);

/* The binned SAH build of illume_bvh_create_ex (BVHAccel::new with
 * SplitMethod::SAH, accelerators/bvh.rs:32-247) run on the device: the same
 * split rule evaluated on the same exact bin boxes and counts, so the tree is
 * the host build's (the same inner nodes and boxes, the same leaves; only the
 * order of primitives inside a leaf may differ, and a node of more than 128
 * primitives below depth 40 is split by count where the host takes an object
 * median). Nodes of more than 128 primitives are split level by level over
 * the whole device, smaller subtrees by one wave each; width 4 is collapsed
 * and quantized on the device too. Traverses exactly as fast as the host
 * tree and builds in milliseconds. Host input arrays; the build runs on
 * `stream` and is complete on return. */
int illume_bvh_create_gpu_sah(const float *verts, uint32_t n_verts, const uint32_t *tri_idx, uint32_t n_tris,
                              const float *spheres, uint32_t n_spheres, int max_prims_in_node, int width,
                              void *stream, illume_bvh **out);

/* Frees the handle and its device memory (Drop of BVHAccel). */
int illume_bvh_destroy(illume_bvh *bvh);

/* Size and shape of a built BVH (the reference's BVH stats counters,
 * accelerators/bvh.rs:676-720). */
int illume_bvh_info(const illume_bvh *bvh, illume_bvh_info_t *out);

/* Closest hit for n rays: batched Primitive::intersect on the aggregate
 * (core/primitive/mod.rs:11, accelerators/bvh.rs:563-612, reached through
 * Scene::intersect core/scene.rs:43-47). Each ray's t_max bounds the search.
 * rays, hits: device pointers. stream: hipStream_t (null = default stream).
 * n <= 2^31 (larger batches: ILLUME_ERR_UNSUPPORTED; split them).
 * The persistent launches — closest and any hit on a four-child tree, any
 * hit on a two-child tree — allocate 8 bytes stream-ordered on `stream` for
 * their ray queue and free them the same way (hipMallocAsync/hipFreeAsync);
 * closest hit on a two-child tree allocates nothing. */
int illume_intersect_closest(const illume_bvh *bvh, const illume_ray *rays, uint32_t n, illume_hit *hits,
                             void *stream);

/* Any hit for n rays: batched Primitive::intersect_p (core/primitive/mod.rs:12,
 * accelerators/bvh.rs:614-662, via Scene::intersect_p core/scene.rs:49-53).
 * occluded[i] = 1 if any primitive is hit in (0, t_max], else 0. n <= 2^31;
 * allocates its queue counter as illume_intersect_closest does. */
int illume_intersect_any(const illume_bvh *bvh, const illume_ray *rays, uint32_t n, uint8_t *occluded,
                         void *stream);

/* No counterpart in the reference: illume_intersect_closest for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments and results, bit for
 * bit; it calls no hipMalloc/hipFree (stream-ordered or not). The persistent
 * launch's ray queue counter lives in a 256-byte array that the library keeps
 * on each device until the process ends (its address is looked up once per
 * device, on the first call with that device current); the counter is reset
 * by a one-thread kernel at the start of every call, so any ray count up to
 * the limit (n <= 2^31, else ILLUME_ERR_UNSUPPORTED) uses the same 8 bytes.
 * Calls that overlap on one device (two streams or two host threads at once)
 * share that counter and race: serialise them. */
int illume_intersect_closest_noalloc(const illume_bvh *bvh, const illume_ray *rays, uint32_t n, illume_hit *hits,
                                     void *stream);

/* One shape against a batch of rays, for parity tests of the shapes alone:
 * Shape::intersect (any == 0; core/shape.rs:8-112, shapes/triangle.rs:142-419,
 * shapes/sphere.rs:105-236) or Shape::intersect_p (any != 0; triangle.rs:421-607,
 * sphere.rs:238-323) of a single Triangle (shape: 9 floats, vertices p0 p1 p2)
 * or full Sphere (shape: 4 floats, centre and radius). shape is a host
 * pointer; rays, hits device pointers. A hit writes prim_id 0, t and (triangles,
 * closest hit) the barycentrics b1, b2; a miss prim_id ILLUME_MISS. The closest
 * hit rejects degenerate triangles as Triangle::intersect does. n <= 2^31. */
#define ILLUME_SHAPE_TRIANGLE 0
#define ILLUME_SHAPE_SPHERE 1
int illume_shape_intersect_batch(int32_t kind, const float *shape, const illume_ray *rays, uint32_t n, int32_t any,
                                 illume_hit *hits, void *stream);

/* Instrumented closest/any traversal that also counts, per ray, node-box tests
 * and primitive tests (the reference's "BVH/..." and "Intersections/..."
 * counters, core/stats_accumulator.rs). counts: device uint32[2*n] (node
 * tests, primitive tests) per ray. any != 0 selects the any-hit traversal. */
int illume_intersect_counts(const illume_bvh *bvh, const illume_ray *rays, uint32_t n, int any, uint32_t *counts,
                            void *stream);

/* ------------------------------------------------------------------ path tracing */

typedef struct illume_scene illume_scene;

/* Materials with constant parameters (core/material.rs:14-16 is a trait only;
 * these are pbrt-v3's Matte, Metal and Glass): */
typedef enum {
    ILLUME_MATTE = 0, /* LambertianReflection(kd) (reflection/lambertian_reflection.rs:17-45) */
    ILLUME_METAL = 1, /* MicrofacetReflection(TrowbridgeReitz(alpha), FresnelConductor(eta, k))
                         (reflection/microfacet_reflection.rs:24-87, microfacet/trowbridge_reitz_distribution.rs) */
    ILLUME_GLASS = 2  /* FresnelSpecular(R=1, T=1, 1, ior) (reflection/fresnel_specular.rs:29-86) */
} illume_material_type;

/* 64 bytes. alpha is the Trowbridge-Reitz alpha itself (no roughness remap). */
typedef struct {
    int32_t type;
    float kd[3];
    float eta[3];
    float k[3];
    float alpha;
    float ior;
    float pad[4];
} illume_material;

/* PerspectiveCamera (cameras/perspective.rs:35-80) from a look-at frame
 * (core/transform.rs:181-208); fov in degrees spans the shorter image axis. */
typedef struct {
    float pos[3];
    float look[3];
    float up[3];
    float fov;
    uint32_t width;
    uint32_t height;
} illume_camera;

/* Per-call render parameters: sample indices [sample_offset, sample_offset+spp)
 * of every pixel are traced (counter-based sampler streams keyed by seed,
 * pixel and sample index, so splitting spp over calls gives the same image). */
typedef struct {
    uint32_t spp;
    uint32_t max_depth;
    uint32_t sample_offset;
    uint32_t flags; /* 0, ILLUME_RENDER_HALTON or ILLUME_RENDER_SOBOL */
    uint64_t seed;
} illume_render_cfg;

/* flags: the Halton sampler instead of the random streams (illume_render,
 * illume_render_filtered and the wavefront renderer; the seed is then unused). pbrt-v3's HaltonSampler
 * (PBR book 3e 7.4) over the reference's radical-inverse primitives
 * (core/low_discrepancy.rs:196-205, 422-466, 1505-1531): sample s of pixel
 * (x, y) is the global index that the pixel map of the sample bounds gives;
 * dimensions 0-1 are the film offset, 2-4 time and lens, then the path's
 * dimensions in order, each a radical inverse in the next prime base with
 * that base's digit permutation (1000 bases; past them a hash of (index,
 * dimension)). */
#define ILLUME_RENDER_HALTON 1u

/* flags: the Sobol' sampler instead of the random streams (illume_render,
 * illume_render_filtered and the wavefront renderer; the seed is then unused). pbrt-v3's SobolSampler
 * (PBR book 3e 7.7) over the reference's sobol_sample_float and
 * sobol_interval_to_index (core/low_discrepancy.rs:354-420) and generator
 * matrices equal to its SOBOL_MATRICES_32 / VD_C_SOBOL_MATRICES(_INV)
 * (core/sobolmatrices.rs), built from the Joe-Kuo direction numbers: sample s
 * of pixel (x, y) is the index whose first two dimensions fall in that pixel of
 * the power-of-two grid over the sample bounds; later dimensions in order
 * (1024; past them a hash of (index, dimension)). The sample bounds' larger
 * side may not exceed 2^25. */
#define ILLUME_RENDER_SOBOL 2u

/* Build a renderable scene: the BVH of illume_bvh_create (max 4 primitives per
 * leaf) plus per-primitive material index and emitted radiance (3 floats per
 * primitive; non-zero makes that TRIANGLE a one-sided DiffuseAreaLight, which
 * the reference lacks; spheres may not emit). Lights are chosen by power
 * (core/lightdistrib.rs, integrator.rs:496-505). Host pointers.
 * Replaces Scene::new (core/scene.rs:20-37). */
int illume_scene_create(const float *verts, uint32_t n_verts, const uint32_t *tri_idx, uint32_t n_tris,
                        const float *spheres, uint32_t n_spheres, const uint32_t *prim_material,
                        const float *prim_emission, const illume_material *materials, uint32_t n_materials,
                        illume_scene **out);

/* Analytic sky: an infinite area light (the role of pbrt-v3's
 * InfiniteAreaLight, which the reference lacks) with radiance
 * L(w) = zenith * t + horizon * (1 - t), t = sqrt(w.y), above the horizon
 * (w.y > 0) and 0 below; sampled cosine-weighted about +y. In the power
 * distribution its power is pi r^2 times its sphere-averaged radiance, r the
 * radius of the scene's bounding sphere. Escaping camera and specular rays
 * see it. */
typedef struct illume_sky {
    float zenith[3];
    float horizon[3];
} illume_sky;

/* Options of illume_scene_create_ex. */
typedef struct illume_scene_opts {
    const illume_sky *sky; /* NULL: no sky */
    int32_t bvh_builder;   /* 0: binned SAH on the host (illume_bvh_create), 1: linear BVH on the device
                              (illume_bvh_create_gpu), 2: binned SAH on the device
                              (illume_bvh_create_gpu_sah) */
    int32_t reserved;
} illume_scene_opts;

/* illume_scene_create with options (opts == NULL: the defaults, no sky). */
int illume_scene_create_ex(const float *verts, uint32_t n_verts, const uint32_t *tri_idx, uint32_t n_tris,
                           const float *spheres, uint32_t n_spheres, const uint32_t *prim_material,
                           const float *prim_emission, const illume_material *materials, uint32_t n_materials,
                           const illume_scene_opts *opts, illume_scene **out);

int illume_scene_destroy(illume_scene *scene);

/* The scene's aggregate, for Scene::intersect / intersect_p queries through
 * illume_intersect_closest / illume_intersect_any (core/scene.rs:43-53). The
 * handle stays owned by the scene. */
int illume_scene_bvh(const illume_scene *scene, const illume_bvh **out);

/* Integrator::render (core/integrator.rs:17-29, SamplerIntegratorBase::render
 * 48-145) with PathIntegrator::li: for every pixel adds the sum of the
 * radiance of cfg->spp camera paths to rgb_sum[3*(y*width+x) + c] (a device
 * float buffer of width*height*3 the caller zeroes and divides by the total
 * sample count; box filter of radius 0.5). ray_counts (device, 2 x uint64, may
 * be null) accumulates the number of closest-hit and any-hit rays traced.
 * Scenes of 2^20 or more primitives run through illume_render_batch and, like
 * it, refuse stream capture (ILLUME_ERR_UNSUPPORTED). */
int illume_render(const illume_scene *scene, const illume_camera *camera, const illume_render_cfg *cfg,
                  float *rgb_sum, uint64_t *ray_counts, void *stream);

/* Integrator::render with AOIntegrator::li (integrators/ao.rs:16-112; the
 * reference's only implemented Li): for every pixel adds, per camera sample,
 * the ambient occlusion at the camera ray's closest hit to all three channels
 * of rgb_sum[3*(y*width+x) + c] (device floats the caller zeroes). Each of
 * n_ao_samples hemisphere directions about the geometric normal facing the
 * ray (cosine-distributed if cos_sample != 0, else uniform; two sampler
 * dimensions each, after the camera sample) adds (wi . n) / (pdf n_ao_samples)
 * when its shadow ray is unoccluded: an unoccluded point gives pi, a miss 0.
 * cfg->max_depth is unused; cfg->flags selects the sampler as in
 * illume_render. n_ao_samples <= 65536. ray_counts as in illume_render. */
int illume_render_ao(const illume_scene *scene, const illume_camera *camera, const illume_render_cfg *cfg,
                     uint32_t n_ao_samples, int32_t cos_sample, float *rgb_sum, uint64_t *ray_counts, void *stream);

/* SamplerIntegrator::li (core/integrator.rs:31-35) for caller-given rays, for
 * parity tests: ray i takes the random sampler stream (cfg->seed, i,
 * cfg->sample_offset), skips the five dimensions of a camera sample and runs
 * PathIntegrator::li (cfg->max_depth) from the ray, so the camera ray of pixel
 * p for sample s gives illume_render's sample s of pixel p. L_out: device
 * floats, 3 per ray, the radiance after SamplerIntegratorBase::render's guard.
 * cfg->flags must be 0; cfg->spp is unused. Latency-bound by design: renders
 * go through illume_render. ray_counts as in illume_render. */
int illume_li_batch(const illume_scene *scene, const illume_ray *rays, uint32_t n, const illume_render_cfg *cfg,
                    float *L_out, uint64_t *ray_counts, void *stream);

/* ------------------------------------------------------------------ reconstruction filters
 * The Filter trait (core/filter.rs) and its implementations: BoxFilter
 * (filters/box_filter.rs), TriangleFilter (triangle.rs), GaussianFilter
 * (gaussian.rs), MitchellFilter (mitchell.rs), LanczosSincFilter (sinc.rs),
 * tabulated as Film::new does (core/film.rs:69-81: 16 x 16 values at
 * ((i + 0.5) r / 16) over the positive quadrant). */

typedef enum illume_filter_type {
    ILLUME_FILTER_BOX = 0,
    ILLUME_FILTER_TRIANGLE = 1,
    ILLUME_FILTER_GAUSSIAN = 2,
    ILLUME_FILTER_MITCHELL = 3,
    ILLUME_FILTER_LANCZOS = 4
} illume_filter_type;

typedef struct illume_filter_desc {
    int32_t type;     /* illume_filter_type */
    float radius[2];  /* x, y radius in pixels, each in (0, 32] */
    float alpha;      /* Gaussian falloff */
    float b, c;       /* Mitchell B, C */
    float tau;        /* Lanczos window: sinc(x) sinc(x / tau) */
} illume_filter_desc;

typedef struct illume_filter illume_filter;

int illume_filter_create(const illume_filter_desc *desc, illume_filter **out);
int illume_filter_destroy(illume_filter *filter);
/* Host copy of the 16 x 16 weight table (row y, column x). */
int illume_filter_table(const illume_filter *filter, float *host_table256);

/* illume_render through a reconstruction filter: camera samples are drawn for
 * every pixel of the sample bounds (Film::get_sample_bounds, film.rs:97-103:
 * the image grown by the filter radius) and each is splatted as
 * FilmTile::add_sample does (film.rs:288-333, with pbrt-v3's upper bound
 * floor(p - 0.5 + r) + 1 and full x loop): rgbw_sum[4*(y*width+x) + c] gets
 * w * L (c < 3) and w (c = 3), w the table weight. The caller zeroes rgbw_sum
 * (device float, width*height*4) and resolves it with illume_film_resolve.
 * With a box filter of radius 0.5 the sample positions, sampler streams and
 * radiance equal illume_render's and every weight is 1. */
int illume_render_filtered(const illume_scene *scene, const illume_camera *camera, const illume_render_cfg *cfg,
                           const illume_filter *filter, float *rgbw_sum, uint64_t *ray_counts, void *stream);

/* Film::write_image's per-pixel step (film.rs:187-226, without XYZ): rgb_out[3p + c] =
 * scale * (w != 0 ? max(0, rgbw[4p + c] / w) : rgbw[4p + c]). Device buffers. */
int illume_film_resolve(const float *rgbw_sum, uint64_t n_pixels, float scale, float *rgb_out, void *stream);

/* ------------------------------------------------------------------ learned proposal
 * No counterpart in the reference: an MLP (144-64-64-64-40, ReLU) from the
 * frequency encoding of a hit's normalised position and normal to an 8-lobe
 * von Mises-Fisher mixture over incident directions, BF16 weights on MFMA with
 * FP32 accumulation, trained online with Adam (beta1 0.9, beta2 0.999,
 * eps 1e-8) on the weighted negative log-likelihood of path-traced radiance
 * samples. Parameters: 20,200 floats, per layer W[in][out] row-major then b. */

typedef struct illume_guide illume_guide;

int illume_guide_create(const float *init_params, float lr, illume_guide **out);
int illume_guide_destroy(illume_guide *guide);
/* Copies the fp32 master parameters to host memory (20,200 floats). */
int illume_guide_get_params(const illume_guide *guide, float *host_out);
/* inputs: device [n][6] (position normalised to [0,1]^3 by the scene bounds,
 * unit normal); out: device [n][40] raw network outputs. */
int illume_guide_infer(const illume_guide *guide, const float *inputs, uint32_t n, float *out, void *stream);
/* records: device [n][10] (position (3), normal (3), unit direction (3),
 * weight); grad_out: device [20200], the gradient of
 * -sum_r (w_r / sum w) log q(dir_r | pos_r, normal_r). */
int illume_guide_grad(const illume_guide *guide, const float *records, uint32_t n, float *grad_out, void *stream);
/* One Adam step on the gradient of that loss over the batch. */
int illume_guide_train_step(illume_guide *guide, const float *records, uint32_t n, void *stream);
/* Adam steps taken so far and the loss of the last one (synchronises the device). */
int illume_guide_stats(const illume_guide *guide, uint64_t *steps, float *last_loss);

/* ------------------------------------------------------------------ wavefront renderer
 * The same Integrator::render as illume_render, executed as per-stage kernels
 * over a pool of path slots whose state lives in HBM: regenerate (camera
 * paths into free slots), extend (closest hit), proposal inference (MFMA),
 * shade (NEE, MIS, continuation; survivors compacted into the next queue).
 * Iterations are captured in a hipGraph and replayed; one render call can
 * batch many scenes and cameras of uneven size (jobs) through one pool.
 * With guide == NULL or guide_frac == 0 the image equals illume_render's. */

typedef struct illume_wavefront illume_wavefront;

typedef struct {
    float guide_frac;      /* probability of sampling the proposal (one-sample MIS with the BSDF) */
    uint32_t train;        /* 0: no training. N > 0: finished paths emit radiance samples and,
                              after every replay of the captured iterations that gathered at
                              least train_batch samples, up to N Adam steps run, each on an
                              interleaved subset of train_batch samples */
    uint32_t train_batch;  /* 0 = 65536 */
    float min_alpha;       /* the proposal is used (and trained) only at diffuse vertices and at
                              microfacet vertices with alpha >= min_alpha; near-specular lobes
                              keep plain BSDF sampling. 0 = every non-specular vertex */
    uint32_t flags;        /* ILLUME_GUIDE_PRODUCT or 0 */
} illume_guide_cfg;

/* illume_guide_cfg.flags: sample the proposal multiplied by a von Mises-Fisher
 * fit of the vertex's BSDF lobe (about the mirror direction for microfacet
 * reflection, about the normal for Lambertian) instead of the proposal alone.
 * The product of two vMF lobes is a vMF lobe, so the product mixture is
 * sampled and its pdf evaluated exactly; the one-sample MIS with the BSDF
 * stays unbiased. */
#define ILLUME_GUIDE_PRODUCT 1u

/* One render of a batch: its image occupies width*height*3 floats of the
 * film buffer starting at film_offset. */
typedef struct {
    const illume_scene *scene;
    illume_camera camera;
    illume_render_cfg cfg;
    uint64_t film_offset;
} illume_job;

/* The jobs of a batch on the megakernel of illume_render, in one launch over
 * all their paths (each thread finds its job): every job's image equals its
 * own illume_render call (cfg.flags selects its sampler; no proposal). For
 * many plain renders of uneven size; the wavefront pool below adds the
 * learned proposal. Integrator::render (core/integrator.rs:17-29) per job.
 * The job table is copied from a host buffer released by a stream-ordered host
 * callback, so the call cannot be captured in a graph: on a capturing stream it
 * returns ILLUME_ERR_UNSUPPORTED before any device work. */
int illume_render_batch(const illume_job *jobs, uint32_t n_jobs, float *film, uint64_t *ray_counts, void *stream);

/* A pool of max_paths path slots (64 .. 2^26, rounded up to a multiple of
 * 1,024) for paths of up to max_depth (1 .. 64) bounces. An unfiltered batch
 * of up to 2^29 paths also takes a per-path sample buffer (16 bytes per path,
 * grown on demand, kept by the pool) that one kernel sums into the film at
 * the end of the call; larger or filtered batches add to the film with
 * atomics. With the null stream a render runs on the pool's own stream and
 * returns once the film is complete. */
int illume_wavefront_create(uint32_t max_paths, uint32_t max_depth, illume_wavefront **out);
int illume_wavefront_destroy(illume_wavefront *wf);
/* Launch shape: threads per block of the stage kernels (64/128/256; each
 * block owns a tile of 4 x block slots), a tile regenerates only while fewer
 * than its share of regen_min slots are live (max_paths+1 or more = always),
 * iterations per captured graph (0 = no graph; rounded up to an even count).
 * Chosen by the autotuner (rs_illume_amd/autotune.py). */
int illume_wavefront_set_tuning(illume_wavefront *wf, uint32_t block, uint32_t regen_min, uint32_t graph_iters);
/* Iterations the last render call ran. */
uint64_t illume_wavefront_last_iterations(const illume_wavefront *wf);
int illume_wavefront_render(illume_wavefront *wf, const illume_scene *scene, const illume_camera *camera,
                            const illume_render_cfg *cfg, illume_guide *guide, const illume_guide_cfg *gcfg,
                            float *rgb_sum, uint64_t *ray_counts, void *stream);
int illume_wavefront_render_batch(illume_wavefront *wf, const illume_job *jobs, uint32_t n_jobs, illume_guide *guide,
                                  const illume_guide_cfg *gcfg, float *film, uint64_t *ray_counts, void *stream);
/* illume_wavefront_render_batch with a proposal network per job: guides[j]
 * (null: no proposal for job j) guides job j's paths and, with gcfg->train,
 * is trained on job j's radiance samples only (each network on the samples of
 * the jobs that name it). Jobs may share a network; at most 16 distinct
 * networks per call (ILLUME_ERR_UNSUPPORTED beyond). One launch sequence for
 * the whole batch: one proposal stage infers every network's rows. With every
 * guides[j] equal to one guide this is illume_wavefront_render_batch. */
int illume_wavefront_render_batch_guides(illume_wavefront *wf, const illume_job *jobs, uint32_t n_jobs,
                                         illume_guide *const *guides, const illume_guide_cfg *gcfg, float *film,
                                         uint64_t *ray_counts, void *stream);
/* The filtered film of illume_render_filtered on the wavefront pool: job j's
 * camera samples cover filters[j]'s sample bounds and are splatted into its
 * image of width*height*4 floats (r, g, b, weight sums) at film_offset;
 * illume_film_resolve turns each image into RGB. Samplers as in
 * illume_render_filtered (cfg.flags). Film::add_sample through
 * Film::get_sample_bounds (core/film.rs:83-105, 125-185). */
int illume_wavefront_render_batch_filtered(illume_wavefront *wf, const illume_job *jobs,
                                           const illume_filter *const *filters, uint32_t n_jobs, illume_guide *guide,
                                           const illume_guide_cfg *gcfg, float *film, uint64_t *ray_counts,
                                           void *stream);
int illume_wavefront_render_filtered(illume_wavefront *wf, const illume_scene *scene, const illume_camera *camera,
                                     const illume_render_cfg *cfg, const illume_filter *filter, illume_guide *guide,
                                     const illume_guide_cfg *gcfg, float *rgbw_sum, uint64_t *ray_counts,
                                     void *stream);

/* Human-readable name of a status code. */
const char *illume_status_string(int status);

/* Message of the last failed call on this host thread ("" if none). */
const char *illume_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ILLUME_HIP_H */
