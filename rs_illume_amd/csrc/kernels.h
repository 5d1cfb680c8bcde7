// Internal launchers shared between the C ABI and the kernel sources.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/illume_hip.h"
#include "film.h"
#include "lowdisc.h"
#include "layout.h"
#include "scene_types.h"

namespace illume {

// Path tracer: adds the radiance sums of spp paths per pixel to rgb.
// hl.table null: random sampler streams; else the Halton or Sobol' sampler.
hipError_t launch_render(const DevScene &sc, const DevCamera &cam, uint32_t spp, uint32_t max_depth, uint64_t seed,
                         uint32_t sample_offset, float *rgb, unsigned long long *counters, const SamplerLaunch &hl,
                         hipStream_t stream);

// Ambient occlusion (AOIntegrator::li): adds the AO sums of spp camera
// samples per pixel, n_ao hemisphere directions each, to rgb (grey).
hipError_t launch_render_ao(const DevScene &sc, const DevCamera &cam, uint32_t spp, uint32_t n_ao, bool cos_sample,
                            uint64_t seed, uint32_t sample_offset, float *rgb, unsigned long long *counters,
                            const SamplerLaunch &hl, hipStream_t stream);

// SamplerIntegrator::li for n caller-given rays (stream i, sample `sample`,
// the camera sample's five dimensions skipped): 3 floats per ray in L_out.
hipError_t launch_li_batch(const DevScene &sc, const illume_ray *rays, uint32_t n, uint32_t max_depth, uint64_t seed,
                           uint32_t sample, float *L_out, unsigned long long *counters, hipStream_t stream);

// Path tracer through a reconstruction filter: samples over fs's sample
// bounds, splatted into rgbw (4 floats per image pixel).
hipError_t launch_render_filtered(const DevScene &sc, const DevCamera &cam, const FilmSplat &fs, uint32_t spp,
                                  uint32_t max_depth, uint64_t seed, uint32_t sample_offset, float *rgbw,
                                  unsigned long long *counters, const SamplerLaunch &hl, hipStream_t stream);

// One image of a batched megakernel render (illume_render_batch).
struct RenderJob {
    const DevScene *scene;  // device copy of the scene record
    DevCamera cam;
    unsigned long long path_begin;   // first path number of the job in the batch
    unsigned long long film_offset;  // float offset of its 3-float pixels
    uint64_t seed;
    uint32_t sample_offset, max_depth;
    SamplerLaunch sl;
};
// Renders every job's width * height * spp paths (jobs ordered by path_begin,
// total paths in all): one thread per path, each finding its job by binary
// search.
hipError_t launch_render_batch(const RenderJob *jobs, uint32_t n_jobs, unsigned long long total, float *film,
                               unsigned long long *counters, hipStream_t stream);

// The sampler parameters of kind kSamplerHalton or kSamplerSobol for the
// sample bounds [x0, x0 + w) x [y0, y0 + h), over this device's table
// (uploaded on first use); false if the table could not be made.
bool sampler_launch(uint32_t kind, int x0, int y0, int w, int h, SamplerLaunch *out);
// Sampler kind of illume_render_cfg.flags: 0 (random streams), kSamplerHalton,
// kSamplerSobol, or ~0u for an invalid combination.
uint32_t sampler_kind(uint32_t flags);

// rgb[3p + c] = scale * (w != 0 ? max(0, rgbw[4p + c] * (1 / w)) : rgbw[4p + c]).
hipError_t launch_resolve(const float *rgbw, uint64_t n, float scale, float *rgb, hipStream_t stream);

// Deepest tree the traversal stack supports (LDS ring + private spill).
int max_supported_depth();

// Largest ray count of one traversal call: grid sizes and the queue's run
// ends are computed in 32 bits and stay exact up to here.
constexpr uint32_t kMaxRaysPerCall = 1u << 31;

// Closest hit (hits), any hit (occluded) or counting traversal (counts != null).
// `empty` = the BVH has no primitives: outputs are filled with misses.
// nodes: GpuNode (width 2) or GpuNode4Q (width 4) array. n <= kMaxRaysPerCall.
// The persistent launches (closest and any hit on four-child trees, any hit
// on two-child trees) allocate their 8-byte queue counter stream-ordered.
hipError_t launch_traverse(const void *nodes, int width, const PrimRecord *prims, bool empty, const illume_ray *rays,
                           uint32_t n, illume_hit *hits, uint8_t *occluded, uint32_t *counts, bool any,
                           hipStream_t stream);

// Words of the device-resident traversal scratch (256 bytes; word 0 is the
// persistent launch's queue counter).
constexpr int kTraverseScratchWords = 32;

// Closest hit like launch_traverse, allocating nothing: the persistent
// launch's queue counter is word 0 of the library's per-device scratch array
// (address looked up once per device). Concurrent calls on one device race on
// that word. n <= kMaxRaysPerCall.
hipError_t launch_closest_noalloc(const void *nodes, int width, const PrimRecord *prims, bool empty,
                                  const illume_ray *rays, uint32_t n, illume_hit *hits, hipStream_t stream);

}  // namespace illume
