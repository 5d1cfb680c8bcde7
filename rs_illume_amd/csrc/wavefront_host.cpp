// illume_wavefront_* entry points: the path-slot pool and the host loop of
// the wavefront path tracer (start, then per bounce extend / infer / shade,
// then training records and Adam steps of the learned proposal).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include "../../include/illume_hip.h"
#include "guide_kernels.h"
#include "wavefront.h"

namespace illume {
int abi_fail(int status, const std::string &msg);
int abi_hip_fail(hipError_t e, const char *what);
const DevScene *scene_dev(const illume_scene *s);
uint32_t scene_prims(const illume_scene *s);
void scene_bounds(const illume_scene *s, float *b);
bool scene_camera(const illume_camera &c, DevCamera *out);
const GuideDev *guide_dev(const illume_guide *g);
}  // namespace illume

using namespace illume;

struct illume_wavefront {
    const illume_scene *scene = nullptr;
    uint32_t capacity = 0, max_depth = 0;
    WfState w;
    float *guide_out = nullptr;
    float *train_rec = nullptr;
    uint32_t *train_count = nullptr;
    uint32_t train_capacity = 0;
    std::vector<void *> allocs;
};

namespace {
template <typename T>
hipError_t walloc(illume_wavefront *wf, T **p, size_t count) {
    hipError_t e = hipMalloc(p, count * sizeof(T));
    if (e == hipSuccess) wf->allocs.push_back(*p);
    return e;
}
}  // namespace

extern "C" {

int illume_wavefront_destroy(illume_wavefront *wf) {
    if (!wf) return ILLUME_OK;
    for (void *p : wf->allocs) hipFree(p);
    delete wf;
    return ILLUME_OK;
}

int illume_wavefront_create(const illume_scene *scene, uint32_t max_paths, uint32_t max_depth,
                            illume_wavefront **out) {
    if (!scene || !out) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (max_paths == 0 || max_paths > (1u << 26)) return abi_fail(ILLUME_ERR_UNSUPPORTED, "max_paths in 1..2^26");
    if (max_depth == 0 || max_depth > 64) return abi_fail(ILLUME_ERR_UNSUPPORTED, "max_depth in 1..64");
    illume_wavefront *wf = new (std::nothrow) illume_wavefront();
    if (!wf) return abi_fail(ILLUME_ERR_OUT_OF_MEMORY, "host allocation failed");
    wf->scene = scene;
    wf->capacity = max_paths;
    wf->max_depth = max_depth;
    WfState &w = wf->w;
    std::memset(&w, 0, sizeof(w));
    const size_t P = max_paths;
    hipError_t e = walloc(wf, &w.o_tmax, P);
    if (e == hipSuccess) e = walloc(wf, &w.d_eta, P);
    if (e == hipSuccess) e = walloc(wf, &w.beta, P);
    if (e == hipSuccess) e = walloc(wf, &w.L, P);
    if (e == hipSuccess) e = walloc(wf, &w.rng, P);
    if (e == hipSuccess) e = walloc(wf, &w.flags, P);
    if (e == hipSuccess) e = walloc(wf, &w.pix, P);
    if (e == hipSuccess) e = walloc(wf, &w.hit0, P);
    if (e == hipSuccess) e = walloc(wf, &w.hit1, P);
    if (e == hipSuccess) e = walloc(wf, &w.guide_slot, P);
    if (e == hipSuccess) e = walloc(wf, &w.guide_in, 6 * P);
    if (e == hipSuccess) e = walloc(wf, &w.guide_count, 64);
    if (e == hipSuccess) e = walloc(wf, &w.recs, P * max_depth);
    if (e == hipSuccess) e = walloc(wf, &w.n_recs, P);
    if (e == hipSuccess) e = walloc(wf, &wf->guide_out, 40 * P);
    wf->train_capacity = (uint32_t)std::min<size_t>(P * max_depth, 1u << 26);
    if (e == hipSuccess) e = walloc(wf, &wf->train_rec, 10 * (size_t)wf->train_capacity);
    if (e == hipSuccess) e = walloc(wf, &wf->train_count, 64);
    if (e == hipSuccess) e = hipMemset(w.guide_count, 0, 64 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(wf->train_count, 0, 64 * sizeof(uint32_t));
    if (e != hipSuccess) {
        illume_wavefront_destroy(wf);
        return abi_hip_fail(e, "wavefront allocation");
    }
    w.max_depth = max_depth;
    *out = wf;
    return ILLUME_OK;
}

int illume_wavefront_render(illume_wavefront *wf, const illume_camera *camera, const illume_render_cfg *cfg,
                            illume_guide *guide, const illume_guide_cfg *gcfg, float *rgb_sum, uint64_t *ray_counts,
                            void *stream) {
    if (!wf || !camera || !cfg || !rgb_sum) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    if (cfg->max_depth > wf->max_depth) return abi_fail(ILLUME_ERR_UNSUPPORTED, "max_depth above the pool's");
    if ((uint64_t)cfg->sample_offset + cfg->spp > (1u << 24))
        return abi_fail(ILLUME_ERR_UNSUPPORTED, "sample index beyond 2^24");
    DevCamera cam;
    if (!scene_camera(*camera, &cam)) return abi_fail(ILLUME_ERR_INVALID_ARG, "invalid camera");
    if (cfg->spp == 0 || scene_prims(wf->scene) == 0) return ILLUME_OK;
    const bool guided = guide && gcfg && gcfg->guide_frac > 0.f;
    const bool train = guide && gcfg && gcfg->train;
    if (gcfg && !(gcfg->guide_frac >= 0.f && gcfg->guide_frac <= 1.f))
        return abi_fail(ILLUME_ERR_INVALID_ARG, "guide_frac must be in [0, 1]");
    const uint32_t batch = (gcfg && gcfg->train_batch) ? gcfg->train_batch : 65536u;
    hipStream_t st = (hipStream_t)stream;
    const DevScene &sc = *scene_dev(wf->scene);
    WfState w = wf->w;
    w.max_depth = cfg->max_depth;
    if (!train) w.recs = nullptr;
    GuideCfg gc;
    std::memset(&gc, 0, sizeof(gc));
    float b[6];
    scene_bounds(wf->scene, b);
    for (int a = 0; a < 3; ++a) {
        gc.bounds_lo[a] = b[a];
        float ext = b[3 + a] - b[a];
        gc.inv_ext[a] = ext > 0.f ? 1.f / ext : 1.f;
    }
    gc.frac = guided ? gcfg->guide_frac : 0.f;
    gc.out = guided ? wf->guide_out : nullptr;
    unsigned long long *cnt = (unsigned long long *)ray_counts;
    const uint64_t n_pix = (uint64_t)cam.width * cam.height;
    const uint64_t first = (uint64_t)cfg->sample_offset * n_pix;
    const uint64_t total = (uint64_t)cfg->spp * n_pix;
    hipError_t e = hipSuccess;
    for (uint64_t done = 0; done < total && e == hipSuccess;) {
        const uint32_t n = (uint32_t)std::min<uint64_t>(wf->capacity, total - done);
        e = wf_launch_start(w, cam, n, cfg->seed, first + done, st);
        for (uint32_t bounce = 0; bounce <= cfg->max_depth && e == hipSuccess; ++bounce) {
            e = wf_launch_extend(w, sc, n, guided, gc, cnt, st);
            if (e == hipSuccess && guided)
                e = guide_launch_infer_count(*guide_dev(guide), w.guide_in, w.guide_count, n, wf->guide_out, st);
            if (e == hipSuccess) e = wf_launch_shade(w, sc, n, gc, rgb_sum, cnt, st);
        }
        if (e == hipSuccess && train) {
            e = wf_launch_records(w, n, wf->train_rec, wf->train_count, wf->train_capacity, st);
            uint32_t n_rec = 0;
            if (e == hipSuccess) e = hipMemcpyAsync(&n_rec, wf->train_count, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            n_rec = std::min(n_rec, wf->train_capacity);
            for (uint32_t s = 0; s < n_rec && e == hipSuccess; s += batch) {
                int r = illume_guide_train_step(guide, wf->train_rec + 10 * (size_t)s, std::min(batch, n_rec - s), st);
                if (r != ILLUME_OK) return r;
            }
            if (e == hipSuccess) e = hipMemsetAsync(wf->train_count, 0, sizeof(uint32_t), st);
        }
        done += n;
    }
    if (e != hipSuccess) return abi_hip_fail(e, "wavefront render");
    return ILLUME_OK;
}

}  // extern "C"
