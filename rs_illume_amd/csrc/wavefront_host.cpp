// illume_wavefront_* entry points: the path-slot pool, the job table of a
// (batched) render, the host loop that replays hipGraph-captured iterations
// of the wavefront path tracer, and the online training of the proposal.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <tuple>
#include <vector>
#include "../../include/illume_hip.h"
#include "guide_kernels.h"
#include "kernels.h"
#include "wavefront.h"

namespace illume {
int abi_fail(int status, const std::string &msg);
int abi_hip_fail(hipError_t e, const char *what);
const DevScene *scene_dev(const illume_scene *s);
const DevScene *scene_dev_device(const illume_scene *s);
uint32_t scene_prims(const illume_scene *s);
void scene_bounds(const illume_scene *s, float *b);
bool scene_camera(const illume_camera &c, DevCamera *out);
void filter_splat(const illume_filter *f, uint32_t width, uint32_t height, FilmSplat *out);
const GuideDev *guide_dev(const illume_guide *g);
int guide_train_strided(illume_guide *g, const float *records, uint32_t n, uint32_t stride, hipStream_t st);
}  // namespace illume

using namespace illume;

namespace {
// A captured graph bakes in its kernels' arguments: the film and counter
// pointers, the record buffer (training on or off) and the proposal's device
// arrays (the whole GuideDev by value, so a new guide that reuses an old one's
// host address never replays a graph holding freed weights).
typedef std::tuple<bool, bool, std::string, float *, unsigned long long *, int, uint32_t> GraphKey;
}

struct illume_wavefront {
    uint32_t capacity = 0, max_depth = 0;
    uint32_t block = 128, regen_min = 0, graph_iters = 8;
    WfState w;
    WfJob *d_jobs = nullptr;
    uint32_t jobs_capacity = 0;
    GuideDev *d_nets = nullptr;  // kMaxNets entries: the networks of the current batch
    uint32_t nets_alloc = 1;     // networks the proposal queue (guide_slot, guide_out) has room for
    std::vector<void *> allocs;
    std::map<GraphKey, hipGraphExec_t> graphs;
    hipStream_t own_stream = nullptr;  // used when the caller passes the null stream (not capturable)
    uint64_t iterations = 0;  // of the last render call
    float4 *smp = nullptr;    // per-path radiance of unfiltered batches
    unsigned long long smp_capacity = 0;
};

namespace {
// Pool sizes are whole multiples of the largest tile (4 x 256 slots); the
// per-tile count arrays have room for the smallest tiles (4 x 64: 16 per
// quantum).
constexpr uint32_t kTileQuantum = 1024;
// Largest batch with a sample buffer: 2^29 paths, 8 GiB (path indices fit the
// 32-bit field they travel in).
constexpr unsigned long long kSampleBufferPaths = 1ull << 29;
template <typename T>
hipError_t walloc(illume_wavefront *wf, T **p, size_t count) {
    hipError_t e = hipMalloc(p, count * sizeof(T));
    if (e == hipSuccess) wf->allocs.push_back(*p);
    return e;
}
void drop_graphs(illume_wavefront *wf) {
    for (auto &kv : wf->graphs) hipGraphExecDestroy(kv.second);
    wf->graphs.clear();
}
}  // namespace

extern "C" {

int illume_wavefront_destroy(illume_wavefront *wf) {
    if (!wf) return ILLUME_OK;
    drop_graphs(wf);
    if (wf->own_stream) hipStreamDestroy(wf->own_stream);
    for (void *p : wf->allocs) hipFree(p);
    delete wf;
    return ILLUME_OK;
}

int illume_wavefront_create(uint32_t max_paths, uint32_t max_depth, illume_wavefront **out) {
    if (!out) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (max_paths < 64 || max_paths > (1u << 26)) return abi_fail(ILLUME_ERR_UNSUPPORTED, "max_paths in 64..2^26");
    if (max_depth == 0 || max_depth > 64) return abi_fail(ILLUME_ERR_UNSUPPORTED, "max_depth in 1..64");
    illume_wavefront *wf = new (std::nothrow) illume_wavefront();
    if (!wf) return abi_fail(ILLUME_ERR_OUT_OF_MEMORY, "host allocation failed");
    // a whole number of tiles for every block size (tiles of 4 x block slots)
    max_paths = (max_paths + kTileQuantum - 1) / kTileQuantum * kTileQuantum;
    wf->capacity = max_paths;
    wf->max_depth = max_depth;
    wf->regen_min = max_paths + 1;  // regenerate every iteration
    WfState &w = wf->w;
    std::memset(&w, 0, sizeof(w));
    const size_t P = max_paths;
    hipError_t e = hipSuccess;
    for (int b = 0; b < 2; ++b) {
        WfPaths &s = w.s[b];
        if (e == hipSuccess) e = walloc(wf, &s.o_tmax, P);
        if (e == hipSuccess) e = walloc(wf, &s.d_eta, P);
        if (e == hipSuccess) e = walloc(wf, &s.beta, P);
        if (e == hipSuccess) e = walloc(wf, &s.L, P);
        if (e == hipSuccess) e = walloc(wf, &s.rng, P);
        if (e == hipSuccess) e = walloc(wf, &s.meta, P);
    }
    if (e == hipSuccess) e = walloc(wf, &w.hit0, P);
    if (e == hipSuccess) e = walloc(wf, &w.hit1, P);
    if (e == hipSuccess) e = walloc(wf, &w.gstage, 2 * P);
    if (e == hipSuccess) e = walloc(wf, &w.fin_L, P);
    if (e == hipSuccess) e = walloc(wf, &w.fin_meta, P);
    if (e == hipSuccess) e = walloc(wf, &w.fin_nrec, P);
    for (int k = 0; k < 7; ++k)
        if (e == hipSuccess) e = walloc(wf, &w.nee[k], P);
    if (e == hipSuccess) e = walloc(wf, &w.guide_slot, P);
    if (e == hipSuccess) e = walloc(wf, &w.guide_out, 40 * P);
    if (e == hipSuccess) e = walloc(wf, &w.recs, P * max_depth);
    if (e == hipSuccess) e = walloc(wf, &w.rec_free, P);
    for (int b = 0; b < 2; ++b)
        if (e == hipSuccess) e = walloc(wf, &w.tile_live[b], P / kTileQuantum * 16);
    if (e == hipSuccess) e = walloc(wf, &w.tile_nee, P / kTileQuantum * 16);
    if (e == hipSuccess) e = walloc(wf, &w.tile_fin, P / kTileQuantum * 16);
    if (e == hipSuccess) e = walloc(wf, &w.tile_free, P / kTileQuantum * 16);
    w.train_capacity = (uint32_t)std::min<size_t>(P * max_depth, 1u << 25);
    if (e == hipSuccess) e = walloc(wf, &w.train_rec, 10 * (size_t)w.train_capacity);
    if (e == hipSuccess) e = walloc(wf, &wf->d_nets, kMaxNets);
    if (e == hipSuccess) e = walloc(wf, &w.ctr, 1);
    if (e == hipSuccess) e = hipMemset(w.ctr, 0, sizeof(WfCounters));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&wf->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        illume_wavefront_destroy(wf);
        return abi_hip_fail(e, "wavefront allocation");
    }
    w.capacity = max_paths;
    w.max_depth = max_depth;
    *out = wf;
    return ILLUME_OK;
}

int illume_wavefront_set_tuning(illume_wavefront *wf, uint32_t block, uint32_t regen_min, uint32_t graph_iters) {
    if (!wf) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    if (block != 64 && block != 128 && block != 256) return abi_fail(ILLUME_ERR_UNSUPPORTED, "block in {64,128,256}");
    if (graph_iters > 64) return abi_fail(ILLUME_ERR_UNSUPPORTED, "graph_iters in 0..64");
    wf->block = block;
    wf->regen_min = std::max(1u, std::min(regen_min, wf->capacity + 1));
    wf->graph_iters = graph_iters;
    drop_graphs(wf);
    return ILLUME_OK;
}

uint64_t illume_wavefront_last_iterations(const illume_wavefront *wf) { return wf ? wf->iterations : 0; }

}  // extern "C"

namespace {

// Iterations per replay: they alternate the state buffer, so a replay of an
// even number of them starts and ends on buffer 0.
uint32_t k_iters_per_replay(const illume_wavefront *wf) {
    return wf->graph_iters > 0 ? (wf->graph_iters + 1) & ~1u : 2u;
}

// Grows the proposal queue to n_nets networks' segments (capacity rows each).
hipError_t ensure_net_rows(illume_wavefront *wf, uint32_t n_nets) {
    if (n_nets <= wf->nets_alloc) return hipSuccess;
    WfState &w = wf->w;
    for (void *p : {(void *)w.guide_slot, (void *)w.guide_out}) {
        hipFree(p);
        wf->allocs.erase(std::remove(wf->allocs.begin(), wf->allocs.end(), p), wf->allocs.end());
    }
    w.guide_slot = nullptr;
    w.guide_out = nullptr;
    wf->nets_alloc = 0;
    drop_graphs(wf);
    hipError_t e = walloc(wf, &w.guide_slot, (size_t)w.capacity * n_nets);
    if (e == hipSuccess) e = walloc(wf, &w.guide_out, 40 * (size_t)w.capacity * n_nets);
    if (e == hipSuccess) wf->nets_alloc = n_nets;
    return e;
}

// The batch render behind the entry points; filters == null: every job
// unfiltered (3 floats per pixel), else filters[j] (non-null) for job j
// (4 floats per pixel, splatted). guides == null: `guide` (may be null) for
// every job; else guides[j] (null: no proposal) for job j.
int render_batch(illume_wavefront *wf, const illume_job *jobs, const illume_filter *const *filters, uint32_t n_jobs,
                 illume_guide *guide, illume_guide *const *guides, const illume_guide_cfg *gcfg, float *film,
                 uint64_t *ray_counts, void *stream) {
    if (!wf || !film || (n_jobs && !jobs)) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    if (gcfg && !(gcfg->guide_frac >= 0.f && gcfg->guide_frac <= 1.f))
        return abi_fail(ILLUME_ERR_INVALID_ARG, "guide_frac must be in [0, 1]");
    if (gcfg && !(gcfg->min_alpha >= 0.f)) return abi_fail(ILLUME_ERR_INVALID_ARG, "min_alpha must be >= 0");
    // the batch's distinct networks in job order; job j uses net_of[j] (-1: none)
    std::vector<illume_guide *> nets;
    std::vector<int> net_of(n_jobs, -1);
    for (uint32_t j = 0; j < n_jobs; ++j) {
        illume_guide *g = guides ? guides[j] : guide;
        if (!g) continue;
        auto it = std::find(nets.begin(), nets.end(), g);
        if (it == nets.end()) {
            if (nets.size() == (size_t)kMaxNets)
                return abi_fail(ILLUME_ERR_UNSUPPORTED, "at most 16 distinct proposal networks per batch");
            nets.push_back(g);
            net_of[j] = (int)nets.size() - 1;
        } else {
            net_of[j] = (int)(it - nets.begin());
        }
    }
    const bool guided = !nets.empty() && gcfg && gcfg->guide_frac > 0.f;
    const bool train = !nets.empty() && gcfg && gcfg->train;
    const uint32_t batch = (gcfg && gcfg->train_batch) ? gcfg->train_batch : 65536u;
    // Work already queued by the caller (e.g. zeroing the film) completes
    // first; the pipeline then runs on the caller's stream, or on the pool's
    // own stream when the caller passes the (uncapturable) null stream.
    hipStream_t caller = (hipStream_t)stream;
    hipStream_t st = caller ? caller : wf->own_stream;
    std::vector<WfJob> hj;
    unsigned long long total = 0, n_pix_all = 0;
    uint32_t max_prims = 0;
    for (uint32_t j = 0; j < n_jobs; ++j) {
        const illume_job &in = jobs[j];
        if (!in.scene) return abi_fail(ILLUME_ERR_INVALID_ARG, "job without a scene");
        if (in.cfg.max_depth > wf->max_depth) return abi_fail(ILLUME_ERR_UNSUPPORTED, "max_depth above the pool's");
        const uint32_t kind = sampler_kind(in.cfg.flags);
        if (kind == ~0u) return abi_fail(ILLUME_ERR_INVALID_ARG, "unknown render flags");
        if ((uint64_t)in.cfg.sample_offset + in.cfg.spp > (1u << 24))
            return abi_fail(ILLUME_ERR_UNSUPPORTED, "sample index beyond 2^24");
        if (filters && !filters[j]) return abi_fail(ILLUME_ERR_INVALID_ARG, "job without a filter");
        WfJob job;
        std::memset(&job, 0, sizeof(job));
        if (!scene_camera(in.camera, &job.cam)) return abi_fail(ILLUME_ERR_INVALID_ARG, "invalid camera");
        unsigned long long n_pix = (unsigned long long)job.cam.width * job.cam.height;
        int x0 = 0, y0 = 0, sw = (int)job.cam.width, sh = (int)job.cam.height;
        if (filters) {
            if ((uint64_t)job.cam.width + 64 >= (1u << 24) || (uint64_t)job.cam.height + 64 >= (1u << 24))
                return abi_fail(ILLUME_ERR_UNSUPPORTED, "image side beyond 2^24 - 64");
            filter_splat(filters[j], job.cam.width, job.cam.height, &job.fs);
            job.filtered = 1;
            n_pix = (unsigned long long)job.fs.wext * job.fs.hext;
            x0 = job.fs.x0;
            y0 = job.fs.y0;
            sw = (int)job.fs.wext;
            sh = (int)job.fs.hext;
        }
        const unsigned long long n = scene_prims(in.scene) ? n_pix * in.cfg.spp : 0ull;
        if (n == 0) continue;
        max_prims = std::max(max_prims, scene_prims(in.scene));
        if (kind && !sampler_launch(kind, x0, y0, sw, sh, &job.sl))
            return abi_fail(ILLUME_ERR_OUT_OF_MEMORY, "low-discrepancy tables");
        job.scene = scene_dev_device(in.scene);
        job.path_begin = total;
        job.pix_begin = n_pix_all;
        if (!filters) n_pix_all += n_pix;
        job.film_offset = in.film_offset;
        job.seed = in.cfg.seed;
        job.spp = in.cfg.spp;
        job.sample_offset = in.cfg.sample_offset;
        job.max_depth = in.cfg.max_depth;
        float b[6];
        scene_bounds(in.scene, b);
        for (int a = 0; a < 3; ++a) {
            job.gc.bounds_lo[a] = b[a];
            const float ext = b[3 + a] - b[a];
            job.gc.inv_ext[a] = ext > 0.f ? 1.f / ext : 1.f;
        }
        job.gc.frac = guided ? gcfg->guide_frac : 0.f;
        job.gc.min_alpha = gcfg ? gcfg->min_alpha : 0.f;
        job.gc.product = gcfg ? (gcfg->flags & ILLUME_GUIDE_PRODUCT) : 0u;
        job.gc.net = guided || train ? net_of[j] : -1;
        job.gc.out = guided && net_of[j] >= 0 ? wf->w.guide_out : nullptr;
        if (!(guided && net_of[j] >= 0)) job.gc.frac = 0.f;
        hj.push_back(job);
        total += n;
    }
    if (total == 0) return ILLUME_OK;
    hipError_t e = hipSuccess;
    if (hj.size() > wf->jobs_capacity) {
        if (wf->d_jobs) {
            hipFree(wf->d_jobs);
            wf->allocs.erase(std::remove(wf->allocs.begin(), wf->allocs.end(), (void *)wf->d_jobs), wf->allocs.end());
        }
        wf->d_jobs = nullptr;
        wf->jobs_capacity = 0;
        e = walloc(wf, &wf->d_jobs, hj.size());
        if (e != hipSuccess) return abi_hip_fail(e, "job table allocation");
        wf->jobs_capacity = (uint32_t)hj.size();
        drop_graphs(wf);
    }
    // Unfiltered batches of up to kSampleBufferPaths paths write each path's
    // radiance to a sample buffer (one 16-byte store) that one kernel sums into
    // the film at the end, instead of three float atomics per path.
    const bool use_smp = !filters && total <= kSampleBufferPaths;
    if (use_smp && total > wf->smp_capacity) {
        if (wf->smp) {
            hipFree(wf->smp);
            wf->allocs.erase(std::remove(wf->allocs.begin(), wf->allocs.end(), (void *)wf->smp), wf->allocs.end());
        }
        wf->smp = nullptr;
        wf->smp_capacity = 0;
        drop_graphs(wf);
        e = walloc(wf, &wf->smp, total);
        if (e != hipSuccess) return abi_hip_fail(e, "sample buffer allocation");
        wf->smp_capacity = total;
    }
    if (guided && (e = ensure_net_rows(wf, (uint32_t)nets.size())) != hipSuccess)
        return abi_hip_fail(e, "proposal queue allocation");
    // guide_out moved if the queue grew: the jobs point into it
    for (WfJob &job : hj)
        if (job.gc.out) job.gc.out = wf->w.guide_out;
    std::vector<GuideDev> hnets;
    for (illume_guide *g : nets) hnets.push_back(*guide_dev(g));
    e = caller ? hipStreamSynchronize(caller) : hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(wf->d_jobs, hj.data(), hj.size() * sizeof(WfJob), hipMemcpyHostToDevice);
    if (e == hipSuccess && guided)
        e = hipMemcpy(wf->d_nets, hnets.data(), hnets.size() * sizeof(GuideDev), hipMemcpyHostToDevice);
    if (e != hipSuccess) return abi_hip_fail(e, "job table upload");
    WfState w = wf->w;
    w.jobs = wf->d_jobs;
    w.n_jobs = (uint32_t)hj.size();
    w.refill = max_prims >= kRefillPrims ? 1u : 0u;
    if (!train) w.recs = nullptr;
    w.nets = wf->d_nets;
    w.n_nets = (uint32_t)nets.size();
    w.smp = use_smp ? wf->smp : nullptr;
    w.rec_cap_net = w.n_nets ? w.train_capacity / w.n_nets : w.train_capacity;
    unsigned long long *cnt = (unsigned long long *)ray_counts;
    w.tile = 4 * wf->block;
    w.n_tiles = w.capacity / w.tile;
    // a tile regenerates while fewer than its share of regen_min are live
    w.regen_thr = (uint32_t)std::min<uint64_t>(((uint64_t)wf->regen_min * w.tile + w.capacity - 1) / w.capacity,
                                               (uint64_t)w.tile + 1);
    // Training uses at most gcfg->train steps of `batch` records per network
    // and replay of k iterations, while every live path may record a vertex
    // per iteration: only a hash-chosen share of the paths records, sized for
    // eight times the records the steps take (all paths for small batches).
    w.rec_thresh = 0xFFFFFFFFu;
    if (train) {
        const double need = 8.0 * std::max(1u, gcfg->train) * batch * std::max<uint32_t>(1u, (uint32_t)w.n_nets);
        const double made = (double)k_iters_per_replay(wf) * (double)std::min<unsigned long long>(w.capacity, total);
        if (need < made) w.rec_thresh = (uint32_t)(need / made * 4294967295.0);
    }
    if (e == hipSuccess) e = wf_launch_reset(w, total, wf->block, st);
    if (e == hipSuccess && train) e = wf_launch_clear_records(w, -1, st);
    // Every path finishes within max_depth + 1 iterations of its start; the
    // pool restarts at least one path per iteration whenever it is idle.
    const uint64_t max_iters = (total / wf->capacity + 2) * (uint64_t)(wf->max_depth + 2) + total + 16;
    const uint32_t k = k_iters_per_replay(wf);
    // A graph holds the launch arguments it was captured with: the state
    // (pointers, tile shape, record share, sample buffer) and, whenever they are
    // used for inference or training records, the networks.
    std::string gbytes((const char *)&w, sizeof(w));
    if (guided || train) gbytes.append((const char *)hnets.data(), hnets.size() * sizeof(GuideDev));
    GraphKey key(guided, train, gbytes, film, cnt, wf->block, w.n_jobs);
    hipGraphExec_t exec = nullptr;
    if (wf->graph_iters > 0 && e == hipSuccess) {
        if (wf->graphs.size() >= 64) drop_graphs(wf);
        auto it = wf->graphs.find(key);
        if (it != wf->graphs.end()) {
            exec = it->second;
        } else {
            hipGraph_t graph;
            e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
            for (uint32_t i = 0; i < k && e == hipSuccess; ++i)
                e = wf_launch_iteration(w, i & 1u, guided, film, cnt, wf->block, st);
            hipError_t e2 = hipStreamEndCapture(st, &graph);
            if (e == hipSuccess) e = e2;
            if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (e2 == hipSuccess) hipGraphDestroy(graph);
            if (e == hipSuccess) wf->graphs[key] = exec;
        }
    }
    WfCounters c;
    uint64_t iters = 0;
    while (e == hipSuccess) {
        if (exec) e = hipGraphLaunch(exec, st);
        else
            for (uint32_t i = 0; i < k && e == hipSuccess; ++i)
                e = wf_launch_iteration(w, i & 1u, guided, film, cnt, wf->block, st);
        iters += k;
        if (e == hipSuccess) e = hipMemcpyAsync(&c, w.ctr, sizeof(c), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) break;
        const bool finished = c.done_paths >= c.total_paths;
        for (uint32_t nk = 0; train && nk < w.n_nets && e == hipSuccess; ++nk) {
            const unsigned long long rc = c.rec_count[nk][0];
            if (!((rc >= batch || finished) && rc > 0)) continue;
            // up to gcfg->train Adam steps per replay for network nk, on its
            // own records; step s reads the records s, s + S*q, s + 2*S*q,
            // ... (S steps, q = spacing), so the steps see disjoint, evenly
            // spread subsets
            const uint32_t n_rec = (uint32_t)std::min<unsigned long long>(rc, w.rec_cap_net);
            const uint32_t rows = std::min(batch, n_rec);
            const uint32_t steps = std::max(1u, std::min(gcfg->train, n_rec / rows));
            const uint32_t q = n_rec / (rows * steps);
            const float *region = w.train_rec + 10 * (size_t)nk * w.rec_cap_net;
            for (uint32_t s = 0; s < steps; ++s) {
                int r = guide_train_strided(nets[nk], region + 10 * (size_t)s * q, rows, steps * q, st);
                if (r != ILLUME_OK) return r;
            }
            e = wf_launch_clear_records(w, (int)nk, st);
        }
        if (finished) break;
        if (iters > max_iters) {
            wf->iterations = iters;
            return abi_fail(ILLUME_ERR_QUEUE_OVERFLOW, "wavefront made no progress");
        }
    }
    wf->iterations = iters;
    if (e == hipSuccess && use_smp) e = wf_launch_film_reduce(w, n_pix_all, film, st);
    // on the pool's own stream (the caller passed the null stream) the film is
    // complete when the call returns
    if (e == hipSuccess && !caller) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return abi_hip_fail(e, "wavefront render");
    return ILLUME_OK;
}

}  // namespace

extern "C" {

int illume_wavefront_render_batch_guides(illume_wavefront *wf, const illume_job *jobs, uint32_t n_jobs,
                                         illume_guide *const *guides, const illume_guide_cfg *gcfg, float *film,
                                         uint64_t *ray_counts, void *stream) {
    if (n_jobs && !guides) return abi_fail(ILLUME_ERR_INVALID_ARG, "guides is null");
    return render_batch(wf, jobs, nullptr, n_jobs, nullptr, guides, gcfg, film, ray_counts, stream);
}

int illume_wavefront_render_batch(illume_wavefront *wf, const illume_job *jobs, uint32_t n_jobs, illume_guide *guide,
                                  const illume_guide_cfg *gcfg, float *film, uint64_t *ray_counts, void *stream) {
    return render_batch(wf, jobs, nullptr, n_jobs, guide, nullptr, gcfg, film, ray_counts, stream);
}

int illume_wavefront_render_batch_filtered(illume_wavefront *wf, const illume_job *jobs,
                                           const illume_filter *const *filters, uint32_t n_jobs, illume_guide *guide,
                                           const illume_guide_cfg *gcfg, float *film, uint64_t *ray_counts,
                                           void *stream) {
    if (n_jobs && !filters) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    return render_batch(wf, jobs, filters, n_jobs, guide, nullptr, gcfg, film, ray_counts, stream);
}

int illume_wavefront_render(illume_wavefront *wf, const illume_scene *scene, const illume_camera *camera,
                            const illume_render_cfg *cfg, illume_guide *guide, const illume_guide_cfg *gcfg,
                            float *rgb_sum, uint64_t *ray_counts, void *stream) {
    if (!scene || !camera || !cfg) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    illume_job job;
    std::memset(&job, 0, sizeof(job));
    job.scene = scene;
    job.camera = *camera;
    job.cfg = *cfg;
    job.film_offset = 0;
    return render_batch(wf, &job, nullptr, 1, guide, nullptr, gcfg, rgb_sum, ray_counts, stream);
}

int illume_wavefront_render_filtered(illume_wavefront *wf, const illume_scene *scene, const illume_camera *camera,
                                     const illume_render_cfg *cfg, const illume_filter *filter, illume_guide *guide,
                                     const illume_guide_cfg *gcfg, float *rgbw_sum, uint64_t *ray_counts,
                                     void *stream) {
    if (!scene || !camera || !cfg || !filter) return abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    illume_job job;
    std::memset(&job, 0, sizeof(job));
    job.scene = scene;
    job.camera = *camera;
    job.cfg = *cfg;
    job.film_offset = 0;
    return render_batch(wf, &job, &filter, 1, guide, nullptr, gcfg, rgbw_sum, ray_counts, stream);
}

}  // extern "C"
