// Path-slot pool of the wavefront path tracer (structure of arrays in HBM)
// and its device-side work counters.
#pragma once
#include <hip/hip_runtime.h>
#include "film.h"
#include "guide_kernels.h"
#include "path.h"

namespace illume {

// Proposal networks one batch may use (one per job at most; jobs may share).
constexpr int kMaxNets = 16;

// One render job: a scene, a camera and a range of sample indices, writing to
// its own region of the film buffer. The paths of all jobs form one global
// index space: job j owns [path_begin, path_begin + width*height*spp).
struct WfJob {
    const DevScene *scene;  // device pointer
    DevCamera cam;
    unsigned long long path_begin;
    unsigned long long film_offset;  // float offset of the job's image in the film buffer
    unsigned long long pix_begin;    // first pixel of the job among the batch's unfiltered pixels
    uint64_t seed;
    uint32_t spp, sample_offset, max_depth;
    uint32_t filtered;  // 1: camera samples over fs's sample bounds, splatted into 4-float pixels
    GuideCfg gc;
    SamplerLaunch sl;   // low-discrepancy sampler of the job (table null: random streams)
    FilmSplat fs;
};

// Batch-wide counters, each on a 256-byte line of its own (the per-tile
// counts live in WfState's tile arrays).
constexpr int kCountStride = 64;  // uint32 stride between the networks' proposal-row counts
struct WfCounters {
    unsigned long long next_path, total_paths;
    alignas(256) unsigned long long done_paths;
    alignas(256) uint32_t guide_count[kMaxNets][kCountStride];  // rows of network k: guide_count[k][0]
    alignas(256) unsigned long long rec_count[kMaxNets][kCountStride / 2];  // training samples of network k
};

// Live paths, structure of arrays. The pool is cut into tiles of `tile`
// slots; tile t (slots t * tile ..) belongs to block t of every stage kernel
// for the whole batch, holds its live paths at its first positions, and the
// shade stage writes the survivors compacted into the same tile of the other
// buffer. No stage uses a batch-wide queue, so no counter is shared by more
// than one block per iteration, and a tile's state is touched by the same
// XCD in every stage.
struct WfPaths {
    float4 *o_tmax;   // ray origin, t_max
    float4 *d_eta;    // ray direction, eta scale
    float4 *beta;     // throughput; w (bits): bounces | specular << 8 | training vertices << 16
    float4 *L;        // radiance; w (bits): job index
    ulonglong2 *rng;  // sampler state, increment
    uint4 *meta;      // sample pixel, film x, film y (bits), training-record slot (bit 31: records off)
};

struct WfState {
    WfPaths s[2];
    float4 *hit0, *hit1;     // closest hit of slot p: prim, slot, t, b0 / b1, b2, proposal row (bits)
    float4 *gstage;          // [2 p], [2 p + 1]: slot p's proposal inputs (position, normal)
    // paths that ended this iteration, from each tile's first slot
    float4 *fin_L;           // radiance; w (bits): job index
    uint4 *fin_meta;         // as WfPaths::meta
    uint32_t *fin_nrec;      // training vertices
    // direct-lighting requests (DeferredDirect), from each tile's first slot:
    // [0] shadow origin, t_max  [1] shadow direction, light-selection pdf
    // [2] shadow term, flags (bit 0 shadow ray, bit 1 BSDF-sample ray) | light << 2
    // [3] path throughput, destination (bits: slot in the next state buffer,
    //     or bit 31 | slot in the finish buffer)
    // [4] BSDF-sample origin, MIS weight  [5] its direction, pdf  [6] f |cos|
    float4 *nee[7];
    uint32_t *tile_live[2];  // per tile: live paths in buffer b
    uint32_t *tile_nee;      // per tile: direct-lighting requests of this iteration
    uint32_t *tile_fin;      // per tile: paths that ended this iteration
    uint32_t *tile_free;     // per tile: free training-record slots
    uint32_t *rec_free;      // per tile: its free record slots (a stack from the tile's first slot)
    uint32_t *guide_slot;    // [n_nets][capacity]: path slot of each proposal row; network k's rows start at row k * capacity
    __half *guide_out;       // [n_nets][capacity][40], fp16
    const GuideDev *nets;    // device array of the batch's networks
    uint32_t n_nets;
    VertexRec *recs;         // [capacity][max_depth] by record slot, null when not training
    float *train_rec;        // [train_capacity][10]: network k's records from k * rec_cap_net
    uint32_t train_capacity;
    uint32_t rec_cap_net;    // record capacity of each network's region
    uint32_t capacity, max_depth;
    uint32_t tile, n_tiles;  // slots per tile (4 x block), tiles
    uint32_t regen_thr;      // a tile starts new paths when fewer than this many are live
    uint32_t rec_thresh;     // with records on, path g records training vertices iff hash(g) <= rec_thresh
    float4 *smp;             // radiance of path g of an unfiltered job at smp[g] (null: added to the film
                             // with atomics when the path ends)
    const WfJob *jobs;
    uint32_t n_jobs;
    uint32_t refill;         // extend refills its lanes (the batch holds a scene of >= kRefillPrims primitives)
    WfCounters *ctr;
};

// Scenes from this many primitives up trace the continuation rays with lane
// refill: on the 64-scene batch (0.25 M triangles a scene) extend's lanes
// were 19 of 64 active and the render 0.266 -> 0.237 s with refill, while the
// Cornell box (12 primitives, rays of a few steps) lost 5-7% to it.
constexpr uint32_t kRefillPrims = 1u << 12;

// Resets the counters and tiles for a new batch of total_paths paths (every
// tile empty); w.tile and w.n_tiles set for `block`.
hipError_t wf_launch_reset(const WfState &w, unsigned long long total_paths, int block, hipStream_t st);
// Clears the training-record counter of network `net`, or of every network
// for net < 0 (a kernel, so it may sit between graph replays).
hipError_t wf_launch_clear_records(const WfState &w, int net, hipStream_t st);
// Sums the sample buffer into the film: every pixel of every unfiltered job
// (n_pix of them in all) gets its samples in sample order.
hipError_t wf_launch_film_reduce(const WfState &w, unsigned long long n_pix, float *film, hipStream_t st);
// One iteration reading state buffer `parity`: regenerate, extend, [infer],
// shade, direct lighting, finish.
hipError_t wf_launch_iteration(const WfState &w, uint32_t parity, bool guided, float *film,
                               unsigned long long *ray_counts, int block, hipStream_t st);

}  // namespace illume
