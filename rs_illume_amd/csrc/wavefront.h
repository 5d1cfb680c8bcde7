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
    uint64_t seed;
    uint32_t spp, sample_offset, max_depth;
    uint32_t filtered;  // 1: camera samples over fs's sample bounds, splatted into 4-float pixels
    GuideCfg gc;
    SamplerLaunch sl;   // low-discrepancy sampler of the job (table null: random streams)
    FilmSplat fs;
};

struct WfCounters {
    uint32_t n_q[2];      // lengths of the two slot queues
    uint32_t n_free;      // free slots
    uint32_t parity;      // queue read this iteration
    uint32_t take;        // slots (re)started by the last regenerate stage
    uint32_t guide_count[kMaxNets];  // rows in each network's segment of the proposal queue
    unsigned long long next_path, total_paths, done_paths;
    unsigned long long rec_count[kMaxNets];  // training samples written for each network
};

struct WfState {
    float4 *o_tmax, *d_eta, *beta, *L;
    ulonglong2 *rng;   // PCG32 state, increment
    uint2 *flags;      // bounces, bit0 active | bit1 specular bounce
    uint32_t *pix, *job;
    float2 *film_xy;   // film position of the path's camera sample (filtered jobs)
    float4 *hit0, *hit1;
    int32_t *guide_slot;     // row of the proposal queue, -1 if none
    float *guide_in;         // [n_nets][capacity][6]: network k's rows start at row k * capacity
    float *guide_out;        // [n_nets][capacity][40]
    const GuideDev *nets;    // device array of the batch's networks
    uint32_t n_nets;
    uint32_t *q[2];          // slot queues
    uint32_t *free_q;        // free slots
    VertexRec *recs;         // [capacity][max_depth], null when not training
    uint32_t *n_recs;
    float *train_rec;        // [train_capacity][10]: network k's records from k * rec_cap_net
    uint32_t train_capacity;
    uint32_t rec_cap_net;    // record capacity of each network's region
    uint32_t capacity, max_depth;
    uint32_t regen_min;      // regenerate only when fewer than this many slots are active
    const WfJob *jobs;
    uint32_t n_jobs;
    WfCounters *ctr;
};

// Resets the counters for a new batch of total_paths paths: all slots free.
hipError_t wf_launch_reset(const WfState &w, unsigned long long total_paths, hipStream_t st);
// Clears the training-record counter of network `net`, or of every network
// for net < 0 (a kernel, so it may sit between graph replays).
hipError_t wf_launch_clear_records(const WfState &w, int net, hipStream_t st);
// One iteration: regenerate, extend, [infer], shade, bookkeeping.
hipError_t wf_launch_iteration(const WfState &w, bool guided, float *film, unsigned long long *ray_counts, int block,
                               hipStream_t st);

}  // namespace illume
