// Path-slot pool of the wavefront path tracer (structure of arrays in HBM).
#pragma once
#include <hip/hip_runtime.h>
#include "path.h"

namespace illume {

struct WfState {
    float4 *o_tmax, *d_eta, *beta, *L;
    ulonglong2 *rng;   // PCG32 state, increment
    uint2 *flags;      // bounces, bit0 active | bit1 specular bounce
    uint32_t *pix;
    float4 *hit0, *hit1;
    int32_t *guide_slot;     // row of the proposal queue, -1 if none
    float *guide_in;         // [capacity][6]
    uint32_t *guide_count;
    VertexRec *recs;         // [capacity][max_depth], null when not training
    uint32_t *n_recs;
    uint32_t max_depth;
};

hipError_t wf_launch_start(const WfState &w, const DevCamera &cam, uint32_t n_paths, uint64_t seed,
                           uint64_t first_path, hipStream_t st);
hipError_t wf_launch_extend(const WfState &w, const DevScene &sc, uint32_t n_paths, bool guided, const GuideCfg &gc,
                            unsigned long long *counters, hipStream_t st);
hipError_t wf_launch_shade(const WfState &w, const DevScene &sc, uint32_t n_paths, const GuideCfg &gc, float *rgb,
                           unsigned long long *counters, hipStream_t st);
hipError_t wf_launch_records(const WfState &w, uint32_t n_paths, float *out, uint32_t *out_count, uint32_t capacity,
                             hipStream_t st);

}  // namespace illume
