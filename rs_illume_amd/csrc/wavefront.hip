// Wavefront path tracer: the bounce loop of PathIntegrator::li split into
// per-stage kernels over a pool of path slots whose state lives in HBM as
// structure-of-arrays, with compacted slot queues and path regeneration
// (Laine, Karras, Aila 2013), so that
//   * lanes only ever hold live paths (the shade stage compacts survivors
//     into the next queue),
//   * finished slots are refilled with new camera paths from any job of the
//     batch (scenes and cameras of very different sizes share one pool),
//   * a batched MFMA stage (the learned proposal's inference) runs between
//     the closest-hit query and shading.
//
// One iteration = regenerate -> extend -> [infer] -> shade -> bookkeeping.
// Every kernel reads its work counts from device memory and exits early, so
// the host launches fixed grids and can capture iterations in a hipGraph.
#include <hip/hip_runtime.h>
#include "guide_kernels.h"
#include "kernels.h"
#include "path.h"
#include "wavefront.h"

#pragma clang fp contract(off)

namespace illume {

namespace {

__device__ __forceinline__ void store_state(const WfState &w, uint32_t p, const PathState &ps) {
    w.o_tmax[p] = make_float4(ps.o.x, ps.o.y, ps.o.z, ps.t_max);
    w.d_eta[p] = make_float4(ps.d.x, ps.d.y, ps.d.z, ps.eta_scale);
    w.beta[p] = make_float4(ps.beta.x, ps.beta.y, ps.beta.z, 0.f);
    w.L[p] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.f);
    w.rng[p] = make_ulonglong2(ps.rng.state, ps.rng.inc);
    w.flags[p] = make_uint2(ps.bounces, (ps.active ? 1u : 0u) | (ps.specular ? 2u : 0u));
}
// table: the job's low-discrepancy table (null for the random streams); the
// sampler's index and next dimension are the stored state and increment.
__device__ __forceinline__ void load_state(const WfState &w, uint32_t p, const uint32_t *table, PathState &ps) {
    const float4 a = w.o_tmax[p], b = w.d_eta[p], c = w.beta[p], l = w.L[p];
    const ulonglong2 r = w.rng[p];
    const uint2 f = w.flags[p];
    ps.o = mk(a.x, a.y, a.z);
    ps.t_max = a.w;
    ps.d = mk(b.x, b.y, b.z);
    ps.eta_scale = b.w;
    ps.beta = mk(c.x, c.y, c.z);
    ps.L = mk(l.x, l.y, l.z);
    ps.rng.state = r.x;
    ps.rng.inc = r.y;
    ps.rng.lds = table;
    ps.bounces = f.x;
    ps.active = (f.y & 1u) != 0;
    ps.specular = (f.y & 2u) != 0;
}

// Slots to (re)start this iteration: all free slots, up to the paths left,
// once fewer than regen_min slots are live.
__device__ __forceinline__ uint32_t regen_take(const WfState &w, const WfCounters &c) {
    if (c.n_q[c.parity] >= w.regen_min) return 0;
    const unsigned long long rem = c.total_paths - c.next_path;
    return (uint32_t)min((unsigned long long)c.n_free, rem);
}

__device__ __forceinline__ uint32_t find_job(const WfState &w, unsigned long long g) {
    uint32_t lo = 0, hi = w.n_jobs;  // last job with path_begin <= g
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (w.jobs[mid].path_begin <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// A slot in a queue for every lane of the wave (among the active lanes) whose
// `pred` holds, with one atomicAdd per wave instead of one per lane (millions
// of lanes adding 1 to one address serialise at the L2): the first such lane
// reserves them all, each takes base + its rank. Every active lane of the wave
// must call it (divergent branches call it with their own active lanes).
__device__ __forceinline__ uint32_t wave_slot(uint32_t *ctr, bool pred) {
    const unsigned long long m = __ballot(pred);
    const uint32_t lane = __lane_id();
    const int leader = m ? __ffsll(m) - 1 : 0;
    uint32_t base = 0;
    if (m && (int)lane == leader) base = atomicAdd(ctr, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

__device__ __forceinline__ unsigned long long wave_slot64(unsigned long long *ctr, bool pred) {
    const unsigned long long m = __ballot(pred);
    const uint32_t lane = __lane_id();
    const int leader = m ? __ffsll(m) - 1 : 0;
    unsigned long long base = 0;
    if (m && (int)lane == leader) base = atomicAdd(ctr, (unsigned long long)__popcll(m));
    base = __shfl(base, leader);
    return base + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
}

__global__ void wf_reset_kernel(WfState w, unsigned long long total) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < w.capacity) w.free_q[i] = i;
    if (i == 0) {
        WfCounters &c = *w.ctr;
        c.n_q[0] = c.n_q[1] = 0;
        c.n_free = w.capacity;
        c.parity = 0;
        for (int k = 0; k < kMaxNets; ++k) c.guide_count[k] = 0;
        c.take = 0;
        c.next_path = 0;
        c.total_paths = total;
        c.done_paths = 0;
    }
}

__global__ void wf_regen_kernel(WfState w) {
    const WfCounters c = *w.ctr;
    const uint32_t take = regen_take(w, c);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= take) return;
    const uint32_t slot = w.free_q[c.n_free - take + i];
    const unsigned long long g = c.next_path + i;
    const uint32_t j = find_job(w, g);
    const WfJob &job = w.jobs[j];
    // filtered jobs sample the filter's sample bounds (stream = index in them),
    // the others one sample pixel per image pixel
    const bool filt = job.filtered != 0;
    const unsigned long long n_pix =
        filt ? (unsigned long long)job.fs.wext * job.fs.hext : (unsigned long long)job.cam.width * job.cam.height;
    const unsigned long long local = g - job.path_begin;
    const uint64_t pix = local % n_pix;
    const int x = filt ? job.fs.x0 + (int)(pix % job.fs.wext) : (int)(pix % job.cam.width);
    const int y = filt ? job.fs.y0 + (int)(pix / job.fs.wext) : (int)(pix / job.cam.width);
    PathState ps;
    float fx, fy;
    path_start_at(job.cam, job.seed, pix, x, y, job.sample_offset + (uint32_t)(local / n_pix), ps, fx, fy, job.sl);
    store_state(w, slot, ps);
    w.film_xy[slot] = make_float2(fx, fy);
    w.pix[slot] = (uint32_t)pix;
    w.job[slot] = j;
    w.n_recs[slot] = 0;
    w.q[c.parity][c.n_q[c.parity] + i] = slot;
}

__global__ void wf_book_regen_kernel(WfState w) {
    WfCounters &c = *w.ctr;
    const uint32_t take = regen_take(w, c);
    c.n_free -= take;
    c.next_path += take;
    c.n_q[c.parity] += take;
    c.take = take;
}

__global__ void wf_book_end_kernel(WfState w) {
    WfCounters &c = *w.ctr;
    c.n_q[c.parity] = 0;
    c.parity ^= 1u;
    for (int k = 0; k < kMaxNets; ++k) c.guide_count[k] = 0;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void wf_extend_kernel(WfState w, int guided,
                                                          unsigned long long *__restrict__ counters) {
    __shared__ int lds_stack[kStackLds * BLOCK];
    const uint32_t parity = w.ctr->parity;
    const uint32_t n_in = w.ctr->n_q[parity];
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    uint32_t n_closest = 0;
    if (i < n_in) {
        const uint32_t p = w.q[parity][i];
        const WfJob &job = w.jobs[w.job[p]];
        const DevScene &sc = *job.scene;
        const float4 a = w.o_tmax[p], b = w.d_eta[p];
        int spill[kSpill];
        TraceStack<BLOCK> st;
        st.reset(lds_stack + threadIdx.x, spill);
        TraceHit h;
        uint32_t nb = 0, np = 0;
        const f3 o = mk(a.x, a.y, a.z), d = mk(b.x, b.y, b.z);
        const bool found = trace_ray4<false, false, BLOCK>(sc.nodes, sc.prims, o, d, a.w, h, st, nb, np);
        ++n_closest;
        w.hit0[p] = make_float4(__uint_as_float(found ? h.prim : 0xFFFFFFFFu), __uint_as_float(h.slot), h.t, h.b0);
        w.hit1[p] = make_float4(h.b1, h.b2, 0.f, 0.f);
        int gslot = -1;
        int net = -1;  // the network whose queue segment takes this vertex
        if (guided && found && job.gc.out && job.gc.net >= 0) {
            const DevMaterial &m = sc.materials[sc.prim_material[h.prim]];
            if (guidable(m, job.gc.min_alpha) && w.flags[p].x < job.max_depth) net = job.gc.net;
        }
        // rows in each network's segment, one reservation per wave and network
        unsigned long long left = __ballot(net >= 0);
        while (left) {
            const int lnet = __shfl(net, __ffsll(left) - 1);
            const bool mine = net == lnet;
            const uint32_t r = wave_slot(&w.ctr->guide_count[lnet], mine);
            if (mine) gslot = lnet * (int)w.capacity + (int)r;
            left &= ~__ballot(mine);
        }
        if (gslot >= 0) {
            {
                Isect it;
                hit_isect(sc, h, o, d, it);
                float *in = w.guide_in + 6 * (size_t)gslot;
                in[0] = (it.p.x - job.gc.bounds_lo[0]) * job.gc.inv_ext[0];
                in[1] = (it.p.y - job.gc.bounds_lo[1]) * job.gc.inv_ext[1];
                in[2] = (it.p.z - job.gc.bounds_lo[2]) * job.gc.inv_ext[2];
                in[3] = it.n.x;
                in[4] = it.n.y;
                in[5] = it.n.z;
            }
        }
        w.guide_slot[p] = gslot;
    }
    if (counters) {
        for (int off = 32; off > 0; off >>= 1) n_closest += __shfl_xor(n_closest, off);
        if ((threadIdx.x & 63) == 0 && n_closest) atomicAdd(counters, (unsigned long long)n_closest);
    }
}

// Training samples of a finished path: for vertex k with a snapshot, the
// radiance that arrived along its sampled direction is
// (L_final - L_snap) / beta_after; the weight is its luminance over the pdf.
// They go to the region of the job's network `net` (-1: none), reserved per
// wave and network. Every active lane of the wave calls it.
__device__ void emit_records(const WfState &w, uint32_t p, uint32_t n, f3 Lf, int net) {
    const float lf[3] = {Lf.x, Lf.y, Lf.z};
    if (net < 0) n = 0;
    for (uint32_t k = 0; __ballot(k < n); ++k) {
        float li[3], wgt = 0.f;
        const VertexRec *r = k < n ? &w.recs[(size_t)p * w.max_depth + k] : nullptr;
        if (r && r->lsnap[0] >= 0.f) {
            for (int c = 0; c < 3; ++c) li[c] = r->beta[c] > 0.f ? fmaxf(0.f, lf[c] - r->lsnap[c]) / r->beta[c] : 0.f;
            wgt = (0.212671f * li[0] + 0.715160f * li[1] + 0.072169f * li[2]) / r->pdf;
        }
        const bool valid = r && wgt > 0.f && __builtin_isfinite(wgt);
        const int vnet = valid ? net : -1;
        unsigned long long left = __ballot(vnet >= 0);
        while (left) {
            const int lnet = __shfl(vnet, __ffsll(left) - 1);
            const bool mine = vnet == lnet;
            const unsigned long long s = wave_slot64(&w.ctr->rec_count[lnet], mine);
            if (mine && s < w.rec_cap_net) {
                float *o = w.train_rec + 10 * ((size_t)lnet * w.rec_cap_net + s);
                for (int c = 0; c < 6; ++c) o[c] = r->in[c];
                o[6] = r->wi[0];
                o[7] = r->wi[1];
                o[8] = r->wi[2];
                o[9] = wgt;
            }
            left &= ~__ballot(mine);
        }
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void wf_shade_kernel(WfState w, float *__restrict__ film,
                                                         unsigned long long *__restrict__ counters) {
    __shared__ int lds_stack[kStackLds * BLOCK];
    const uint32_t parity = w.ctr->parity;
    const uint32_t n_in = w.ctr->n_q[parity];
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    PathCounters pc{0, 0};
    if (i < n_in) {
        const uint32_t p = w.q[parity][i];
        const WfJob &job = w.jobs[w.job[p]];
        int spill[kSpill];
        TraceStack<BLOCK> st;
        st.reset(lds_stack + threadIdx.x, spill);
        PathState ps;
        load_state(w, p, job.sl.table, ps);
        const float4 h0 = w.hit0[p], h1 = w.hit1[p];
        TraceHit h;
        h.prim = __float_as_uint(h0.x);
        h.slot = __float_as_uint(h0.y);
        h.t = h0.z;
        h.b0 = h0.w;
        h.b1 = h1.x;
        h.b2 = h1.y;
        const bool found = h.prim != 0xFFFFFFFFu;
        const int gslot = w.guide_slot[p];
        const float *gout = (job.gc.out && gslot >= 0) ? job.gc.out + 40 * (size_t)gslot : nullptr;
        uint32_t n_recs = w.n_recs[p];
        VertexRec *recs = w.recs ? w.recs + (size_t)p * w.max_depth : nullptr;
        path_shade<BLOCK>(*job.scene, job.max_depth, ps, found, h, gout, job.gc, recs, n_recs, st, pc);
        w.n_recs[p] = n_recs;
        if (ps.active) {
            store_state(w, p, ps);
            w.q[parity ^ 1u][wave_slot(&w.ctr->n_q[parity ^ 1u], true)] = p;
        } else {
            const f3 L = guard_radiance(ps.L);
            if (job.filtered) {
                const float2 fxy = w.film_xy[p];
                film_splat(job.fs, film + job.film_offset, (int)job.cam.width, (int)job.cam.height, fxy.x, fxy.y, L.x,
                           L.y, L.z);
            } else {
                float *px = film + job.film_offset + 3 * (size_t)w.pix[p];
                atomicAdd(px, L.x);
                atomicAdd(px + 1, L.y);
                atomicAdd(px + 2, L.z);
            }
            if (recs) emit_records(w, p, n_recs, L, job.gc.net);
            w.flags[p] = make_uint2(ps.bounces, 0u);
            const uint32_t f = wave_slot(&w.ctr->n_free, true);
            w.free_q[f] = p;
            // the wave's finished paths counted once
            const unsigned long long done = __ballot(true);
            if ((int)__lane_id() == __ffsll(done) - 1) atomicAdd(&w.ctr->done_paths, (unsigned long long)__popcll(done));
        }
    }
    if (counters) {
        uint32_t c = pc.closest, a = pc.any;
        for (int off = 32; off > 0; off >>= 1) {
            c += __shfl_xor(c, off);
            a += __shfl_xor(a, off);
        }
        if ((threadIdx.x & 63) == 0) {
            if (c) atomicAdd(counters, (unsigned long long)c);
            if (a) atomicAdd(counters + 1, (unsigned long long)a);
        }
    }
}

template <int BLOCK>
hipError_t iteration(const WfState &w, bool guided, float *film, unsigned long long *cnt, hipStream_t st) {
    const uint32_t grid = (w.capacity + BLOCK - 1) / BLOCK;
    wf_regen_kernel<<<(w.capacity + 255) / 256, 256, 0, st>>>(w);
    wf_book_regen_kernel<<<1, 1, 0, st>>>(w);
    wf_extend_kernel<BLOCK><<<grid, BLOCK, 0, st>>>(w, guided ? 1 : 0, cnt);
    if (guided) {
        hipError_t e = guide_launch_infer_nets(w.nets, w.n_nets, w.guide_in, w.ctr->guide_count, w.capacity,
                                               w.guide_out, st);
        if (e != hipSuccess) return e;
    }
    wf_shade_kernel<BLOCK><<<grid, BLOCK, 0, st>>>(w, film, cnt);
    wf_book_end_kernel<<<1, 1, 0, st>>>(w);
    return hipGetLastError();
}

}  // namespace

hipError_t wf_launch_reset(const WfState &w, unsigned long long total_paths, hipStream_t st) {
    wf_reset_kernel<<<(w.capacity + 255) / 256, 256, 0, st>>>(w, total_paths);
    return hipGetLastError();
}

namespace {
__global__ void wf_clear_records_kernel(WfCounters *c, int net) {
    for (int k = 0; k < kMaxNets; ++k)
        if (net < 0 || k == net) c->rec_count[k] = 0ull;
}
}  // namespace

hipError_t wf_launch_clear_records(const WfState &w, int net, hipStream_t st) {
    wf_clear_records_kernel<<<1, 1, 0, st>>>(w.ctr, net);
    return hipGetLastError();
}

hipError_t wf_launch_iteration(const WfState &w, bool guided, float *film, unsigned long long *ray_counts, int block,
                               hipStream_t st) {
    switch (block) {
        case 64: return iteration<64>(w, guided, film, ray_counts, st);
        case 256: return iteration<256>(w, guided, film, ray_counts, st);
        default: return iteration<128>(w, guided, film, ray_counts, st);
    }
}

}  // namespace illume
