// Wavefront path tracer: the bounce loop of PathIntegrator::li split into
// per-stage kernels over a pool of path slots whose state lives in HBM as
// structure-of-arrays, with path regeneration (Laine, Karras, Aila 2013), so
// that
//   * lanes only ever hold live paths (the shade stage compacts survivors),
//   * finished slots are refilled with new camera paths from any job of the
//     batch (scenes and cameras of very different sizes share one pool),
//   * a batched MFMA stage (the learned proposal's inference) runs between
//     the closest-hit query and shading.
//
// One iteration = regenerate -> extend -> [infer] -> shade -> direct
// lighting -> finish. Shading traces nothing: the continuation ray is the next
// iteration's extend stage, and the two rays of the direct-lighting estimate
// (the light sample's shadow ray, the BSDF sample's ray toward the light) are
// traced by a kernel that holds only traversal state; paths that ended are
// filmed after it, once their last vertex's direct light is in.
//
// The pool is cut into tiles of 4 x block slots, and block t of every stage
// kernel owns tile t for the whole batch: it reads its tile's counts, keeps
// its live paths at the tile's first slots (survivors compacted with a
// block-wide scan into the other state buffer), and writes its requests and
// finished paths from the tile's first slot. Only starting new paths, counting
// finished ones and reserving proposal rows and training samples touch
// batch-wide counters, once per block (and network). Every kernel launches one block per tile and reads its work
// counts from device memory, so the host can capture iterations in a
// hipGraph (parity, which buffer holds the live paths, alternates with the
// iteration and is a launch argument).
#include <hip/hip_runtime.h>
#include "guide_kernels.h"
#include "kernels.h"
#include "path.h"
#include "wavefront.h"

#pragma clang fp contract(off)

namespace illume {

namespace {

constexpr uint32_t kSpecularBit = 1u << 8;
constexpr uint32_t kFinishDest = 1u << 31;  // direct-lighting destination in the finish buffer
constexpr int kTileRounds = 4;             // slots per tile = kTileRounds x block
constexpr uint32_t kNoRecords = 1u << 31;  // record-slot bit: this path keeps no training vertices

// Integer hash of a path index (the 64-bit finaliser of MurmurHash3, folded).
__device__ __forceinline__ uint32_t path_hash(unsigned long long g) {
    g ^= g >> 33;
    g *= 0xff51afd7ed558ccdull;
    g ^= g >> 33;
    g *= 0xc4ceb9fe1a85ec53ull;
    g ^= g >> 33;
    return (uint32_t)g;
}

__device__ __forceinline__ uint32_t path_bits(const PathState &ps, uint32_t n_recs) {
    return ps.bounces | (ps.specular ? kSpecularBit : 0u) | (n_recs << 16);
}

// Path at slot p of buffer s (the sampler's low-discrepancy table comes from
// the path's job; null for the random streams, whose index and next dimension
// are the stored state and increment).
__device__ __forceinline__ void load_path(const WfPaths &s, uint32_t p, const WfJob *jobs, PathState &ps,
                                          uint32_t &job, uint32_t &n_recs, uint4 &meta) {
    const float4 a = s.o_tmax[p], b = s.d_eta[p], c = s.beta[p], l = s.L[p];
    const ulonglong2 r = s.rng[p];
    meta = s.meta[p];
    const uint32_t bits = __float_as_uint(c.w);
    job = __float_as_uint(l.w);
    ps.o = mk(a.x, a.y, a.z);
    ps.t_max = a.w;
    ps.d = mk(b.x, b.y, b.z);
    ps.eta_scale = b.w;
    ps.beta = mk(c.x, c.y, c.z);
    ps.L = mk(l.x, l.y, l.z);
    ps.rng.state = r.x;
    ps.rng.inc = r.y;
    ps.rng.lds = jobs[job].sl.table;
    ps.bounces = bits & 0xFFu;
    ps.specular = (bits & kSpecularBit) != 0;
    ps.active = true;
    n_recs = bits >> 16;
}

__device__ __forceinline__ void store_path(const WfPaths &s, uint32_t p, const PathState &ps, uint32_t job,
                                           uint32_t n_recs, uint4 meta) {
    s.o_tmax[p] = make_float4(ps.o.x, ps.o.y, ps.o.z, ps.t_max);
    s.d_eta[p] = make_float4(ps.d.x, ps.d.y, ps.d.z, ps.eta_scale);
    s.beta[p] = make_float4(ps.beta.x, ps.beta.y, ps.beta.z, __uint_as_float(path_bits(ps, n_recs)));
    s.L[p] = make_float4(ps.L.x, ps.L.y, ps.L.z, __uint_as_float(job));
    s.rng[p] = make_ulonglong2(ps.rng.state, ps.rng.inc);
    s.meta[p] = meta;
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

// Block-wide append to N lists: for each list q, the threads whose f[q]
// holds get consecutive positions run[q] + (their rank in the block), and
// run[q] advances by the block's count. Every thread of the block calls it
// (two barriers); run is the same in every thread.
template <int BLOCK, int N>
__device__ __forceinline__ void block_append(const bool (&f)[N], uint32_t (&pos)[N], uint32_t (&run)[N],
                                             uint32_t *lds) {
    constexpr int W = BLOCK / 64;
    const int wave = threadIdx.x / 64;
    const uint32_t lane = __lane_id();
    uint32_t rank[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const unsigned long long m = __ballot(f[q]);
        rank[q] = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) lds[q * W + wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N; ++q) {
        uint32_t off = run[q], tot = 0;
#pragma unroll
        for (int v = 0; v < W; ++v) {
            const uint32_t c = lds[q * W + v];
            if (v < wave) off += c;
            tot += c;
        }
        pos[q] = off + rank[q];
        run[q] += tot;
    }
    __syncthreads();
}

// Block-wide reservation of slots from per-network counters, one atomic per
// (block, network): a thread's round k (of kTileRounds) names network
// key(k) - 1 (key 0: none) and needs count(k) consecutive slots. Networks are
// served in increasing order; for each, emit(k, first) is called for every
// round of that network with the first of its slots, counter(key, total)
// returning the first of the block's total slots. Every thread of the block
// calls it (four barriers per network, two more at the end).
template <int BLOCK, typename KeyF, typename CountF, typename CtrF, typename EmitF>
__device__ __forceinline__ void block_reserve(uint32_t *lds, KeyF key, CountF count, CtrF counter, EmitF emit) {
    constexpr int W = BLOCK / 64;
    const int wave = threadIdx.x / 64;
    const int lane = (int)__lane_id();
    uint32_t done = 0;  // rounds already served (bit k)
    for (int guard = 0; guard <= kMaxNets; ++guard) {
        uint32_t mine = 0xffffffffu;
#pragma unroll
        for (int k = 0; k < kTileRounds; ++k) {
            const uint32_t v = key(k);
            if (v && !((done >> k) & 1u)) mine = min(mine, v);
        }
        for (int o = 32; o > 0; o >>= 1) mine = min(mine, (uint32_t)__shfl_xor((int)mine, o));
        if (lane == 0) lds[wave] = mine;
        __syncthreads();
        uint32_t cur = 0xffffffffu;
#pragma unroll
        for (int v = 0; v < W; ++v) cur = min(cur, lds[v]);
        __syncthreads();
        if (cur == 0xffffffffu) return;  // the same in every thread
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < kTileRounds; ++k)
            if (key(k) == cur && !((done >> k) & 1u)) cnt += count(k);
        uint32_t x = cnt;  // inclusive prefix sum over the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) lds[wave] = x;
        __syncthreads();
        uint32_t before = 0, total = 0;
#pragma unroll
        for (int v = 0; v < W; ++v) {
            const uint32_t c = lds[v];
            if (v < wave) before += c;
            total += c;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned long long b = total ? counter(cur, total) : 0ull;
            lds[W] = (uint32_t)b;
            lds[W + 1] = (uint32_t)(b >> 32);
        }
        __syncthreads();
        unsigned long long first = ((unsigned long long)lds[W + 1] << 32 | lds[W]) + before + x - cnt;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kTileRounds; ++k)
            if (key(k) == cur && !((done >> k) & 1u)) {
                emit(k, first);
                first += count(k);
                done |= 1u << k;
            }
    }
}

// Adds the block's sum of v (one value per thread) to *dst with one atomic.
// Every thread of the block calls it.
template <int BLOCK>
__device__ __forceinline__ void block_add(unsigned long long *dst, uint32_t v, uint32_t *lds) {
    constexpr int W = BLOCK / 64;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (__lane_id() == 0) lds[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < W; ++k) t += lds[k];
        if (t) atomicAdd(dst, t);
    }
    __syncthreads();
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void wf_reset_kernel(WfState w, unsigned long long total) {
    const uint32_t t = blockIdx.x, S = kTileRounds * BLOCK;
    if (w.recs)
        for (uint32_t j = threadIdx.x; j < S; j += BLOCK) w.rec_free[t * S + j] = t * S + j;
    if (threadIdx.x == 0) {
        w.tile_live[0][t] = w.tile_live[1][t] = 0;
        w.tile_free[t] = S;
    }
    if (t == 0 && threadIdx.x == 0) {
        WfCounters &c = *w.ctr;
        c.next_path = 0;
        c.total_paths = total;
        c.done_paths = 0;
    }
}

// New camera paths after tile t's live ones in buffer `parity`, once fewer
// than regen_thr are live: a run of path indices from the batch-wide counter
// (one atomic per block). Block 0 also clears the proposal-row counts.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void wf_regen_kernel(WfState w, uint32_t parity) {
    __shared__ unsigned long long s_base;
    __shared__ uint32_t s_take;
    const uint32_t t = blockIdx.x, S = kTileRounds * BLOCK;
    const uint32_t live = w.tile_live[parity][t];
    if (threadIdx.x == 0) {
        uint32_t take = 0;
        unsigned long long base = 0;
        WfCounters &c = *w.ctr;
        const unsigned long long total = c.total_paths;
        if (live < w.regen_thr && c.next_path < total) {
            const uint32_t want = S - live;
            base = atomicAdd(&c.next_path, (unsigned long long)want);
            if (base < total) take = (uint32_t)min((unsigned long long)want, total - base);
        }
        s_base = base;
        s_take = take;
        if (t == 0)
            for (int k = 0; k < kMaxNets; ++k) c.guide_count[k][0] = 0;
    }
    __syncthreads();
    const uint32_t take = s_take;
    if (take == 0) return;
    const unsigned long long base = s_base;
    const uint32_t free0 = w.recs ? w.tile_free[t] : 0u;
    for (uint32_t i = threadIdx.x; i < take; i += BLOCK) {
        const unsigned long long g = base + i;
        const uint32_t j = find_job(w, g);
        const WfJob &job = w.jobs[j];
        // filtered jobs sample the filter's sample bounds (stream = index in
        // them), the others one sample pixel per image pixel
        const bool filt = job.filtered != 0;
        const unsigned long long n_pix =
            filt ? (unsigned long long)job.fs.wext * job.fs.hext : (unsigned long long)job.cam.width * job.cam.height;
        const unsigned long long local = g - job.path_begin;
        const uint64_t pix = local % n_pix;
        const int x = filt ? job.fs.x0 + (int)(pix % job.fs.wext) : (int)(pix % job.cam.width);
        const int y = filt ? job.fs.y0 + (int)(pix / job.fs.wext) : (int)(pix / job.cam.width);
        PathState ps;
        float fx, fy;
        path_start_at(job.cam, job.seed, pix, x, y, job.sample_offset + (uint32_t)(local / n_pix), ps, fx, fy,
                      job.sl);
        uint32_t rec = 0u;
        if (w.recs) {
            rec = w.rec_free[t * S + free0 - take + i];
            if (path_hash(g) > w.rec_thresh) rec |= kNoRecords;
        }
        // unfiltered paths with a sample buffer keep their path index in place
        // of the film position
        const uint32_t my = (!filt && w.smp) ? (uint32_t)g : __float_as_uint(fy);
        store_path(w.s[parity], t * S + live + i, ps, j, 0u,
                   make_uint4((uint32_t)pix, __float_as_uint(fx), my, rec));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        w.tile_live[parity][t] = live + take;
        if (w.recs) w.tile_free[t] = free0 - take;
    }
}

// The tail of both extend kernels: proposal rows for the guided vertices of
// tile t (pend: round k's 5 bits hold 1 + the network, 0 none), one
// reservation per network for the whole block (per-wave atomics on the
// batch-wide row counter serialised), and the block's closest-hit count.
template <int BLOCK>
__device__ __forceinline__ void extend_tail(const WfState &w, uint32_t t, bool guided, uint32_t pend,
                                            uint32_t n_closest, uint32_t *lds_red,
                                            unsigned long long *__restrict__ counters) {
    constexpr uint32_t S = kTileRounds * BLOCK;
    if (guided) {
        block_reserve<BLOCK>(
            lds_red, [&](int k) { return (pend >> (5 * k)) & 31u; }, [&](int) { return 1u; },
            [&](uint32_t key, uint32_t total) {
                return (unsigned long long)atomicAdd(&w.ctr->guide_count[key - 1][0], total);
            },
            [&](int k, unsigned long long first) {
                const uint32_t p = t * S + (uint32_t)k * BLOCK + threadIdx.x;
                const int gslot = (int)(((pend >> (5 * k)) & 31u) - 1u) * (int)w.capacity + (int)first;
                w.guide_slot[gslot] = p;
                reinterpret_cast<float *>(&w.hit1[p])[2] = __int_as_float(gslot);
            });
    }
    if (counters) block_add<BLOCK>(counters, n_closest, lds_red);
}

// Closest hit of every live path's ray, one path per thread and round;
// guided vertices get a row in their network's segment of the proposal queue.
// Small scenes (rays end after a few steps) take this kernel.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void wf_extend_kernel(WfState w, uint32_t parity, int guided,
                                                          unsigned long long *__restrict__ counters) {
    __shared__ int lds_stack[kStackLds * BLOCK];
    __shared__ uint32_t lds_red[BLOCK / 64 + 2];
    const uint32_t t = blockIdx.x, S = kTileRounds * BLOCK;
    const uint32_t n_in = w.tile_live[parity][t];
    const WfPaths &s = w.s[parity];
    uint32_t n_closest = 0;
    uint32_t pend = 0;  // round k: 1 + the network that takes the vertex's proposal row (5 bits), 0: none
    for (uint32_t k = 0, i = threadIdx.x; i < n_in; ++k, i += BLOCK) {
        const uint32_t p = t * S + i;
        const float4 a = s.o_tmax[p], b = s.d_eta[p];
        const WfJob &job = w.jobs[__float_as_uint(s.L[p].w)];
        const DevScene &sc = *job.scene;
        int spill[kSpill];
        TraceStack<BLOCK> st;
        st.reset(lds_stack + threadIdx.x, spill);
        TraceHit h;
        uint32_t nb = 0, np = 0;
        const f3 o = mk(a.x, a.y, a.z), d = mk(b.x, b.y, b.z);
        const bool found = trace_ray4<false, false, BLOCK>(sc.nodes, sc.prims, o, d, a.w, h, st, nb, np);
        ++n_closest;
        int gslot = -1;
        int net = -1;  // the network whose queue segment takes this vertex
        if (guided && found && job.gc.out && job.gc.net >= 0) {
            const DevMaterial &m = sc.materials[sc.prim_material[h.prim]];
            const uint32_t bounces = __float_as_uint(s.beta[p].w) & 0xFFu;
            if (guidable(m, job.gc.min_alpha) && bounces < job.max_depth) net = job.gc.net;
        }
        if (net >= 0) {
            // the proposal's inputs stay at the path's slot: its row, once
            // the block has reserved the rows, records the slot
            Isect it;
            hit_isect(sc, h, o, d, it);
            w.gstage[2 * (size_t)p] = make_float4((it.p.x - job.gc.bounds_lo[0]) * job.gc.inv_ext[0],
                                                  (it.p.y - job.gc.bounds_lo[1]) * job.gc.inv_ext[1],
                                                  (it.p.z - job.gc.bounds_lo[2]) * job.gc.inv_ext[2], 0.f);
            w.gstage[2 * (size_t)p + 1] = make_float4(it.n.x, it.n.y, it.n.z, 0.f);
            pend |= (uint32_t)(net + 1) << (5 * k);
        }
        w.hit0[p] = make_float4(__uint_as_float(found ? h.prim : 0xFFFFFFFFu), __uint_as_float(h.slot), h.t, h.b0);
        w.hit1[p] = make_float4(h.b1, h.b2, __int_as_float(gslot), 0.f);
    }
    extend_tail<BLOCK>(w, t, guided != 0, pend, n_closest, lds_red, counters);
}

// The same for scenes large enough that rays take many traversal steps: the
// traversal refills its lanes: a wave takes the tile's next paths (a
// counter in LDS) into its idle lanes once at least kExtendRefill of them are
// idle, so a lane whose ray ended early (a miss, a short path through the
// tree) takes another instead of waiting for the wave's slowest ray. An ended
// ray leaves only its primitive record and distance in LDS; a second pass
// over the tile, one path per thread, forms the hit records and the guided
// vertices' proposal inputs with every lane busy. On the 64-scene batch (2^23
// slots, 256-lane tiles) 20 idle lanes gave 0.2210-0.2214 s per render, 24
// 0.2213-0.2222 and 28 0.2220-0.2229 (the extend rays are shorter than the
// 2^24-ray closest-hit workload's, which is fastest at 24).
constexpr int kExtendRefill = 20;
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void wf_extend_refill_kernel(WfState w, uint32_t parity, int guided,
                                                                 unsigned long long *__restrict__ counters) {
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    constexpr uint32_t S = kTileRounds * BLOCK;
    __shared__ int lds_stack[kStackLds * BLOCK];
    __shared__ uint32_t lds_red[BLOCK / 64 + 2];
    __shared__ uint32_t lds_slot[S];  // path i: primitive record of its closest hit, kNoSlot: none
    __shared__ float lds_t[S];        // path i: its distance
    __shared__ uint32_t lds_next;     // the tile's next path not yet taken
    const uint32_t t = blockIdx.x;
    const uint32_t n_in = w.tile_live[parity][t];
    const WfPaths &s = w.s[parity];
    if (threadIdx.x == 0) lds_next = 0;
    __syncthreads();
    const uint32_t lane = __lane_id();
    uint32_t n_closest = 0;
    {
        int spill[kSpill];
        TraceStack<BLOCK> st;
        st.reset(lds_stack + threadIdx.x, spill);
        uint32_t i = kNone;  // the lane's path (index in the tile), kNone: none
        bool busy = false, more = n_in > 0;
        const GpuNode4Q *nodes = nullptr;
        const PrimRecord *prims = nullptr;
        f3 o = mk(0.f, 0.f, 0.f), d = o, inv = o;
        Shear sh{0, 0.f, 0.f, 0.f};
        float t_max = 0.f;
        int cur = kTraceDone;
        uint32_t pf = 0, pn = 0, slot = kNoSlot, n_box = 0;
        while (true) {
            const uint64_t idle = __ballot(!busy);
            const int n_idle = __popcll(idle);
            if (n_idle >= kExtendRefill || n_idle == 64) {
                if (!busy && i != kNone) {
                    lds_slot[i] = slot;
                    lds_t[i] = t_max;
                    i = kNone;
                }
                if (more) {
                    uint32_t base = 0;
                    if (lane == 0) base = atomicAdd(&lds_next, (uint32_t)n_idle);
                    base = (uint32_t)__shfl((int)base, 0);
                    if (base + (uint32_t)n_idle >= n_in) more = false;
                    const uint32_t r = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                    if (!busy && r < n_in) {
                        const uint32_t p = t * S + r;
                        const float4 a = s.o_tmax[p], b = s.d_eta[p];
                        const DevScene &sc = *w.jobs[__float_as_uint(s.L[p].w)].scene;
                        nodes = sc.nodes;
                        prims = sc.prims;
                        o = mk(a.x, a.y, a.z);
                        d = mk(b.x, b.y, b.z);
                        t_max = a.w;
                        inv = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
                        sh = make_shear(d);
                        slot = kNoSlot;
                        cur = 0;
                        pf = pn = 0;
                        st.sp = 0;
                        i = r;
                        busy = true;
                        ++n_closest;
                    }
                }
                if (__ballot(busy) == 0) break;  // no path left, every result in LDS
            }
            if (busy) {
                if (pn != 0) {
                    float th;
                    if (prim_step<false>(prims, pf, o, d, sh, t_max, th)) {
                        slot = pf;
                        t_max = th;
                    }
                    ++pf;
                    --pn;
                } else {
                    cur = node_step4<false, BLOCK>(nodes, cur, o, inv, t_max, st, n_box);
                }
                take_leaf(cur, pf, pn, st);
                if (pn == 0 && cur == kTraceDone) busy = false;
            }
        }
    }
    __syncthreads();  // every path's result in LDS
    uint32_t pend = 0;  // round k: 1 + the network that takes the vertex's proposal row (5 bits), 0: none
    for (uint32_t k = 0, i = threadIdx.x; i < n_in; ++k, i += BLOCK) {
        const uint32_t p = t * S + i;
        const uint32_t slot = lds_slot[i];
        TraceHit h;
        h.prim = 0xFFFFFFFFu;
        h.slot = 0;
        h.t = h.b0 = h.b1 = h.b2 = 0.f;
        const bool found = slot != kNoSlot;
        if (found) {
            const float4 a = s.o_tmax[p], b = s.d_eta[p];
            const WfJob &job = w.jobs[__float_as_uint(s.L[p].w)];
            const DevScene &sc = *job.scene;
            const f3 o = mk(a.x, a.y, a.z), d = mk(b.x, b.y, b.z);
            finish_hit(sc.prims, slot, o, make_shear(d), lds_t[i], h);
            if (guided && job.gc.out && job.gc.net >= 0) {
                const DevMaterial &m = sc.materials[sc.prim_material[h.prim]];
                const uint32_t bounces = __float_as_uint(s.beta[p].w) & 0xFFu;
                if (guidable(m, job.gc.min_alpha) && bounces < job.max_depth) {
                    // the proposal's inputs stay at the path's slot: its row,
                    // once the block has reserved the rows, records the slot
                    Isect it;
                    hit_isect(sc, h, o, d, it);
                    w.gstage[2 * (size_t)p] = make_float4((it.p.x - job.gc.bounds_lo[0]) * job.gc.inv_ext[0],
                                                          (it.p.y - job.gc.bounds_lo[1]) * job.gc.inv_ext[1],
                                                          (it.p.z - job.gc.bounds_lo[2]) * job.gc.inv_ext[2], 0.f);
                    w.gstage[2 * (size_t)p + 1] = make_float4(it.n.x, it.n.y, it.n.z, 0.f);
                    pend |= (uint32_t)(job.gc.net + 1) << (5 * k);
                }
            }
        }
        w.hit0[p] = make_float4(__uint_as_float(h.prim), __uint_as_float(h.slot), h.t, h.b0);
        w.hit1[p] = make_float4(h.b1, h.b2, __int_as_float(-1), 0.f);
    }
    extend_tail<BLOCK>(w, t, guided != 0, pend, n_closest, lds_red, counters);
}

// One bounce of tile t's live paths, tracing nothing. GUIDED (a proposal or
// training records in use) reads the proposal rows and keeps training
// vertices; the plain kernel compiles without them. Survivors go compacted to
// tile t of the other state buffer, ended paths to the finish buffer, and a
// vertex's direct-lighting rays to the request buffer, each from the tile's
// first slot.
// At least 3 waves per SIMD (168 VGPRs; the natural 182 / 203 give 2): the
// Cornell box's 4-spp pass 2.73 -> 2.55 ms plain, 4.93 -> 4.80 ms guided;
// 4 waves spill too much in the guided kernel (5.43 ms).
template <int BLOCK, bool GUIDED>
__global__ __launch_bounds__(BLOCK, 3) void wf_shade_kernel(WfState w, uint32_t parity) {
    __shared__ uint32_t lds[3 * (BLOCK / 64)];
    // the decoded proposal of each thread's vertex (48 KB at 256 threads:
    // three blocks per CU still fit)
    __shared__ float4 mix_mk[GUIDED ? 8 * BLOCK : 1];
    __shared__ float2 mix_pc[GUIDED ? 8 * BLOCK : 1];
    static_assert(!GUIDED || sizeof(mix_mk) + sizeof(mix_pc) == kMixtureLds<BLOCK>, "mixture LDS layout");
    const uint32_t t = blockIdx.x, S = kTileRounds * BLOCK;
    const uint32_t n_in = w.tile_live[parity][t];
    const uint32_t nxt = parity ^ 1u;
    uint32_t run[3] = {0u, 0u, 0u};  // survivors, ended, requests
    for (uint32_t r0 = 0; r0 < n_in; r0 += BLOCK) {
        const uint32_t i = r0 + threadIdx.x;
        bool f[3] = {false, false, false};
        PathState ps;
        uint32_t j = 0, n_recs = 0;
        uint4 meta;
        DeferredDirect dd;
        if (i < n_in) {
            const uint32_t p = t * S + i;
            load_path(w.s[parity], p, w.jobs, ps, j, n_recs, meta);
            const WfJob &job = w.jobs[j];
            const float4 h0 = w.hit0[p], h1 = w.hit1[p];
            TraceHit h;
            h.prim = __float_as_uint(h0.x);
            h.slot = __float_as_uint(h0.y);
            h.t = h0.z;
            h.b0 = h0.w;
            h.b1 = h1.x;
            h.b2 = h1.y;
            const bool found = h.prim != 0xFFFFFFFFu;
            const __half *gout = nullptr;
            VertexRec *recs = nullptr;
            if (GUIDED) {
                const int gslot = __float_as_int(h1.z);
                gout = (job.gc.out && gslot >= 0) ? job.gc.out + 40 * (size_t)gslot : nullptr;
                if (w.recs && !(meta.w & kNoRecords)) recs = w.recs + (size_t)meta.w * w.max_depth;
            }
            TraceStack<BLOCK> st;  // unused: shading traces nothing
            PathCounters pc{0, 0};
            Mixture<BLOCK> mix{nullptr, nullptr};
            if (GUIDED) mix = Mixture<BLOCK>{mix_mk + threadIdx.x, mix_pc + threadIdx.x};
            path_shade<BLOCK, true>(*job.scene, job.max_depth, ps, found, h, gout, job.gc, recs, n_recs, st, pc,
                                    &dd, mix);
            f[0] = ps.active;
            f[1] = !ps.active;
            // a non-finite throughput turns even a zero direct term into NaN,
            // as in the megakernel, so such a vertex is queued as well
            f[2] = dd.pending &&
                   (dd.r.shadow || dd.r.mis || !__builtin_isfinite(dd.beta.x + dd.beta.y + dd.beta.z));
        }
        uint32_t pos[3];
        block_append<BLOCK, 3>(f, pos, run, lds);
        if (f[0]) store_path(w.s[nxt], t * S + pos[0], ps, j, n_recs, meta);
        if (f[1]) {
            w.fin_L[t * S + pos[1]] = make_float4(ps.L.x, ps.L.y, ps.L.z, __uint_as_float(j));
            w.fin_meta[t * S + pos[1]] = meta;
            w.fin_nrec[t * S + pos[1]] = n_recs;
        }
        if (f[2]) {
            const DirectRays &r = dd.r;
            const uint32_t q = t * S + pos[2];
            const uint32_t bits = (r.shadow ? 1u : 0u) | (r.mis ? 2u : 0u) | ((uint32_t)r.li << 2);
            const uint32_t dest = f[0] ? t * S + pos[0] : (kFinishDest | (t * S + pos[1]));
            if (r.shadow) w.nee[0][q] = make_float4(r.so.x, r.so.y, r.so.z, r.s_tmax);
            w.nee[1][q] = make_float4(r.sd.x, r.sd.y, r.sd.z, dd.lpdf);
            w.nee[2][q] = make_float4(r.cs.x, r.cs.y, r.cs.z, __uint_as_float(bits));
            w.nee[3][q] = make_float4(dd.beta.x, dd.beta.y, dd.beta.z, __uint_as_float(dest));
            if (r.mis) {
                w.nee[4][q] = make_float4(r.mo.x, r.mo.y, r.mo.z, r.mw);
                w.nee[5][q] = make_float4(r.md.x, r.md.y, r.md.z, r.mpdf);
                w.nee[6][q] = make_float4(r.mf.x, r.mf.y, r.mf.z, 0.f);
            }
        }
    }
    if (threadIdx.x == 0) {
        w.tile_live[nxt][t] = run[0];
        w.tile_fin[t] = run[1];
        w.tile_nee[t] = run[2];
    }
}

// Tile t's direct-lighting requests: the shadow ray (any hit) and the BSDF
// sample's ray (closest hit), then L += beta (ld / lpdf) into the radiance of
// the request's path (in the next state buffer or the finish buffer).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void wf_nee_kernel(WfState w, uint32_t parity,
                                                       unsigned long long *__restrict__ counters) {
    __shared__ int lds_stack[kStackLds * BLOCK];
    __shared__ uint32_t lds_red[BLOCK / 64];
    const uint32_t t = blockIdx.x, S = kTileRounds * BLOCK;
    const uint32_t n_in = w.tile_nee[t];
    float4 *next_L = w.s[parity ^ 1u].L;
    PathCounters pc{0, 0};
    for (uint32_t i = threadIdx.x; i < n_in; i += BLOCK) {
        const uint32_t q = t * S + i;
        const float4 n1 = w.nee[1][q], n2 = w.nee[2][q], n3 = w.nee[3][q];
        const uint32_t dest = __float_as_uint(n3.w);
        float4 *lp = (dest & kFinishDest) ? w.fin_L + (dest & ~kFinishDest) : next_L + dest;
        const float4 l = *lp;
        const DevScene &sc = *w.jobs[__float_as_uint(l.w)].scene;
        int spill[kSpill];
        TraceStack<BLOCK> st;
        st.reset(lds_stack + threadIdx.x, spill);
        DirectRays r;
        const uint32_t bits = __float_as_uint(n2.w);
        r.shadow = (bits & 1u) != 0;
        r.mis = (bits & 2u) != 0;
        r.li = (int)(bits >> 2);
        if (r.shadow) {
            const float4 n0 = w.nee[0][q];
            r.so = mk(n0.x, n0.y, n0.z);
            r.s_tmax = n0.w;
            r.sd = mk(n1.x, n1.y, n1.z);
            r.cs = mk(n2.x, n2.y, n2.z);
        }
        if (r.mis) {
            const float4 n4 = w.nee[4][q], n5 = w.nee[5][q], n6 = w.nee[6][q];
            r.mo = mk(n4.x, n4.y, n4.z);
            r.mw = n4.w;
            r.md = mk(n5.x, n5.y, n5.z);
            r.mpdf = n5.w;
            r.mf = mk(n6.x, n6.y, n6.z);
        }
        const f3 ld = direct_resolve<BLOCK>(sc, r, st, pc);
        const f3 L = add_direct(mk(l.x, l.y, l.z), mk(n3.x, n3.y, n3.z), ld, n1.w);
        *lp = make_float4(L.x, L.y, L.z, l.w);
    }
    if (counters) {
        block_add<BLOCK>(counters, pc.closest, lds_red);
        block_add<BLOCK>(counters + 1, pc.any, lds_red);
    }
}

// Tile t's paths that ended this iteration: the guarded radiance into the
// film (or its splat over the filter footprint), training records, record
// slots back to the tile's free stack, the block's count added to done_paths.
// Training sample of vertex k of record slot `rec` for a path that ended with
// radiance lf: the radiance that arrived along its sampled direction is
// (L_final - L_snap) / beta_after, its weight the luminance of that over the
// pdf. False if the vertex has no snapshot or the weight is not positive and
// finite; else the 10 floats (inputs, direction, weight) go to o (if not null).
__device__ __forceinline__ bool vertex_record(const WfState &w, uint32_t rec, uint32_t k, const float *lf, float *o) {
    const VertexRec &r = w.recs[(size_t)rec * w.max_depth + k];
    if (!(r.lsnap[0] >= 0.f)) return false;
    float li[3];
    for (int c = 0; c < 3; ++c) li[c] = r.beta[c] > 0.f ? fmaxf(0.f, lf[c] - r.lsnap[c]) / r.beta[c] : 0.f;
    const float wgt = (0.212671f * li[0] + 0.715160f * li[1] + 0.072169f * li[2]) / r.pdf;
    if (!(wgt > 0.f && __builtin_isfinite(wgt))) return false;
    if (o) {
        for (int c = 0; c < 6; ++c) o[c] = r.in[c];
        o[6] = r.wi[0];
        o[7] = r.wi[1];
        o[8] = r.wi[2];
        o[9] = wgt;
    }
    return true;
}

// Ended paths of tile t: film (or sample buffer), then their training
// samples, reserved once per block and network (per-wave atomics on the
// batch-wide record counter serialised).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void wf_finish_kernel(WfState w, float *__restrict__ film) {
    __shared__ uint32_t lds[BLOCK / 64 + 2];
    const uint32_t t = blockIdx.x, S = kTileRounds * BLOCK;
    const uint32_t n_in = w.tile_fin[t];
    const uint32_t free0 = w.recs ? w.tile_free[t] : 0u;
    uint32_t pend = 0;  // round k: 1 + the network of its training samples (5 bits), 0: none
    unsigned long long vmask[kTileRounds];  // round k: its vertices with a training sample
#pragma unroll
    for (int k = 0; k < kTileRounds; ++k) {
        vmask[k] = 0ull;
        const uint32_t i = threadIdx.x + (uint32_t)k * BLOCK;
        if (i >= n_in) continue;
        const uint32_t q = t * S + i;
        const float4 l = w.fin_L[q];
        const uint4 meta = w.fin_meta[q];
        const WfJob &job = w.jobs[__float_as_uint(l.w)];
        const f3 L = guard_radiance(mk(l.x, l.y, l.z));
        if (job.filtered) {
            film_splat(job.fs, film + job.film_offset, (int)job.cam.width, (int)job.cam.height,
                       __uint_as_float(meta.y), __uint_as_float(meta.z), L.x, L.y, L.z);
        } else if (w.smp) {
            w.smp[meta.z] = make_float4(L.x, L.y, L.z, 0.f);
        } else {
            float *px = film + job.film_offset + 3 * (size_t)meta.x;
            atomicAdd(px, L.x);
            atomicAdd(px + 1, L.y);
            atomicAdd(px + 2, L.z);
        }
        if (w.recs) {
            const uint32_t rec = meta.w & ~kNoRecords;
            const uint32_t n = (meta.w & kNoRecords) ? 0u : w.fin_nrec[q];
            const int net = job.gc.net;
            if (net >= 0) {
                const float lf[3] = {L.x, L.y, L.z};
                unsigned long long m = 0ull;
                for (uint32_t v = 0; v < n; ++v)
                    if (vertex_record(w, rec, v, lf, nullptr)) m |= 1ull << v;
                vmask[k] = m;
                if (m) pend |= (uint32_t)(net + 1) << (5 * k);
            }
            w.rec_free[t * S + free0 + i] = rec;
        }
    }
    if (w.recs)
        block_reserve<BLOCK>(
            lds, [&](int k) { return (pend >> (5 * k)) & 31u; },
            [&](int k) { return (uint32_t)__popcll(vmask[k]); },
            [&](uint32_t key, uint32_t total) {
                return atomicAdd(&w.ctr->rec_count[key - 1][0], (unsigned long long)total);
            },
            [&](int k, unsigned long long first) {
                const uint32_t q = t * S + threadIdx.x + (uint32_t)k * BLOCK;
                const float4 l = w.fin_L[q];
                const f3 L = guard_radiance(mk(l.x, l.y, l.z));
                const float lf[3] = {L.x, L.y, L.z};
                const uint32_t rec = w.fin_meta[q].w & ~kNoRecords;
                const size_t region = (size_t)(((pend >> (5 * k)) & 31u) - 1u) * w.rec_cap_net;
                unsigned long long m = vmask[k], slot = first;
                while (m) {
                    const uint32_t v = (uint32_t)(__ffsll(m) - 1);
                    m &= m - 1ull;
                    if (slot < w.rec_cap_net) vertex_record(w, rec, v, lf, w.train_rec + 10 * (region + slot));
                    ++slot;
                }
            });
    if (threadIdx.x == 0 && n_in) {
        atomicAdd(&w.ctr->done_paths, (unsigned long long)n_in);
        if (w.recs) w.tile_free[t] = free0 + n_in;
    }
}

template <int BLOCK>
hipError_t iteration(const WfState &w, uint32_t parity, bool guided, float *film, unsigned long long *cnt,
                     hipStream_t st) {
    const uint32_t grid = w.n_tiles;
    wf_regen_kernel<BLOCK><<<grid, BLOCK, 0, st>>>(w, parity);
    if (w.refill)
        wf_extend_refill_kernel<BLOCK><<<grid, BLOCK, 0, st>>>(w, parity, guided ? 1 : 0, cnt);
    else
        wf_extend_kernel<BLOCK><<<grid, BLOCK, 0, st>>>(w, parity, guided ? 1 : 0, cnt);
    if (guided) {
        hipError_t e = guide_launch_infer_nets(w.nets, w.n_nets, w.gstage, w.guide_slot, &w.ctr->guide_count[0][0],
                                               kCountStride, w.capacity, w.guide_out, st);
        if (e != hipSuccess) return e;
    }
    if (guided || w.recs)
        wf_shade_kernel<BLOCK, true><<<grid, BLOCK, 0, st>>>(w, parity);
    else
        wf_shade_kernel<BLOCK, false><<<grid, BLOCK, 0, st>>>(w, parity);
    wf_nee_kernel<BLOCK><<<grid, BLOCK, 0, st>>>(w, parity, cnt);
    wf_finish_kernel<BLOCK><<<grid, BLOCK, 0, st>>>(w, film);
    return hipGetLastError();
}

// Pixel i of the batch's unfiltered pixels: its job (the last with
// pix_begin <= i; filtered jobs span no pixels) and its samples
// path_begin + s * n_pix + p, s = 0 .. spp - 1, summed in that order and added
// to the film.
__global__ void wf_film_reduce_kernel(WfState w, unsigned long long n_pix_all, float *__restrict__ film) {
    const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pix_all) return;
    uint32_t lo = 0, hi = w.n_jobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (w.jobs[mid].pix_begin <= i) lo = mid;
        else hi = mid;
    }
    const WfJob &job = w.jobs[lo];
    const unsigned long long p = i - job.pix_begin;
    const unsigned long long n_pix = (unsigned long long)job.cam.width * job.cam.height;
    float r = 0.f, g = 0.f, b = 0.f;
    for (uint32_t s = 0; s < job.spp; ++s) {
        const float4 v = w.smp[job.path_begin + s * n_pix + p];
        r += v.x;
        g += v.y;
        b += v.z;
    }
    float *px = film + job.film_offset + 3 * p;
    px[0] += r;
    px[1] += g;
    px[2] += b;
}

__global__ void wf_clear_records_kernel(WfCounters *c, int net) {
    for (int k = 0; k < kMaxNets; ++k)
        if (net < 0 || k == net) c->rec_count[k][0] = 0ull;
}

}  // namespace

hipError_t wf_launch_reset(const WfState &w, unsigned long long total_paths, int block, hipStream_t st) {
    switch (block) {
        case 64: wf_reset_kernel<64><<<w.n_tiles, 64, 0, st>>>(w, total_paths); break;
        case 256: wf_reset_kernel<256><<<w.n_tiles, 256, 0, st>>>(w, total_paths); break;
        default: wf_reset_kernel<128><<<w.n_tiles, 128, 0, st>>>(w, total_paths); break;
    }
    return hipGetLastError();
}

hipError_t wf_launch_film_reduce(const WfState &w, unsigned long long n_pix, float *film, hipStream_t st) {
    if (n_pix == 0) return hipSuccess;
    wf_film_reduce_kernel<<<(unsigned)((n_pix + 255) / 256), 256, 0, st>>>(w, n_pix, film);
    return hipGetLastError();
}

hipError_t wf_launch_clear_records(const WfState &w, int net, hipStream_t st) {
    wf_clear_records_kernel<<<1, 1, 0, st>>>(w.ctr, net);
    return hipGetLastError();
}

hipError_t wf_launch_iteration(const WfState &w, uint32_t parity, bool guided, float *film,
                               unsigned long long *ray_counts, int block, hipStream_t st) {
    switch (block) {
        case 64: return iteration<64>(w, parity, guided, film, ray_counts, st);
        case 256: return iteration<256>(w, parity, guided, film, ray_counts, st);
        default: return iteration<128>(w, parity, guided, film, ray_counts, st);
    }
}

}  // namespace illume
