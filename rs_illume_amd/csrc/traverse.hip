// Batched ray queries over the traversal steps of trace.h: closest and any
// hit on the four-child tree, and any hit on the two-child tree, by persistent
// waves that refill their lanes from a static run of the batch, then from runs
// handed out by a device queue; closest hit on the two-child tree and the
// instrumented (counting) traversal one ray per lane.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <mutex>
#include "kernels.h"
#include "trace.h"

#pragma clang fp contract(off)

// Device-resident scratch of the non-allocating closest-hit entry point
// (illume_intersect_closest_noalloc): word 0 is the persistent launch's ray
// queue counter. Each device holds its own copy until the process ends; host
// code uses the address hipGetSymbolAddress returns for the current device,
// never this name. Calls that overlap on one device share it and race.
static __device__ __align__(256) unsigned long long scratch[illume::kTraverseScratchWords];

namespace illume {

namespace {

// Launch shape of the one-ray-per-lane kernel (measured on the 2^20-triangle
// / 2^24-ray workload): 64-lane blocks limited to 64 VGPRs (8 waves per SIMD),
// which keeps more node fetches in flight (two-child tree: any hit 9.9 ->
// 9.3 ms against 128-lane blocks at its natural 72-80 VGPRs).
template <int WIDTH>
struct TraverseShape {
    static constexpr int kBlock = 64, kWaves = 8;
};

template <bool ANY, bool COUNT, int WIDTH, int kBlock = TraverseShape<WIDTH>::kBlock>
__global__ __launch_bounds__(kBlock, TraverseShape<WIDTH>::kWaves) void traverse_kernel(
    const void *__restrict__ nodes, const PrimRecord *__restrict__ prims, const illume_ray *__restrict__ rays,
    uint32_t n, illume_hit *__restrict__ hits, uint8_t *__restrict__ occluded, uint32_t *__restrict__ counts) {
    __shared__ int lds_stack[kStackLds * kBlock];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 *rp = reinterpret_cast<const float4 *>(rays + i);
    const float4 r0 = rp[0], r1 = rp[1];
    int spill[kSpill];
    TraceStack<kBlock> st;
    st.reset(lds_stack + threadIdx.x, spill);
    TraceHit h;
    uint32_t n_box = 0, n_prim = 0;
    const f3 o = mk(r0.x, r0.y, r0.z), d = mk(r0.w, r1.x, r1.y);
    const bool found =
        WIDTH == 4
            ? trace_ray4<ANY, COUNT, kBlock>((const GpuNode4Q *)nodes, prims, o, d, r1.z, h, st, n_box, n_prim)
            : trace_ray<ANY, COUNT, kBlock>((const GpuNode *)nodes, prims, o, d, r1.z, h, st, n_box, n_prim);
    if (COUNT) {
        counts[2 * (size_t)i] = n_box;
        counts[2 * (size_t)i + 1] = n_prim;
    } else if (ANY) {
        occluded[i] = found ? 1 : 0;
    } else {
        reinterpret_cast<float4 *>(hits)[i] = make_float4(__uint_as_float(h.prim), h.t, h.b1, h.b2);
    }
}

// Persistent batch traversal: each wave (one 64-lane block) starts on a
// contiguous run of rays. A lane that finishes its ray keeps its result until
// at least kRefillMin lanes of the wave are idle; then those lanes write their
// results and take the run's next rays (in lane order), so a wave's trip count
// follows its rays' total work instead of each group of 64 rays' slowest ray.
// The grid is one wave per wave slot of the device. Waiting for three eighths
// of the wave to be idle amortises the refill (ray load, reciprocal direction,
// shear, result write) over many lanes: on the 2^20-triangle / 2^24-ray
// workload 24 idle lanes gave 7.08 ms per launch, 20 7.16, 28 7.11, 32 7.19
// and 16 7.28 (closest hit; any hit 4.35, 4.39, 4.34, 4.36, 4.49). Before
// the four-child visit chose its near and far planes by the direction's sign
// (a cheaper node step) the best was 20.
constexpr int kRefillMin = 24;

// Load balance: the static runs cover 1/kStaticDiv of the batch (an equal run
// per wave); the rest is handed out kQueueRun rays at a time from a counter in
// device memory, taken by a wave whose run cannot feed its idle lanes. Waves
// end within one run of each other instead of with their static shares'
// spread: on the 2^20-triangle / 2^24-ray workload closest hit 8.65 -> 7.39 ms
// per launch and any hit 5.23 -> 4.56 ms (static fractions 0.4-0.55 within 1%
// of each other, 0.7 at 7.98 ms, all-static at 8.65; runs of 128 or 256 rays
// 1-2% slower than 64; with the current node visit a static quarter times the
// same as a static half, 7.155 against 7.150 ms).
constexpr uint32_t kStaticDiv = 2;
constexpr uint32_t kQueueRun = 64;  // at least one wave, so one run feeds every idle lane

// Waves per SIMD: at most 64 VGPRs, no spills. Fewer waves with more registers
// lose latency hiding: 7 waves (72 VGPRs) 7.21 ms and 6 waves (80) 7.60 ms per
// closest-hit launch against 7.15 ms at 8 on the 2^24-ray workload.
constexpr int kRefillWaves = 8;

template <bool ANY, int WIDTH>
__global__ __launch_bounds__(64, kRefillWaves) void traverse_refill_kernel(
    const void *__restrict__ nodes, const PrimRecord *__restrict__ prims, const illume_ray *__restrict__ rays,
    uint32_t n, uint32_t per_wave, uint32_t grab, unsigned long long *__restrict__ queue,
    illume_hit *__restrict__ hits, uint8_t *__restrict__ occluded) {
    constexpr int kBlock = 64;
    constexpr uint32_t kNoRay = 0xFFFFFFFFu;
    __shared__ int lds_stack[kStackLds * kBlock];
    const uint32_t lane = threadIdx.x;
    const uint32_t begin = min(n, blockIdx.x * per_wave);
    uint32_t end = min(n, begin + per_wave);
    uint32_t next = begin;  // the wave's next ray and the end of its run (wave-uniform)
    bool more = true;       // the shared queue may still hold rays
    int spill[kSpill];
    TraceStack<kBlock> st;
    st.reset(lds_stack + lane, spill);
    uint32_t ray = kNoRay, slot = kNoSlot, pf = 0, pn = 0, n_box = 0;
    bool busy = false;
    f3 o = mk(0.f, 0.f, 0.f), d = o, inv = o;
    Shear sh{0, 0.f, 0.f, 0.f};
    float t_max = 0.f;
    int cur = kTraceDone;
    while (true) {
        const uint64_t idle = __ballot(!busy);
        const int n_idle = __popcll(idle);
        if (n_idle >= kRefillMin || n_idle == kBlock) {
            // The run cannot feed every idle lane: take the queue's next run.
            const uint32_t left = end - next;
            uint32_t run = kNoRay;
            if (left < (uint32_t)n_idle && more) {
                unsigned long long v = 0;
                if (lane == 0) v = atomicAdd(queue, (unsigned long long)grab);
                v = __shfl(v, 0);
                if (v >= n) more = false;
                else run = (uint32_t)v;
            }
            const uint32_t run_end = run == kNoRay ? 0u : min(n, run + grab);
            if (!busy) {
                if (ray != kNoRay) {
                    if (ANY) {
                        occluded[ray] = slot != kNoSlot ? 1 : 0;
                    } else {
                        TraceHit h;
                        h.prim = 0xFFFFFFFFu;
                        h.t = h.b0 = h.b1 = h.b2 = 0.f;
                        if (slot != kNoSlot) finish_hit(prims, slot, o, sh, t_max, h);
                        reinterpret_cast<float4 *>(hits)[ray] = make_float4(__uint_as_float(h.prim), h.t, h.b1, h.b2);
                    }
                    ray = kNoRay;
                }
                const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                uint32_t r = kNoRay;
                if (rank < left) r = next + rank;
                else if (run != kNoRay && run + (rank - left) < run_end) r = run + (rank - left);
                if (r != kNoRay) {
                    const float4 *rp = reinterpret_cast<const float4 *>(rays + r);
                    const float4 r0 = rp[0], r1 = rp[1];
                    o = mk(r0.x, r0.y, r0.z);
                    d = mk(r0.w, r1.x, r1.y);
                    t_max = r1.z;
                    inv = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
                    sh = make_shear(d);
                    slot = kNoSlot;
                    cur = 0;
                    pf = pn = 0;
                    st.sp = 0;
                    ray = r;
                    busy = true;
                }
            }
            if (run != kNoRay) {
                next = min(run + ((uint32_t)n_idle - left), run_end);
                end = run_end;
            } else {
                next = min(next + (uint32_t)n_idle, end);
            }
            if (__ballot(busy) == 0) break;  // queue drained, every result written
        }
        if (busy) {
            if (pn != 0) {
                float t;
                if (prim_step<ANY>(prims, pf, o, d, sh, t_max, t)) {
                    slot = pf;
                    t_max = t;
                    if (ANY) {
                        pn = 1;
                        cur = kTraceDone;
                    }
                }
                ++pf;
                --pn;
            } else if (WIDTH == 4) {
                cur = node_step4<false, kBlock>((const GpuNode4Q *)nodes, cur, o, inv, t_max, st, n_box);
            } else {
                cur = node_step2<false, kBlock>((const GpuNode *)nodes, cur, o, inv, t_max, st, n_box);
            }
            take_leaf(cur, pf, pn, st);
            if (pn == 0 && cur == kTraceDone) busy = false;
        }
    }
}

__global__ void queue_start_kernel(unsigned long long *queue, unsigned long long start) { *queue = start; }

__global__ void fill_miss_kernel(uint32_t n, illume_hit *hits, uint8_t *occluded, uint32_t *counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (hits) reinterpret_cast<float4 *>(hits)[i] = make_float4(__uint_as_float(ILLUME_MISS), 0.f, 0.f, 0.f);
    if (occluded) occluded[i] = 0;
    if (counts) {
        counts[2 * (size_t)i] = 0;
        counts[2 * (size_t)i + 1] = 0;
    }
}

}  // namespace

int max_supported_depth() { return kMaxTraceDepth; }

// Persistent launch (either tree width): one 64-lane wave per wave slot
// (compute units x 4 SIMDs x waves per SIMD), each starting on an equal static
// run; the queue counter (stream-ordered allocation, set by a one-thread
// kernel so that the launch sequence can be captured in a graph) starts past
// the static runs.
//
// The counter lives in an 8-byte stream-ordered allocation made and freed by
// the call, or, when `static_queue` is given, in that device word (the
// non-allocating entry point's scratch).
template <int W>
hipError_t launch_refill(const void *nodes, const PrimRecord *prims, const illume_ray *rays, uint32_t n,
                         illume_hit *hits, uint8_t *occluded, bool any, unsigned long long *static_queue,
                         hipStream_t stream) {
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    const uint32_t slots = (uint32_t)std::max(cus, 1) * 4u * (uint32_t)kRefillWaves;
    const uint32_t waves = std::min((n + 63u) / 64u, slots);
    const uint32_t per_wave = (n + waves - 1) / waves / kStaticDiv;
    const uint32_t grab = kQueueRun;
    unsigned long long *queue = static_queue;
    if (!queue) {
        e = hipMallocAsync((void **)&queue, sizeof(unsigned long long), stream);
        if (e != hipSuccess) return e;
    }
    queue_start_kernel<<<1, 1, 0, stream>>>(queue, (unsigned long long)waves * per_wave);
    if (any)
        traverse_refill_kernel<true, W><<<waves, 64, 0, stream>>>(nodes, prims, rays, n, per_wave, grab, queue,
                                                                  nullptr, occluded);
    else
        traverse_refill_kernel<false, W><<<waves, 64, 0, stream>>>(nodes, prims, rays, n, per_wave, grab, queue,
                                                                   hits, nullptr);
    e = hipGetLastError();
    if (static_queue) return e;
    const hipError_t f = hipFreeAsync(queue, stream);
    return e != hipSuccess ? e : f;
}

// Address of this device's `scratch` array, looked up once per device (with
// the device current) and kept; hipGetSymbolAddress allocates nothing.
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned long long *g_scratch_ptr[kMaxScratchDevices] = {};

hipError_t scratch_address(unsigned long long **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<unsigned long long *>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

// One ray per lane: counting queries, and closest hit on the two-child tree
// (whose cheaper steps do not repay the refill: 13.5 against 12.6 ms per
// launch on the 2^20-triangle / 2^24-ray workload, with the queue; any hit
// there takes the refill kernel, 8.84 against 9.3 ms).
template <int W>
hipError_t launch_per_ray(const void *nodes, const PrimRecord *prims, const illume_ray *rays, uint32_t n,
                          illume_hit *hits, uint8_t *occluded, uint32_t *counts, bool any, hipStream_t stream) {
    constexpr int kBlock = TraverseShape<W>::kBlock;
    const uint32_t grid = (n + kBlock - 1) / kBlock;
    if (counts) {
        if (any)
            traverse_kernel<true, true, W><<<grid, kBlock, 0, stream>>>(nodes, prims, rays, n, nullptr, nullptr, counts);
        else
            traverse_kernel<false, true, W><<<grid, kBlock, 0, stream>>>(nodes, prims, rays, n, nullptr, nullptr, counts);
    } else if (any) {
        traverse_kernel<true, false, W><<<grid, kBlock, 0, stream>>>(nodes, prims, rays, n, nullptr, occluded, nullptr);
    } else {
        traverse_kernel<false, false, W><<<grid, kBlock, 0, stream>>>(nodes, prims, rays, n, hits, nullptr, nullptr);
    }
    return hipGetLastError();
}

hipError_t launch_traverse(const void *nodes, int width, const PrimRecord *prims, bool empty, const illume_ray *rays,
                           uint32_t n, illume_hit *hits, uint8_t *occluded, uint32_t *counts, bool any,
                           hipStream_t stream) {
    if (n == 0) return hipSuccess;
    if (empty) {
        fill_miss_kernel<<<(n + 255) / 256, 256, 0, stream>>>(n, counts ? nullptr : (any ? nullptr : hits),
                                                              counts ? nullptr : (any ? occluded : nullptr), counts);
        return hipGetLastError();
    }
    if (width == 4 && !counts) return launch_refill<4>(nodes, prims, rays, n, hits, occluded, any, nullptr, stream);
    if (any && !counts) return launch_refill<2>(nodes, prims, rays, n, hits, occluded, any, nullptr, stream);
    if (width == 4) return launch_per_ray<4>(nodes, prims, rays, n, hits, occluded, counts, any, stream);
    return launch_per_ray<2>(nodes, prims, rays, n, hits, occluded, counts, any, stream);
}

hipError_t launch_closest_noalloc(const void *nodes, int width, const PrimRecord *prims, bool empty,
                                  const illume_ray *rays, uint32_t n, illume_hit *hits, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    if (empty) {
        fill_miss_kernel<<<(n + 255) / 256, 256, 0, stream>>>(n, hits, nullptr, nullptr);
        return hipGetLastError();
    }
    // the two-child tree's closest hit runs one ray per lane and needs no queue
    if (width != 4) return launch_per_ray<2>(nodes, prims, rays, n, hits, nullptr, nullptr, false, stream);
    unsigned long long *queue = nullptr;
    const hipError_t e = scratch_address(&queue);
    if (e != hipSuccess) return e;
    return launch_refill<4>(nodes, prims, rays, n, hits, nullptr, false, queue, stream);
}

}  // namespace illume
