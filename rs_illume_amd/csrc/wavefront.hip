// Wavefront path tracer: the bounce loop of PathIntegrator::li split into
// per-stage kernels over a pool of path slots whose state lives in HBM as
// structure-of-arrays, so a batched MFMA stage (the learned proposal's
// inference) can run between the closest-hit query and shading.
//
//   start   camera samples -> path slots
//   extend  closest hit of every active slot; slots whose hit needs the
//           proposal append (slot, normalised position, normal) to a queue
//   infer   MLP over the queue (guide.hip), count read on the device
//   shade   one bounce of path_shade (NEE + MIS rays traced inline); finished
//           slots add their radiance to the film
//   records after the last bounce: per-vertex training samples
//           (incident radiance / pdf) appended to the training buffer
//
// Every kernel reads its work count from device memory and exits early, so
// the host can launch fixed grids (and capture the loop in a graph).
#include <hip/hip_runtime.h>
#include "kernels.h"
#include "path.h"
#include "wavefront.h"

#pragma clang fp contract(off)

namespace illume {

namespace {

constexpr int kBlock = 128;

__device__ __forceinline__ void store_state(const WfState &w, uint32_t p, const PathState &ps) {
    w.o_tmax[p] = make_float4(ps.o.x, ps.o.y, ps.o.z, ps.t_max);
    w.d_eta[p] = make_float4(ps.d.x, ps.d.y, ps.d.z, ps.eta_scale);
    w.beta[p] = make_float4(ps.beta.x, ps.beta.y, ps.beta.z, 0.f);
    w.L[p] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.f);
    w.rng[p] = make_ulonglong2(ps.rng.state, ps.rng.inc);
    w.flags[p] = make_uint2(ps.bounces, (ps.active ? 1u : 0u) | (ps.specular ? 2u : 0u));
}
__device__ __forceinline__ void load_state(const WfState &w, uint32_t p, PathState &ps) {
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
    ps.bounces = f.x;
    ps.active = (f.y & 1u) != 0;
    ps.specular = (f.y & 2u) != 0;
}

__global__ __launch_bounds__(kBlock) void wf_start_kernel(WfState w, DevCamera cam, uint32_t n_paths, uint64_t seed,
                                                          uint64_t first_path) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n_paths) return;
    const uint64_t n_pix = (uint64_t)cam.width * cam.height;
    const uint64_t g = first_path + p;
    PathState ps;
    path_start(cam, seed, g % n_pix, (uint32_t)(g / n_pix), ps);
    store_state(w, p, ps);
    w.pix[p] = (uint32_t)(g % n_pix);
    w.n_recs[p] = 0;
    if (p == 0) *w.guide_count = 0;
}

__global__ __launch_bounds__(kBlock) void wf_extend_kernel(WfState w, DevScene sc, uint32_t n_paths, int guided,
                                                           GuideCfg gc, unsigned long long *__restrict__ counters) {
    __shared__ int lds_stack[kStackLds * kBlock];
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    uint32_t n_closest = 0;
    if (p < n_paths && (w.flags[p].y & 1u)) {
        const float4 a = w.o_tmax[p], b = w.d_eta[p];
        TraceStack<kBlock> st;
        st.reset(lds_stack + threadIdx.x);
        TraceHit h;
        uint32_t nb = 0, np = 0;
        const f3 o = mk(a.x, a.y, a.z), d = mk(b.x, b.y, b.z);
        const bool found = trace_ray<false, false, kBlock>(sc.nodes, sc.prims, o, d, a.w, h, st, nb, np);
        ++n_closest;
        w.hit0[p] = make_float4(__uint_as_float(found ? h.prim : 0xFFFFFFFFu), __uint_as_float(h.slot), h.t, h.b0);
        w.hit1[p] = make_float4(h.b1, h.b2, 0.f, 0.f);
        int gslot = -1;
        if (guided && found) {
            const DevMaterial &m = sc.materials[sc.prim_material[h.prim]];
            const bool nonspec = (m.type == kMetal) ||
                                 (m.type == kMatte && !(m.kd[0] == 0.f && m.kd[1] == 0.f && m.kd[2] == 0.f));
            if (nonspec && (w.flags[p].x < w.max_depth)) {
                Isect it;
                hit_isect(sc, h, o, d, it);
                gslot = (int)atomicAdd(w.guide_count, 1u);
                float *in = w.guide_in + 6 * (size_t)gslot;
                in[0] = (it.p.x - gc.bounds_lo[0]) * gc.inv_ext[0];
                in[1] = (it.p.y - gc.bounds_lo[1]) * gc.inv_ext[1];
                in[2] = (it.p.z - gc.bounds_lo[2]) * gc.inv_ext[2];
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

__global__ __launch_bounds__(kBlock) void wf_shade_kernel(WfState w, DevScene sc, uint32_t n_paths, GuideCfg gc,
                                                          float *__restrict__ rgb,
                                                          unsigned long long *__restrict__ counters) {
    __shared__ int lds_stack[kStackLds * kBlock];
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p == 0) *w.guide_count = 0;  // the next extend stage refills the proposal queue
    PathCounters pc{0, 0};
    if (p < n_paths && (w.flags[p].y & 1u)) {
        TraceStack<kBlock> st;
        st.reset(lds_stack + threadIdx.x);
        PathState ps;
        load_state(w, p, ps);
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
        const float *gout = (gc.out && gslot >= 0) ? gc.out + 40 * (size_t)gslot : nullptr;
        uint32_t n_recs = w.n_recs[p];
        VertexRec *recs = w.recs ? w.recs + (size_t)p * w.max_depth : nullptr;
        path_shade<kBlock>(sc, w.max_depth, ps, found, h, gout, gc, recs, n_recs, st, pc);
        w.n_recs[p] = n_recs;
        store_state(w, p, ps);
        if (!ps.active) {
            const f3 L = guard_radiance(ps.L);
            const uint32_t pix = w.pix[p];
            atomicAdd(rgb + 3 * (size_t)pix, L.x);
            atomicAdd(rgb + 3 * (size_t)pix + 1, L.y);
            atomicAdd(rgb + 3 * (size_t)pix + 2, L.z);
            w.L[p] = make_float4(L.x, L.y, L.z, 0.f);
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

// Training samples: for vertex k with a snapshot, the radiance that arrived
// along its sampled direction is (L_final - L_snap) / beta_after; the sample
// weight is its luminance over the sampling pdf.
__global__ __launch_bounds__(kBlock) void wf_records_kernel(WfState w, uint32_t n_paths, float *__restrict__ out,
                                                            uint32_t *__restrict__ out_count, uint32_t capacity) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n_paths || !w.recs) return;
    const float4 Lf = w.L[p];
    const uint32_t n = w.n_recs[p];
    for (uint32_t k = 0; k < n; ++k) {
        const VertexRec &r = w.recs[(size_t)p * w.max_depth + k];
        if (r.lsnap[0] < 0.f) continue;
        float li[3];
        const float lf[3] = {Lf.x, Lf.y, Lf.z};
        for (int c = 0; c < 3; ++c) li[c] = r.beta[c] > 0.f ? fmaxf(0.f, lf[c] - r.lsnap[c]) / r.beta[c] : 0.f;
        const float wgt = (0.212671f * li[0] + 0.715160f * li[1] + 0.072169f * li[2]) / r.pdf;
        if (!(wgt > 0.f) || !__builtin_isfinite(wgt)) continue;
        const uint32_t slot = atomicAdd(out_count, 1u);
        if (slot >= capacity) continue;
        float *o = out + 10 * (size_t)slot;
        for (int c = 0; c < 6; ++c) o[c] = r.in[c];
        o[6] = r.wi[0];
        o[7] = r.wi[1];
        o[8] = r.wi[2];
        o[9] = wgt;
    }
}

}  // namespace

hipError_t wf_launch_start(const WfState &w, const DevCamera &cam, uint32_t n_paths, uint64_t seed,
                           uint64_t first_path, hipStream_t st) {
    wf_start_kernel<<<(n_paths + kBlock - 1) / kBlock, kBlock, 0, st>>>(w, cam, n_paths, seed, first_path);
    return hipGetLastError();
}
hipError_t wf_launch_extend(const WfState &w, const DevScene &sc, uint32_t n_paths, bool guided, const GuideCfg &gc,
                            unsigned long long *counters, hipStream_t st) {
    wf_extend_kernel<<<(n_paths + kBlock - 1) / kBlock, kBlock, 0, st>>>(w, sc, n_paths, guided ? 1 : 0, gc,
                                                                         counters);
    return hipGetLastError();
}
hipError_t wf_launch_shade(const WfState &w, const DevScene &sc, uint32_t n_paths, const GuideCfg &gc, float *rgb,
                           unsigned long long *counters, hipStream_t st) {
    wf_shade_kernel<<<(n_paths + kBlock - 1) / kBlock, kBlock, 0, st>>>(w, sc, n_paths, gc, rgb, counters);
    return hipGetLastError();
}
hipError_t wf_launch_records(const WfState &w, uint32_t n_paths, float *out, uint32_t *out_count, uint32_t capacity,
                             hipStream_t st) {
    wf_records_kernel<<<(n_paths + kBlock - 1) / kBlock, kBlock, 0, st>>>(w, n_paths, out, out_count, capacity);
    return hipGetLastError();
}

}  // namespace illume
