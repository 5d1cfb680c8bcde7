// Device-side BVH traversal shared by the intersection kernels and the path
// tracer: closest hit (BVHAccel::intersect, accelerators/bvh.rs:563-612) and
// any hit (BVHAccel::intersect_p, bvh.rs:614-662) over the 64-byte
// two-child layout of layout.h.
//
// Each visit of an inner node reads both child boxes (one 64-byte record),
// tests them with the conservative slab test, descends into the nearer hit
// child and pushes the farther. The stack is a ring in LDS (kStackLds entries
// per lane, laid out [slot][lane] so a wave's pushes and pops hit 64 distinct
// banks); entries deeper than the ring spill to a private array that only
// rays with unusually deep stacks touch.
#pragma once
#include <hip/hip_runtime.h>
#include "geom.h"
#include "layout.h"

#pragma clang fp contract(off)

namespace illume {

constexpr int kStackLds = 16;  // power of two
constexpr int kSpill = 64;
constexpr int kMaxTraceDepth = kStackLds + kSpill;

// The overflow array is the caller's own local array, not a member: an
// array indexed by a run-time value lives in scratch memory, and as a member
// it would take the stack pointer and the LDS pointer with it (a scratch load
// and store of sp on every push and pop).
template <int BLOCK>
struct TraceStack {
    int *lds;    // this lane's column: lds[slot * BLOCK]
    int *spill;  // this lane's kSpill-entry overflow
    int sp;
    __device__ __forceinline__ void reset(int *column, int *overflow) {
        lds = column;
        spill = overflow;
        sp = 0;
    }
    __device__ __forceinline__ void push(int v) {
        int slot = sp & (kStackLds - 1);
        if (sp >= kStackLds) spill[sp - kStackLds] = lds[slot * BLOCK];
        lds[slot * BLOCK] = v;
        ++sp;
    }
    __device__ __forceinline__ int pop() {
        --sp;
        int slot = sp & (kStackLds - 1);
        int v = lds[slot * BLOCK];
        if (sp >= kStackLds) lds[slot * BLOCK] = spill[sp - kStackLds];
        return v;
    }
};

struct TraceHit {
    uint32_t prim;  // caller's primitive id, ILLUME_MISS-equivalent 0xFFFFFFFF on a miss
    uint32_t slot;  // primitive record index (leaf order)
    float t, b0, b1, b2;
};

// One primitive test of the traversal loop: on a hit nearer than t_max
// (any hit for ANY; closest hit ignores degenerate triangles, as
// Triangle::intersect does), returns true with its distance.
template <bool ANY>
__device__ __forceinline__ bool prim_step(const PrimRecord *__restrict__ prims, uint32_t pf, f3 o, f3 d,
                                          const Shear &sh, float t_max, float &t) {
    const float4 *pp = reinterpret_cast<const float4 *>(prims + pf);
    const float4 pa = pp[0];
    const uint32_t tag = __float_as_uint(pa.w);
    if (tag & kSphereBit) {
        const float4 pb = pp[1];
        return sphere_test(mk(pa.x, pa.y, pa.z), pb.x, o, d, t_max, t);
    }
    const float4 pb = pp[1], pc = pp[2];
    const f3 p0 = mk(pa.x, pa.y, pa.z), p1 = mk(pb.x, pb.y, pb.z), p2 = mk(pc.x, pc.y, pc.z);
    float b0, b1, b2;
    return triangle_test_sheared(sh, p0, p1, p2, o, t_max, t, b0, b1, b2) &&
           (ANY || !triangle_degenerate(p0, p1, p2));
}

// The hit record of primitive record `slot` hit at distance t: the triangle
// test again on the same operands gives the same barycentrics (nothing in it
// but the accept/reject comparisons depends on t_max).
__device__ __forceinline__ void finish_hit(const PrimRecord *__restrict__ prims, uint32_t slot, f3 o, const Shear &sh,
                                           float t, TraceHit &hit) {
    const float4 *pp = reinterpret_cast<const float4 *>(prims + slot);
    const float4 pa = pp[0];
    const uint32_t tag = __float_as_uint(pa.w);
    hit.slot = slot;
    hit.t = t;
    if (tag & kSphereBit) {
        hit.prim = tag & ~kSphereBit;
        return;
    }
    const float4 pb = pp[1], pc = pp[2];
    float tt;
    triangle_test_sheared(sh, mk(pa.x, pa.y, pa.z), mk(pb.x, pb.y, pb.z), mk(pc.x, pc.y, pc.z), o, __builtin_inff(), tt,
                          hit.b0, hit.b1, hit.b2);
    hit.prim = tag;
}

constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
constexpr int kTraceDone = (int)0x80000000;  // neither a node (>= 0) nor a leaf reference

// A leaf reached (or popped) becomes the current leaf: slots pf .. pf + pn - 1;
// the node after it is popped at once (the stack order is unchanged).
template <int BLOCK>
__device__ __forceinline__ void take_leaf(int &cur, uint32_t &pf, uint32_t &pn, TraceStack<BLOCK> &st) {
    if (pn == 0 && cur < 0 && cur != kTraceDone) {
        pf = leaf_first(cur);
        pn = leaf_count(cur);
        cur = st.sp == 0 ? kTraceDone : st.pop();
    }
}

// One visit of a two-child node: tests both child boxes, pushes the farther
// of two entered children and returns the nearer; with none entered, returns
// the top of the stack or kTraceDone.
template <bool COUNT, int BLOCK>
__device__ __forceinline__ int node_step2(const GpuNode *__restrict__ nodes, int cur, f3 o, f3 inv, float t_max,
                                          TraceStack<BLOCK> &st, uint32_t &n_box) {
    const float4 *np = reinterpret_cast<const float4 *>(nodes + cur);
    const float4 a = np[0], b = np[1], c = np[2];
    const int4 l = reinterpret_cast<const int4 *>(np)[3];
    if (COUNT) n_box += 2;
    float t0, t1;
    const bool h0 = slab(a.x, a.y, a.z, a.w, c.x, c.y, o, inv, t_max, t0);
    const bool h1 = slab(b.x, b.y, b.z, b.w, c.z, c.w, o, inv, t_max, t1);
    if (h0 && h1) {
        const bool swap = t1 < t0;
        st.push(swap ? l.x : l.y);
        return swap ? l.y : l.x;
    }
    if (h0) return l.x;
    if (h1) return l.y;
    return st.sp == 0 ? kTraceDone : st.pop();
}

// Closest hit (ANY = false) or any hit (ANY = true) of the ray (o, d) in
// (0, t_max] over the two-child tree. Returns true on a hit. COUNT adds box /
// primitive tests to the counters. One step per iteration, as in trace_ray4
// below.
template <bool ANY, bool COUNT, int BLOCK>
__device__ __forceinline__ bool trace_ray(const GpuNode *__restrict__ nodes, const PrimRecord *__restrict__ prims, f3 o,
                                          f3 d, float t_max, TraceHit &hit, TraceStack<BLOCK> &st, uint32_t &n_box,
                                          uint32_t &n_prim) {
    const f3 inv = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
    const Shear sh = make_shear(d);
    hit.prim = 0xFFFFFFFFu;
    hit.slot = 0;
    hit.t = hit.b0 = hit.b1 = hit.b2 = 0.f;
    uint32_t slot = kNoSlot;
    int cur = 0;
    uint32_t pf = 0, pn = 0;
    st.sp = 0;
    while (true) {
        if (pn != 0) {
            if (COUNT) ++n_prim;
            float t;
            if (prim_step<ANY>(prims, pf, o, d, sh, t_max, t)) {
                slot = pf;
                t_max = t;
                if (ANY) {
                    // answered: leave through the loop's one exit
                    pn = 1;
                    cur = kTraceDone;
                }
            }
            ++pf;
            --pn;
        } else {
            cur = node_step2<COUNT, BLOCK>(nodes, cur, o, inv, t_max, st, n_box);
        }
        take_leaf(cur, pf, pn, st);
        if (pn == 0 && cur == kTraceDone) break;
    }
    if (slot == kNoSlot) return false;
    hit.slot = slot;
    if (!ANY) finish_hit(prims, slot, o, sh, t_max, hit);
    return true;
}

// Compare-exchange of (distance, reference) pairs for the 4-child order.
__device__ __forceinline__ void cswap(float &ta, int &ra, float &tb, int &rb) {
    const bool s = tb < ta;
    const float t = s ? tb : ta;
    tb = s ? ta : tb;
    ta = t;
    const int r = s ? rb : ra;
    rb = s ? ra : rb;
    ra = r;
}

// Box of child C of a quantized node (words w1 = lo_x hi_x lo_y hi_y, w2.x/y =
// lo_z hi_z; byte C of each) entered by the ray? Planes are fma(q, step, origin).
template <int C>
__device__ __forceinline__ bool qslab(const uint4 &w1, const uint4 &w2, float ox, float oy, float oz, float sx,
                                      float sy, float sz, f3 o, f3 inv, float t_max, float &t) {
    constexpr int sh = 8 * C;
    const float lx = fmaf((float)((w1.x >> sh) & 0xffu), sx, ox), hx = fmaf((float)((w1.y >> sh) & 0xffu), sx, ox);
    const float ly = fmaf((float)((w1.z >> sh) & 0xffu), sy, oy), hy = fmaf((float)((w1.w >> sh) & 0xffu), sy, oy);
    const float lz = fmaf((float)((w2.x >> sh) & 0xffu), sz, oz), hz = fmaf((float)((w2.y >> sh) & 0xffu), sz, oz);
    return slab(lx, hx, ly, hy, lz, hz, o, inv, t_max, t);
}

// The same test with each axis's near and far plane already chosen (near.x,
// far.x: the bytes of x's entry and exit planes, likewise y, z): an entry is
// (near - o) inv and an exit (far - o) inv, the values the min and max of
// qslab pick whenever neither is NaN. A NaN (0 * inf: the origin on a plane
// the ray runs along) drops out of the entry maximum and the exit minimum,
// so such a box is entered whenever qslab enters it.
template <int C>
__device__ __forceinline__ bool qslab_nf(const uint4 &near, const uint4 &far, float ox, float oy, float oz, float sx,
                                         float sy, float sz, f3 o, f3 inv, float t_max, float &t) {
    constexpr int sh = 8 * C;
    constexpr float k = 1.0f + 2.0f * gamma_f(3);
    const float ax = (fmaf((float)((near.x >> sh) & 0xffu), sx, ox) - o.x) * inv.x;
    const float bx = (fmaf((float)((far.x >> sh) & 0xffu), sx, ox) - o.x) * inv.x;
    const float ay = (fmaf((float)((near.y >> sh) & 0xffu), sy, oy) - o.y) * inv.y;
    const float by = (fmaf((float)((far.y >> sh) & 0xffu), sy, oy) - o.y) * inv.y;
    const float az = (fmaf((float)((near.z >> sh) & 0xffu), sz, oz) - o.z) * inv.z;
    const float bz = (fmaf((float)((far.z >> sh) & 0xffu), sz, oz) - o.z) * inv.z;
    const float t0 = fmaxf(fmaxf(ax, ay), az);
    const float t1 = fminf(fminf(bx, by), bz) * k;
    t = t0;
    return t0 <= t1 && t0 < t_max && t1 > 0.f;
}

// One visit of a four-child node: tests the four dequantized child boxes,
// sorts the entered children by entry distance, pushes all but the nearest
// (farthest first) and returns the nearest; with none entered, returns the
// top of the stack or kTraceDone.
template <bool COUNT, int BLOCK>
__device__ __forceinline__ int node_step4(const GpuNode4Q *__restrict__ nodes, int cur, f3 o, f3 inv, float t_max,
                                          TraceStack<BLOCK> &st, uint32_t &n_box) {
    const float inf = __builtin_inff();
    const uint4 *np = reinterpret_cast<const uint4 *>(nodes + cur);
    const uint4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3];
    if (COUNT) n_box += 4;
    const float ox = __uint_as_float(w0.x), oy = __uint_as_float(w0.y), oz = __uint_as_float(w0.z);
    const float sx = __uint_as_float((w0.w & 0xffu) << 23), sy = __uint_as_float(((w0.w >> 8) & 0xffu) << 23),
                sz = __uint_as_float(((w0.w >> 16) & 0xffu) << 23);
    const int4 ch = make_int4((int)w2.z, (int)w2.w, (int)w3.x, (int)w3.y);
    // the near and far plane bytes of each axis, chosen once per visit by the
    // direction's sign (BVHAccel's dir_is_neg) instead of a min and a max per
    // axis and child
    const bool nx = __float_as_uint(inv.x) >> 31, ny = __float_as_uint(inv.y) >> 31, nz = __float_as_uint(inv.z) >> 31;
    const uint4 near = make_uint4(nx ? w1.y : w1.x, ny ? w1.w : w1.z, nz ? w2.y : w2.x, 0u);
    const uint4 far = make_uint4(nx ? w1.x : w1.y, ny ? w1.z : w1.w, nz ? w2.x : w2.y, 0u);
    float t0, t1, t2, t3;
    const bool h0 = qslab_nf<0>(near, far, ox, oy, oz, sx, sy, sz, o, inv, t_max, t0) && ch.x != kEmptyChild;
    const bool h1 = qslab_nf<1>(near, far, ox, oy, oz, sx, sy, sz, o, inv, t_max, t1) && ch.y != kEmptyChild;
    const bool h2 = qslab_nf<2>(near, far, ox, oy, oz, sx, sy, sz, o, inv, t_max, t2) && ch.z != kEmptyChild;
    const bool h3 = qslab_nf<3>(near, far, ox, oy, oz, sx, sy, sz, o, inv, t_max, t3) && ch.w != kEmptyChild;
    float ta = h0 ? t0 : inf, tb = h1 ? t1 : inf, tc = h2 ? t2 : inf, td = h3 ? t3 : inf;
    int ra = ch.x, rb = ch.y, rc = ch.z, rd = ch.w;
    const int n_hit = (int)h0 + (int)h1 + (int)h2 + (int)h3;
    if (n_hit == 0) return st.sp == 0 ? kTraceDone : st.pop();
    // sorting network for four keys; misses (inf) sink to the end
    cswap(ta, ra, tb, rb);
    cswap(tc, rc, td, rd);
    cswap(ta, ra, tc, rc);
    cswap(tb, rb, td, rd);
    cswap(tb, rb, tc, rc);
    if (n_hit > 3) st.push(rd);
    if (n_hit > 2) st.push(rc);
    if (n_hit > 1) st.push(rb);
    return ra;
}

// Traversal of the 4-wide tree: each visit reads one 64-byte quantized node
// (node_step4). A stack of kMaxTraceDepth entries holds 3 per level (the
// builder caps the depth).
//
// The loop body is one step: a lane either visits an inner node or tests one
// primitive of its current leaf, so a wave pays for at most one node visit and
// one primitive test per iteration whatever mix of the two its lanes are in.
// (Nesting a leaf loop inside the node loop makes every pop a wave-wide
// rendezvous that runs the whole leaf loop whenever any lane holds a leaf; on
// incoherent rays that left a sixth of the lanes active.) A nearer hit only
// moves t_max and the record index; the rest of the hit record is filled in
// once, after the loop. An answered any-hit query leaves through the loop's
// one exit: a second exit from inside the primitive branch made the
// shadow-ray loops of the wavefront shading kernel twenty times slower for the
// same node and primitive tests.
template <bool ANY, bool COUNT, int BLOCK>
__device__ __forceinline__ bool trace_ray4(const GpuNode4Q *__restrict__ nodes, const PrimRecord *__restrict__ prims,
                                           f3 o, f3 d, float t_max, TraceHit &hit, TraceStack<BLOCK> &st,
                                           uint32_t &n_box, uint32_t &n_prim) {
    const f3 inv = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
    const Shear sh = make_shear(d);
    hit.prim = 0xFFFFFFFFu;
    hit.slot = 0;
    hit.t = hit.b0 = hit.b1 = hit.b2 = 0.f;
    uint32_t slot = kNoSlot;
    int cur = 0;              // next node: inner node (>= 0), leaf (< 0) or kTraceDone
    uint32_t pf = 0, pn = 0;  // primitives left in the current leaf
    st.sp = 0;
    while (true) {
        if (pn != 0) {
            if (COUNT) ++n_prim;
            float t;
            if (prim_step<ANY>(prims, pf, o, d, sh, t_max, t)) {
                slot = pf;
                t_max = t;
                if (ANY) {
                    // answered: leave through the loop's one exit
                    pn = 1;
                    cur = kTraceDone;
                }
            }
            ++pf;
            --pn;
        } else {
            cur = node_step4<COUNT, BLOCK>(nodes, cur, o, inv, t_max, st, n_box);
        }
        take_leaf(cur, pf, pn, st);
        if (pn == 0 && cur == kTraceDone) break;
    }
    if (slot == kNoSlot) return false;
    hit.slot = slot;
    if (!ANY) finish_hit(prims, slot, o, sh, t_max, hit);
    return true;
}

}  // namespace illume
