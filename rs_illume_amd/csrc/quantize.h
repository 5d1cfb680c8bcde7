// Quantization of one four-child node (GpuNode4 -> GpuNode4Q), shared by the
// host path (quantize_width4) and the device collapse of the device builder,
// which therefore write the same bytes for the same node.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include "layout.h"

namespace illume {

// Plane of byte q: origin + q * step in float32 (the product is exact, the
// sum rounds once); the traversal evaluates planes with the same fused
// multiply-add.
__host__ __device__ inline float q_plane(float origin, float step, int q) { return fmaf((float)q, step, origin); }
__host__ __device__ inline float pow2f(int e) { return ldexpf(1.f, e); }

// Surface area used when the collapse picks the child to open.
__host__ __device__ inline float collapse_area(const float *b) {
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return 2.f * (dx * dy + dy * dz + dz * dx);
}

// Per axis: origin = the smallest lower plane of the used children, step =
// the smallest power of two whose byte 255 reaches the largest upper plane;
// each lower plane takes the largest byte whose plane lies on or below it,
// each upper plane the smallest byte whose plane lies on or above it, so the
// dequantized box contains the float box. False if a used child's box is not
// finite or cannot be covered.
__host__ __device__ inline bool quantize_node(const GpuNode4 &g, GpuNode4Q *out) {
    GpuNode4Q q;
    q.ox = q.oy = q.oz = 0.f;
    q.ex = q.ey = q.ez = q.pad0 = 0;
    q.pad1[0] = q.pad1[1] = 0;
    const float *lo[3] = {g.lo_x, g.lo_y, g.lo_z}, *hi[3] = {g.hi_x, g.hi_y, g.hi_z};
    uint8_t *qlo[3] = {q.lo_x, q.lo_y, q.lo_z}, *qhi[3] = {q.hi_x, q.hi_y, q.hi_z};
    float *origin[3] = {&q.ox, &q.oy, &q.oz};
    uint8_t *expo[3] = {&q.ex, &q.ey, &q.ez};
    bool used[4];
    for (int c = 0; c < 4; ++c) {
        used[c] = g.child[c] != kEmptyChild;
        q.child[c] = g.child[c];
    }
    for (int a = 0; a < 3; ++a) {
        float mn = INFINITY, mx = -INFINITY;
        for (int c = 0; c < 4; ++c)
            if (used[c]) {
                if (!isfinite(lo[a][c]) || !isfinite(hi[a][c])) return false;
                mn = fminf(mn, lo[a][c]);
                mx = fmaxf(mx, hi[a][c]);
            }
        if (!(mn <= mx)) mn = mx = 0.f;  // no used child: any valid encoding
        *origin[a] = mn;
        // the smallest exponent that reaches mx (the plane is monotone in e):
        // start from the real-arithmetic guess and step either way
        int e = -126;
        const double ext = (double)mx - (double)mn;
        if (ext > 0) e = (int)ceil(log2(ext / 255.0));
        e = e < -126 ? -126 : (e > 127 ? 127 : e);
        while (e < 127 && q_plane(mn, pow2f(e), 255) < mx) ++e;
        while (e > -126 && q_plane(mn, pow2f(e - 1), 255) >= mx) --e;
        if (q_plane(mn, pow2f(e), 255) < mx) return false;
        *expo[a] = (uint8_t)(e + 127);
        const float step = pow2f(e);
        for (int c = 0; c < 4; ++c) {
            if (!used[c]) {
                qlo[a][c] = 255;
                qhi[a][c] = 0;
                continue;
            }
            // largest q with plane(q) <= lo, smallest q with plane(q) >= hi
            int ql = (int)floor(((double)lo[a][c] - (double)mn) / (double)step);
            ql = ql < 0 ? 0 : (ql > 255 ? 255 : ql);
            while (ql > 0 && q_plane(mn, step, ql) > lo[a][c]) --ql;
            while (ql < 255 && q_plane(mn, step, ql + 1) <= lo[a][c]) ++ql;
            int qh = (int)ceil(((double)hi[a][c] - (double)mn) / (double)step);
            qh = qh < 0 ? 0 : (qh > 255 ? 255 : qh);
            while (qh < 255 && q_plane(mn, step, qh) < hi[a][c]) ++qh;
            while (qh > 0 && q_plane(mn, step, qh - 1) >= hi[a][c]) --qh;
            if (q_plane(mn, step, ql) > lo[a][c] || q_plane(mn, step, qh) < hi[a][c]) return false;
            qlo[a][c] = (uint8_t)ql;
            qhi[a][c] = (uint8_t)qh;
        }
    }
    *out = q;
    return true;
}

}  // namespace illume
