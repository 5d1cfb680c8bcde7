// One shape against a batch of rays: Shape::intersect / Shape::intersect_p
// (core/shape.rs:8-112) of a single Triangle (shapes/triangle.rs:142-607) or
// full Sphere (shapes/sphere.rs:105-323), with the same device tests the
// traversal runs on a leaf's primitive record. For parity tests of the shapes
// alone; rendering reaches them through the aggregate.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <string>
#include "../../include/illume_hip.h"
#include "trace.h"

#pragma clang fp contract(off)

namespace illume {
int abi_fail(int status, const std::string &msg);
int abi_hip_fail(hipError_t e, const char *what);

namespace {

// rec: the shape as the leaf record of primitive 0 (triangle: the three
// vertices; sphere: centre and radius with the sphere tag). Closest hit
// rejects degenerate triangles (Triangle::intersect), any hit does not
// (Triangle::intersect_p); a hit writes prim 0, t and the barycentrics of
// vertices 1 and 2.
template <bool ANY>
__global__ __launch_bounds__(256) void shape_kernel(PrimRecord rec, const illume_ray *__restrict__ rays, uint32_t n,
                                                    illume_hit *__restrict__ hits) {
    __shared__ PrimRecord s_rec;
    if (threadIdx.x == 0) s_rec = rec;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 *rp = reinterpret_cast<const float4 *>(rays + i);
    const float4 r0 = rp[0], r1 = rp[1];
    const f3 o = mk(r0.x, r0.y, r0.z), d = mk(r0.w, r1.x, r1.y);
    const Shear sh = make_shear(d);
    float t;
    TraceHit h;
    h.prim = ILLUME_MISS;
    h.t = h.b0 = h.b1 = h.b2 = 0.f;
    if (prim_step<ANY>(&s_rec, 0, o, d, sh, r1.z, t)) {
        finish_hit(&s_rec, 0, o, sh, t, h);
        h.prim = 0;
    }
    reinterpret_cast<float4 *>(hits)[i] = make_float4(__uint_as_float(h.prim), h.t, h.b1, h.b2);
}

}  // namespace
}  // namespace illume

namespace illume {
namespace {
// Triangle t's vertices (tri_verts[3 t + k], w = 0) from the primitive
// records, which hold them in mesh index order with t as the first record's
// id; spheres' records are skipped.
__global__ __launch_bounds__(256) void tri_verts_kernel(const PrimRecord *__restrict__ prims, uint32_t n_records,
                                                        uint32_t n_tris, float4 *__restrict__ tv) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_records) return;
    const float4 *pp = reinterpret_cast<const float4 *>(prims + s);
    const float4 a = pp[0], b = pp[1], c = pp[2];
    const uint32_t id = __float_as_uint(a.w);
    if ((id & kSphereBit) || id >= n_tris) return;
    tv[3 * (size_t)id] = make_float4(a.x, a.y, a.z, 0.f);
    tv[3 * (size_t)id + 1] = make_float4(b.x, b.y, b.z, 0.f);
    tv[3 * (size_t)id + 2] = make_float4(c.x, c.y, c.z, 0.f);
}
}  // namespace

// A scene's per-triangle vertex table built on the device from its BVH's
// primitive records (no host copy or upload of the 48 bytes per triangle);
// entries of triangles without a record stay zero. Queued on st.
hipError_t scene_tri_verts(const PrimRecord *prims, uint32_t n_records, uint32_t n_tris, float4 *tv, hipStream_t st) {
    hipError_t e = hipMemsetAsync(tv, 0, 3 * (size_t)n_tris * sizeof(float4), st);
    if (e != hipSuccess || n_records == 0) return e;
    tri_verts_kernel<<<(n_records + 255) / 256, 256, 0, st>>>(prims, n_records, n_tris, tv);
    return hipGetLastError();
}
}  // namespace illume

extern "C" int illume_shape_intersect_batch(int32_t kind, const float *shape, const illume_ray *rays, uint32_t n,
                                            int32_t any, illume_hit *hits, void *stream) {
    using namespace illume;
    if (kind != ILLUME_SHAPE_TRIANGLE && kind != ILLUME_SHAPE_SPHERE)
        return abi_fail(ILLUME_ERR_INVALID_ARG, "kind must be ILLUME_SHAPE_TRIANGLE or ILLUME_SHAPE_SPHERE");
    if (!shape) return abi_fail(ILLUME_ERR_INVALID_ARG, "shape is null");
    if (n > (1u << 31)) return abi_fail(ILLUME_ERR_UNSUPPORTED, "at most 2^31 rays per call");
    if (n == 0) return ILLUME_OK;
    if (!rays || !hits) return abi_fail(ILLUME_ERR_INVALID_ARG, "rays or hits is null");
    PrimRecord rec{};
    if (kind == ILLUME_SHAPE_TRIANGLE) {
        for (int k = 0; k < 9; ++k)
            if (!std::isfinite(shape[k])) return abi_fail(ILLUME_ERR_INVALID_ARG, "non-finite triangle vertex");
        float *dst[3] = {rec.a, rec.b, rec.c};
        for (int v = 0; v < 3; ++v)
            for (int c = 0; c < 3; ++c) dst[v][c] = shape[3 * v + c];
        rec.a[3] = 0.f;  // primitive id 0
    } else {
        for (int k = 0; k < 4; ++k)
            if (!std::isfinite(shape[k])) return abi_fail(ILLUME_ERR_INVALID_ARG, "non-finite sphere");
        if (!(shape[3] > 0.f)) return abi_fail(ILLUME_ERR_INVALID_ARG, "sphere radius must be positive");
        for (int c = 0; c < 3; ++c) rec.a[c] = shape[c];
        uint32_t tag = kSphereBit;  // primitive id 0 with the sphere tag
        std::memcpy(&rec.a[3], &tag, 4);
        rec.b[0] = shape[3];
    }
    const uint32_t grid = (n + 255) / 256;
    if (any)
        shape_kernel<true><<<grid, 256, 0, (hipStream_t)stream>>>(rec, rays, n, hits);
    else
        shape_kernel<false><<<grid, 256, 0, (hipStream_t)stream>>>(rec, rays, n, hits);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? ILLUME_OK : abi_hip_fail(e, "shape intersect launch");
}
