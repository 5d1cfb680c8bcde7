// C ABI entry points (include/illume_hip.h). Status codes, never exceptions,
// cross the boundary; the message of the last failure is kept per host thread.
#include <hip/hip_runtime.h>
#include <chrono>
#include <cstring>
#include <new>
#include <string>
#include "../../include/illume_hip.h"
#include "bvh_builder.h"
#include "kernels.h"
#include "lbvh.h"

namespace {
thread_local std::string g_last_error;

int fail(int status, const std::string &msg) {
    g_last_error = msg;
    return status;
}
int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? ILLUME_ERR_OUT_OF_MEMORY : ILLUME_ERR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

struct illume_bvh {
    void *d_nodes = nullptr;  // GpuNode[] (width 2) or GpuNode4Q[] (width 4)
    illume::PrimRecord *d_prims = nullptr;
    int width = 2;
    illume_bvh_info_t info;
    int device = 0;
};

namespace illume {
int abi_fail(int status, const std::string &msg) { return fail(status, msg); }
int abi_hip_fail(hipError_t e, const char *what) { return hip_fail(e, what); }
void bvh_device_arrays(const illume_bvh *bvh, const GpuNode4Q **nodes, const PrimRecord **prims) {
    *nodes = bvh->width == 4 ? (const GpuNode4Q *)bvh->d_nodes : nullptr;
    *prims = bvh->d_prims;
}
}  // namespace illume

extern "C" {

const char *illume_status_string(int status) {
    switch (status) {
        case ILLUME_OK: return "ok";
        case ILLUME_ERR_INVALID_ARG: return "invalid argument";
        case ILLUME_ERR_HIP: return "HIP runtime error";
        case ILLUME_ERR_OUT_OF_MEMORY: return "out of device memory";
        case ILLUME_ERR_UNSUPPORTED: return "unsupported parameter";
        case ILLUME_ERR_QUEUE_OVERFLOW: return "work queue overflow";
        default: return "unknown status";
    }
}

const char *illume_last_error(void) { return g_last_error.c_str(); }

int illume_bvh_create_ex(const float *verts, uint32_t n_verts, const uint32_t *tri_idx, uint32_t n_tris,
                         const float *spheres, uint32_t n_spheres, int max_prims_in_node, int width,
                         illume_bvh **out) {
    if (!out) return fail(ILLUME_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    if (max_prims_in_node < 1 || max_prims_in_node > illume::kMaxLeafPrims)
        return fail(ILLUME_ERR_UNSUPPORTED, "max_prims_in_node must be in 1..15");
    if (width != 2 && width != 4) return fail(ILLUME_ERR_UNSUPPORTED, "width must be 2 or 4");
    auto t0 = std::chrono::steady_clock::now();
    illume::BuildInput in;
    in.verts = verts;
    in.n_verts = n_verts;
    in.idx = tri_idx;
    in.n_tris = n_tris;
    in.spheres = spheres;
    in.n_spheres = n_spheres;
    in.max_prims_in_node = max_prims_in_node;
    in.width = width;
    illume::BuildOutput bo;
    std::string err;
    try {
        err = illume::build_bvh(in, &bo);
    } catch (const std::bad_alloc &) {
        return fail(ILLUME_ERR_OUT_OF_MEMORY, "host allocation failed during BVH build");
    }
    if (!err.empty()) return fail(ILLUME_ERR_INVALID_ARG, err);
    // stack entries: up to width-1 pending siblings per level
    if ((int)bo.max_depth * (width - 1) >= illume::max_supported_depth())
        return fail(ILLUME_ERR_UNSUPPORTED, "BVH deeper than the traversal stack");
    auto t1 = std::chrono::steady_clock::now();

    illume_bvh *h = new (std::nothrow) illume_bvh();
    if (!h) return fail(ILLUME_ERR_OUT_OF_MEMORY, "host allocation failed");
    std::memset(&h->info, 0, sizeof(h->info));
    h->width = width;
    hipGetDevice(&h->device);
    std::vector<illume::GpuNode4Q> nq;
    if (width == 4 && !illume::quantize_width4(bo.nodes4, &nq)) {
        delete h;
        return fail(ILLUME_ERR_INVALID_ARG, "node boxes cannot be quantized (non-finite bounds)");
    }
    const void *nsrc = width == 4 ? (const void *)nq.data() : (const void *)bo.nodes.data();
    size_t n_nodes = width == 4 ? nq.size() : bo.nodes.size();
    size_t nb = width == 4 ? n_nodes * sizeof(illume::GpuNode4Q) : n_nodes * sizeof(illume::GpuNode);
    size_t pb = bo.prims.size() * sizeof(illume::PrimRecord);
    hipError_t e;
    if (nb) {
        if ((e = hipMalloc(&h->d_nodes, nb)) != hipSuccess) {
            delete h;
            return hip_fail(e, "hipMalloc(nodes)");
        }
        if ((e = hipMemcpy(h->d_nodes, nsrc, nb, hipMemcpyHostToDevice)) != hipSuccess) {
            illume_bvh_destroy(h);
            return hip_fail(e, "hipMemcpy(nodes)");
        }
    }
    if (pb) {
        if ((e = hipMalloc(&h->d_prims, pb)) != hipSuccess) {
            illume_bvh_destroy(h);
            return hip_fail(e, "hipMalloc(prims)");
        }
        if ((e = hipMemcpy(h->d_prims, bo.prims.data(), pb, hipMemcpyHostToDevice)) != hipSuccess) {
            illume_bvh_destroy(h);
            return hip_fail(e, "hipMemcpy(prims)");
        }
    }
    h->info.n_prims = (uint64_t)n_tris + n_spheres;
    h->info.n_inner_nodes = n_nodes;
    h->info.n_leaves = bo.n_leaves;
    h->info.max_depth = bo.max_depth;
    h->info.device_bytes = nb + pb;
    h->info.build_seconds = std::chrono::duration<double>(t1 - t0).count();
    std::memcpy(h->info.bounds, bo.bounds, sizeof(bo.bounds));
    *out = h;
    return ILLUME_OK;
}

int illume_bvh_create_gpu(const float *verts, uint32_t n_verts, const uint32_t *tri_idx, uint32_t n_tris,
                          const float *spheres, uint32_t n_spheres, int max_prims_in_node, int width, void *stream,
                          illume_bvh **out) {
    if (!out) return fail(ILLUME_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    if (max_prims_in_node < 1 || max_prims_in_node > illume::kMaxLeafPrims)
        return fail(ILLUME_ERR_UNSUPPORTED, "max_prims_in_node must be in 1..15");
    if (width != 2 && width != 4) return fail(ILLUME_ERR_UNSUPPORTED, "width must be 2 or 4");
    const uint64_t n = (uint64_t)n_tris + n_spheres;
    // a handful of primitives: the host build is as fast and handles a root leaf
    if (n < 2 * (uint64_t)max_prims_in_node)
        return illume_bvh_create_ex(verts, n_verts, tri_idx, n_tris, spheres, n_spheres, max_prims_in_node, width,
                                    out);
    auto t0 = std::chrono::steady_clock::now();
    std::string err = illume::validate_prims(verts, n_verts, tri_idx, n_tris, spheres, n_spheres);
    if (!err.empty()) return fail(ILLUME_ERR_INVALID_ARG, err);
    illume::LbvhResult r;
    hipError_t e = illume::lbvh_build(verts, n_verts, tri_idx, n_tris, spheres, n_spheres, max_prims_in_node,
                                      (hipStream_t)stream, &r);
    if (e != hipSuccess) return hip_fail(e, "device BVH build");
    illume_bvh *h = new (std::nothrow) illume_bvh();
    if (!h) {
        hipFree(r.nodes);
        hipFree(r.prims);
        return fail(ILLUME_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    std::memset(&h->info, 0, sizeof(h->info));
    h->width = width;
    hipGetDevice(&h->device);
    h->d_prims = r.prims;
    size_t nb = (size_t)r.n_nodes * sizeof(illume::GpuNode);
    // the levels above the Morton treelets rebuilt by SAH on the host
    std::vector<illume::GpuNode> n2(r.n_nodes);
    e = hipMemcpy(n2.data(), r.nodes, nb, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        hipFree(r.nodes);
        illume_bvh_destroy(h);
        return hip_fail(e, "hipMemcpy(binary nodes)");
    }
    uint32_t depth = illume::sah_top(n2, r.upper);
    // stack entries: up to width-1 pending siblings per level of the traced tree
    if (width == 2 && (int)depth >= illume::max_supported_depth()) {
        hipFree(r.nodes);
        illume_bvh_destroy(h);
        return fail(ILLUME_ERR_UNSUPPORTED, "BVH deeper than the traversal stack");
    }
    if (width == 2) {
        h->d_nodes = r.nodes;
        e = hipMemcpy(h->d_nodes, n2.data(), nb, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            illume_bvh_destroy(h);
            return hip_fail(e, "hipMemcpy(binary nodes)");
        }
    } else {
        std::vector<illume::GpuNode4> n4;
        hipFree(r.nodes);
        depth = illume::collapse_to_width4(n2, &n4);
        if ((int)depth * 3 >= illume::max_supported_depth()) {
            illume_bvh_destroy(h);
            return fail(ILLUME_ERR_UNSUPPORTED, "BVH deeper than the traversal stack");
        }
        std::vector<illume::GpuNode4Q> nq;
        if (!illume::quantize_width4(n4, &nq)) {
            illume_bvh_destroy(h);
            return fail(ILLUME_ERR_INVALID_ARG, "node boxes cannot be quantized (non-finite bounds)");
        }
        nb = nq.size() * sizeof(illume::GpuNode4Q);
        if ((e = hipMalloc(&h->d_nodes, nb)) == hipSuccess)
            e = hipMemcpy(h->d_nodes, nq.data(), nb, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            illume_bvh_destroy(h);
            return hip_fail(e, "four-child nodes");
        }
        r.n_nodes = (uint32_t)n4.size();
    }
    auto t1 = std::chrono::steady_clock::now();
    h->info.n_prims = n;
    h->info.n_inner_nodes = r.n_nodes;
    h->info.n_leaves = r.n_leaves;
    h->info.max_depth = depth;
    h->info.device_bytes = nb + n * sizeof(illume::PrimRecord);
    h->info.build_seconds = std::chrono::duration<double>(t1 - t0).count();
    std::memcpy(h->info.bounds, r.bounds, sizeof(r.bounds));
    *out = h;
    return ILLUME_OK;
}

int illume_bvh_create_gpu_sah(const float *verts, uint32_t n_verts, const uint32_t *tri_idx, uint32_t n_tris,
                              const float *spheres, uint32_t n_spheres, int max_prims_in_node, int width,
                              void *stream, illume_bvh **out) {
    if (!out) return fail(ILLUME_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    if (max_prims_in_node < 1 || max_prims_in_node > illume::kMaxLeafPrims)
        return fail(ILLUME_ERR_UNSUPPORTED, "max_prims_in_node must be in 1..15");
    if (width != 2 && width != 4) return fail(ILLUME_ERR_UNSUPPORTED, "width must be 2 or 4");
    const uint64_t n = (uint64_t)n_tris + n_spheres;
    if (n >= illume::kMaxPrimSlots) return fail(ILLUME_ERR_INVALID_ARG, "too many primitives (limit 2^27)");
    // a handful of primitives: the host build (the same tree) handles a root leaf
    if (n < 2 * (uint64_t)max_prims_in_node)
        return illume_bvh_create_ex(verts, n_verts, tri_idx, n_tris, spheres, n_spheres, max_prims_in_node, width,
                                    out);
    auto t0 = std::chrono::steady_clock::now();
    std::string err = illume::validate_prims(verts, n_verts, tri_idx, n_tris, spheres, n_spheres);
    if (!err.empty()) return fail(ILLUME_ERR_INVALID_ARG, err);
    illume::SahGpuResult r;
    hipError_t e = illume::sah_gpu_build(verts, n_verts, tri_idx, n_tris, spheres, n_spheres, max_prims_in_node,
                                         width, (hipStream_t)stream, &r);
    if (e == hipErrorInvalidValue) return fail(ILLUME_ERR_INVALID_ARG, "node boxes cannot be quantized (non-finite bounds)");
    if (e != hipSuccess) return hip_fail(e, "device SAH build");
    void *nodes = width == 4 ? (void *)r.nodes4 : (void *)r.nodes;
    // stack entries: up to width-1 pending siblings per level of the traced tree
    if ((int)r.depth * (width - 1) >= illume::max_supported_depth()) {
        (void)hipFree(nodes);
        (void)hipFree(r.prims);
        return fail(ILLUME_ERR_UNSUPPORTED, "BVH deeper than the traversal stack");
    }
    illume_bvh *h = new (std::nothrow) illume_bvh();
    if (!h) {
        (void)hipFree(nodes);
        (void)hipFree(r.prims);
        return fail(ILLUME_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    std::memset(&h->info, 0, sizeof(h->info));
    h->width = width;
    hipGetDevice(&h->device);
    h->d_nodes = nodes;
    h->d_prims = r.prims;
    const size_t nb = (size_t)r.n_nodes * (width == 4 ? sizeof(illume::GpuNode4Q) : sizeof(illume::GpuNode));
    auto t1 = std::chrono::steady_clock::now();
    h->info.n_prims = n;
    h->info.n_inner_nodes = r.n_nodes;
    h->info.n_leaves = r.n_leaves;
    h->info.max_depth = r.depth;
    h->info.device_bytes = nb + n * sizeof(illume::PrimRecord);
    h->info.build_seconds = std::chrono::duration<double>(t1 - t0).count();
    std::memcpy(h->info.bounds, r.bounds, sizeof(r.bounds));
    *out = h;
    return ILLUME_OK;
}

int illume_bvh_create(const float *verts, uint32_t n_verts, const uint32_t *tri_idx, uint32_t n_tris,
                      const float *spheres, uint32_t n_spheres, int max_prims_in_node, illume_bvh **out) {
    return illume_bvh_create_ex(verts, n_verts, tri_idx, n_tris, spheres, n_spheres, max_prims_in_node, 2, out);
}

int illume_bvh_destroy(illume_bvh *bvh) {
    if (!bvh) return ILLUME_OK;
    if (bvh->d_nodes) hipFree(bvh->d_nodes);
    if (bvh->d_prims) hipFree(bvh->d_prims);
    delete bvh;
    return ILLUME_OK;
}

int illume_bvh_info(const illume_bvh *bvh, illume_bvh_info_t *out) {
    if (!bvh || !out) return fail(ILLUME_ERR_INVALID_ARG, "null argument");
    *out = bvh->info;
    return ILLUME_OK;
}

static int check_ray_count(uint32_t n) {
    if (n > illume::kMaxRaysPerCall)
        return fail(ILLUME_ERR_UNSUPPORTED, "at most 2^31 rays per traversal call (split the batch)");
    return ILLUME_OK;
}

static int run_traverse(const illume_bvh *bvh, const illume_ray *rays, uint32_t n, illume_hit *hits,
                        uint8_t *occluded, uint32_t *counts, bool any, void *stream) {
    if (int st = check_ray_count(n)) return st;
    if (!bvh) return fail(ILLUME_ERR_INVALID_ARG, "bvh is null");
    if (n == 0) return ILLUME_OK;
    if (!rays) return fail(ILLUME_ERR_INVALID_ARG, "rays is null");
    if (!hits && !occluded && !counts) return fail(ILLUME_ERR_INVALID_ARG, "output is null");
    hipError_t e = illume::launch_traverse(bvh->d_nodes, bvh->width, bvh->d_prims, bvh->info.n_prims == 0, rays, n, hits,
                                           occluded, counts, any, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "traversal launch");
    return ILLUME_OK;
}

int illume_intersect_closest(const illume_bvh *bvh, const illume_ray *rays, uint32_t n, illume_hit *hits,
                             void *stream) {
    if (int st = check_ray_count(n)) return st;
    if (n && !hits) return fail(ILLUME_ERR_INVALID_ARG, "hits is null");
    return run_traverse(bvh, rays, n, hits, nullptr, nullptr, false, stream);
}

int illume_intersect_closest_noalloc(const illume_bvh *bvh, const illume_ray *rays, uint32_t n, illume_hit *hits,
                                     void *stream) {
    if (int st = check_ray_count(n)) return st;
    if (!bvh) return fail(ILLUME_ERR_INVALID_ARG, "bvh is null");
    if (n == 0) return ILLUME_OK;
    if (!rays) return fail(ILLUME_ERR_INVALID_ARG, "rays is null");
    if (!hits) return fail(ILLUME_ERR_INVALID_ARG, "hits is null");
    hipError_t e = illume::launch_closest_noalloc(bvh->d_nodes, bvh->width, bvh->d_prims, bvh->info.n_prims == 0,
                                                  rays, n, hits, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "traversal launch");
    return ILLUME_OK;
}

int illume_intersect_any(const illume_bvh *bvh, const illume_ray *rays, uint32_t n, uint8_t *occluded,
                         void *stream) {
    if (int st = check_ray_count(n)) return st;
    if (n && !occluded) return fail(ILLUME_ERR_INVALID_ARG, "occluded is null");
    return run_traverse(bvh, rays, n, nullptr, occluded, nullptr, true, stream);
}

int illume_intersect_counts(const illume_bvh *bvh, const illume_ray *rays, uint32_t n, int any, uint32_t *counts,
                            void *stream) {
    if (int st = check_ray_count(n)) return st;
    if (n && !counts) return fail(ILLUME_ERR_INVALID_ARG, "counts is null");
    return run_traverse(bvh, rays, n, nullptr, nullptr, counts, any != 0, stream);
}

}  // extern "C"
