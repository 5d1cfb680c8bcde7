// Counter calibration probe: random 64-byte line reads issued the way the
// traversal fetches nodes (each lane reads its own line with float4 loads).
// Reading a known number of bytes lets hardware read counters (FETCH_SIZE)
// be converted to bytes for that access pattern.
#include <hip/hip_runtime.h>
#include <string>
#include "../../include/illume_hip.h"
#include "../../include/illume_probe.h"

namespace illume {
int abi_fail(int status, const std::string &msg);
int abi_hip_fail(hipError_t e, const char *what);
}  // namespace illume

namespace {

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// lane t reads lines mix32(t * 7919 + s * 104729) % n_lines, s < steps; every
// index is reduced modulo n_lines, so every access is inside the buffer.
__global__ __launch_bounds__(256) void probe_kernel(const float4 *__restrict__ buf, uint32_t n_lines, uint32_t n,
                                                    uint32_t steps, float *__restrict__ out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float acc = 0.f;
    for (uint32_t s = 0; s < steps; ++s) {
        const float4 *p = buf + 4 * (size_t)(mix32(t * 7919u + s * 104729u) % n_lines);
        const float4 a = p[0], b = p[1], c = p[2], d = p[3];
        acc += a.x + b.y + c.z + d.w;
    }
    out[t] = acc;
}

}  // namespace

extern "C" int illume_probe_random_reads(const void *buf, uint64_t buf_bytes, uint32_t n_lanes, uint32_t steps,
                                         float *out, void *stream) {
    if (!buf || !out) return illume::abi_fail(ILLUME_ERR_INVALID_ARG, "null argument");
    const uint64_t n_lines = buf_bytes / 64;
    if (n_lines == 0 || n_lines > 0xFFFFFFFFull || ((uintptr_t)buf & 15))
        return illume::abi_fail(ILLUME_ERR_INVALID_ARG, "buffer must hold 1..2^32-1 64-byte lines, 16-byte aligned");
    if (n_lanes == 0) return ILLUME_OK;
    probe_kernel<<<(n_lanes + 255) / 256, 256, 0, (hipStream_t)stream>>>((const float4 *)buf, (uint32_t)n_lines,
                                                                         n_lanes, steps, out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? ILLUME_OK : illume::abi_hip_fail(e, "probe launch");
}
