/*
 * illume_probe.h — measurement-only entry points of librs_illume_amd.so.
 *
 * Not part of the drop-in boundary (include/illume_hip.h) and with no
 * counterpart in the reference: probes used to calibrate hardware counters
 * and to measure the memory system's ceiling for the traversal's access
 * pattern. Status codes as in illume_hip.h.
 */
#ifndef ILLUME_PROBE_H
#define ILLUME_PROBE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* n_lanes lanes each read `steps` random 64-byte lines of buf (device,
 * buf_bytes bytes, 16-byte aligned) with four float4 loads per line, the way
 * the traversal fetches nodes; out: device [n_lanes] floats (sums, so the
 * loads are kept). Bytes read = n_lanes * steps * 64. */
int illume_probe_random_reads(const void *buf, uint64_t buf_bytes, uint32_t n_lanes, uint32_t steps, float *out,
                              void *stream);

#ifdef __cplusplus
}
#endif

#endif /* ILLUME_PROBE_H */
