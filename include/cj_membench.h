/*
 * cj_membench.h -- memory microbenchmarks of libcalamarijoins_amd.so, the
 * calibration behind the roofline figures (no reference counterpart: the
 * reference publishes no hardware measurements).
 *
 *   cj_membench_copy:   a streaming 16-byte-per-lane device copy of `bytes`
 *                       (read + write); the per-byte scale of the profiler's
 *                       FETCH_SIZE / WRITE_SIZE counters comes from it.
 *   cj_membench_gather: `n` dependent-free loads of `record_bytes` (4, 8 or 16)
 *                       at table + idx[i] * record_bytes from a table of
 *                       `table_bytes`, 8 loads in flight per lane, one 4-byte
 *                       result per lane written at the end; idx is host memory.
 *   cj_membench_gather16_mode: the gather of 16-byte records (table_bytes a
 *                       table) with mode 0 plain loads, 1 non-temporal loads,
 *                       2 the lookups grouped by table slice (eighths) and
 *                       workgroup b taking slice b % 8, so that each XCD's L2
 *                       serves one eighth of the table.
 *
 * All allocate their buffers, run `reps` times and report the best time in
 * milliseconds (hipEvents around each launch).  Return CJ_OK or a negative
 * status with cj_last_error().
 */
#ifndef CJ_MEMBENCH_H
#define CJ_MEMBENCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int cj_membench_copy(int32_t device, uint64_t bytes, uint32_t reps, double *best_ms);
int cj_membench_gather(int32_t device, const uint32_t *idx, uint64_t n, uint64_t table_bytes, uint32_t record_bytes,
                       uint32_t reps, double *best_ms);
int cj_membench_gather16_mode(int32_t device, const uint32_t *idx, uint64_t n, uint64_t table_bytes, uint32_t mode,
                              uint32_t reps, double *best_ms);

#ifdef __cplusplus
}
#endif
#endif
