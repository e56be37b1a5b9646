/*
 * cj_oracle.h -- CPU oracle for the join pipeline (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark load it.  It
 * restates, in plain C, the reference's algorithm for
 * Contest::execute(const Plan&, void*) (reference src/execute.cpp:207-325)
 * and the reference's canonical page codec (src/build_table.cpp:520-924).
 *
 * Plans and tables use the structs of include/cj_api.h so that the same
 * descriptors drive the oracle and the HIP engine.
 */
#ifndef CJ_ORACLE_H
#define CJ_ORACLE_H

#include "../include/cj_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- hashing primitives (reference join_execution/hashtable.h:137-169,
 *      join_execution/bloom_tags.h:35-263) ------------------------------- */
uint32_t cjo_crc32c_u32(uint32_t crc, uint32_t v);   /* table-driven CRC-32C */
uint32_t cjo_crc32c_bytes(uint32_t crc, const uint8_t *p, size_t n);
uint16_t cjo_bloom_tag(uint32_t idx);                 /* idx in [0, 2048)     */

/* ---- canonical page codec ------------------------------------------------
 * Encoders restate Table::to_columnar (build_table.cpp:679-924); they write
 * pages back to back into `out` (capacity `max_pages` pages) and return the
 * number of pages, or (size_t)-1 if `max_pages` is too small.
 *   fixed: values has n_rows entries of 4 (INT32) or 8 (INT64/FP64) bytes;
 *          NULL rows (valid[i]==0) ignore their value slot.
 *   varchar: offsets[n_rows+1] into bytes. */
size_t cjo_encode_fixed(int32_t type, uint64_t n_rows, const void *values,
                        const uint8_t *valid, uint8_t *out, size_t max_pages);
size_t cjo_encode_varchar(uint64_t n_rows, const uint64_t *offsets,
                          const char *bytes, const uint8_t *valid, uint8_t *out,
                          size_t max_pages);

/* Decoders restate Table::from_columnar (build_table.cpp:520-659).  They
 * return the number of rows decoded or -1 on a malformed column.  Pass
 * NULL outputs to only count rows / string bytes. */
int64_t cjo_decode_fixed(int32_t type, const void *const *pages,
                         uint32_t num_pages, void *values, uint8_t *valid);
int64_t cjo_decode_varchar(const void *const *pages, uint32_t num_pages,
                           uint64_t *offsets, char *bytes, uint64_t bytes_cap,
                           uint8_t *valid, uint64_t *total_bytes);

/* Per-row 64-bit hash of a whole table (all columns, NULL-aware, raw bytes
 * for VARCHAR, raw bits for FP64); used for order-free multiset compares. */
int cjo_row_hashes(const cj_table_desc *table, uint64_t *out);

/* ---- execute() restatement ------------------------------------------------
 * Runs the plan on the CPU with `threads` worker threads (0 = all cores).
 * The result pages are malloc'd by the oracle; release with
 * cjo_result_free.  Returns 0 or a negative status (cjo_last_error). */
typedef struct cjo_result {
    uint64_t num_rows;
    uint32_t num_columns;
    int32_t *types;
    uint32_t *num_pages;
    void ***pages;
} cjo_result;

int cjo_execute(const cj_plan_desc *plan, int threads, cjo_result *out);
void cjo_result_free(cjo_result *res);
const char *cjo_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
