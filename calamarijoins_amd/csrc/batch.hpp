// batch.hpp -- batched execution of many independent single-join plans.
//
// Every phase of the join (radix partition histogram and scatter, LDS hash
// build, probe count, pair write, column decode, page materialization) is
// one launch of a persistent kernel that walks a device work list covering
// all plans: global unit g (a key page, a partition, an output page) maps to
// (plan, local unit) through a per-plan prefix array.  Per-plan geometry
// lives in a device descriptor array, and all per-plan buffers are slices of
// global arrays, so one global scan yields every plan's offsets.  The two
// phase sequences (before and after the read-back of the per-plan result
// sizes) are captured into hipGraphs once per batch template and replayed.
#pragma once
#include "cj_internal.hpp"

namespace cj {

// One plan as the batch kernels see it.  The build side is already chosen.
struct BPlan {
    const uint8_t *bk_pages;  // build key column
    const uint64_t *bk_rs;    // its row starts
    const uint8_t *pk_pages;  // probe key column
    const uint64_t *pk_rs;
    uint32_t log_slots;       // directory = 2^log_slots entries
    uint32_t log_part;        // slots per partition = 2^log_part
    uint64_t dir_base;        // first directory entry of the plan
};

struct BatchArgs {
    const BPlan *plans;
    uint32_t nplans;
    const uint32_t *bpage_start;  // [nplans + 1] prefix of build key pages
    const uint32_t *ppage_start;  // [nplans + 1] prefix of probe key pages
    const uint32_t *part_start;   // [nplans + 1] prefix of partitions
    uint32_t max_parts;           // largest partition count of a plan (LDS histogram)
    uint64_t *part_count;         // [nparts + 1]
    uint64_t *part_cur;           // scatter cursors (copy of the partition offsets)
    const uint64_t *part_off;     // [nparts + 1] exclusive scan of part_count
    int32_t *stage_keys;          // partitioned (key, row) tuples, all plans
    uint32_t *stage_rids;
    uint64_t *dir;                // directories, all plans
    int32_t *keys;                // tuples in slot order, all plans
    uint32_t *rids;
    uint64_t *pcount;             // [nppages + 1] matches per probe key page
    const uint64_t *poff;         // [nppages + 1] exclusive scan of pcount
    uint32_t *out_b, *out_p;      // (build row, probe row) pairs, all plans
};

// Dense decode of a base column (values + validity bitmap).
struct BDecode {
    const uint8_t *pages;
    const uint64_t *rs;
    int32_t width;
    void *values;
    uint32_t *valid;  // zeroed
};

// One output column of one plan.
struct BMat {
    const uint32_t *ids;  // the plan's build or probe rows, one per result row
    uint64_t rows;
    const void *values;   // dense decoded source column
    const uint32_t *valid;
    int32_t width;
    int32_t is_int32;
    uint8_t *out;         // page block
};

constexpr uint32_t kBatchMaxLogPart = 12;             // partitions of <= 4096 slots (LDS build)
constexpr uint64_t kBatchMaxBuildRows = 1ull << 22;   // larger builds run one plan at a time

// Each launcher takes the persistent grid size (workgroups).
void batch_part_count(const BatchArgs &a, uint32_t units, uint32_t grid, hipStream_t s);
void batch_part_scatter(const BatchArgs &a, uint32_t units, uint32_t grid, hipStream_t s);
void batch_build(const BatchArgs &a, uint32_t nparts, uint32_t grid, hipStream_t s);
void batch_probe_count(const BatchArgs &a, uint32_t units, uint32_t grid, hipStream_t s);
void batch_probe_write(const BatchArgs &a, uint32_t units, uint32_t grid, hipStream_t s);
void batch_decode(const BDecode *jobs, const uint32_t *page_start, uint32_t njobs, uint32_t units, uint32_t grid,
                  hipStream_t s);
void batch_materialize(const BMat *cols, const uint32_t *page_start, uint32_t ncols, uint32_t units, uint32_t grid,
                       hipStream_t s);
// totals[k] = poff[ppage_start[k + 1]] - poff[ppage_start[k]]
void batch_totals(const uint64_t *poff, const uint32_t *ppage_start, uint32_t n, uint64_t *totals, hipStream_t s);
void batch_copy_u64(const uint64_t *src, uint64_t *dst, uint64_t n, hipStream_t s);

}  // namespace cj
