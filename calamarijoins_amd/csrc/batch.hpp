// batch.hpp -- batched execution of many independent single-join plans.
//
// A radix join over all plans at once.  Both sides of every plan are
// partitioned by the top bits of the key hash; a build partition covers
// 2^log_part directory slots (2^8 .. 2^12, tuned), so its tuples (a few
// thousand at most) fit in LDS.  The output columns' values travel with the
// keys through the partition pass ("carried" columns), so no phase gathers
// by row id: the probe side streams through HBM and every lookup is an LDS
// access.  Phases:
//
//   decode of the carried columns not read in place -> partition histogram
//   and scatter of both sides -> per-partition build (slot order,
//   partition-local ends) -> join units (probe chunks of a partition) ->
//   match count per unit -> [read-back of per-plan result sizes] -> page
//   init -> match write (values into their 8 KiB page slots) -> page fix
//   (pages with NULLs packed).
//
// Every phase is one launch of a persistent kernel that walks a work list
// covering all plans: global unit g (a key page, a partition, a probe
// chunk, an output page) maps to (plan, local unit) through a prefix array.
// Workgroups take units strided by the grid; the join phases may instead
// pull them from a device work queue (one atomic counter per phase).  The
// phase sequence before the read-back is a hipGraph captured once per batch
// template and replayed.  The autotuner picks the hash load factor, the
// radix bits (slots per partition), the join tile (probe tuples per unit),
// the join scheduling and the grid (each kernel's resident grid, or a fixed
// workgroup count for every launch); cj_batch_config_set pins all of them.
#pragma once
#include "cj_internal.hpp"

namespace cj {

// Work queue counters of the join phases (used when the configuration
// schedules join units dynamically).
enum { kQJoinCount = 0, kQJoinWrite, kBatchQueues };

constexpr uint32_t kBatchMaxCarry = 4;                // carried columns per side
constexpr uint32_t kBatchMaxOut = 8;                  // output columns per plan
constexpr uint32_t kBatchMaxLogPart = 12;             // slots per partition: 2^8 .. 2^12
constexpr uint64_t kBatchMaxBuildRows = 1ull << 22;   // larger builds run one plan at a time
constexpr uint32_t kBatchMaxChunk = 16384;            // probe tuples per join unit
constexpr uint32_t kBatchLdsBytes = 64 * 1024;        // LDS of a join workgroup
constexpr uint32_t kBatchSortCap = 16384;             // partitions of at most this many build tuples are
                                                      // ordered by (slot, key) (partition-local indices fit 16 bits)

// One side of one plan.  Carried column j holds, in partition order, the
// value of the source column for each tuple: INT32 with NULL folded to
// INT32_MIN (the output writer turns INT32_MIN into NULL either way), 8-byte
// values with a validity byte beside them.
struct BSide {
    const uint8_t *k_pages;  // key column pages
    const uint64_t *k_rs;    // their row starts
    uint32_t ncar;
    uint32_t car_w[kBatchMaxCarry];             // 4 or 8
    const uint8_t *src[kBatchMaxCarry];         // dense decoded source column (by row), or its pages
    const uint32_t *src_valid[kBatchMaxCarry];  // its validity bitmap (by row)
    uint32_t src_rows[kBatchMaxCarry];          // > 0: INT32 without NULLs read in place from its pages,
                                                // row r = value r % src_rows of page r / src_rows
    uint8_t *car[kBatchMaxCarry];               // carried values, plan-local tuple index
    uint8_t *car_vld[kBatchMaxCarry];           // validity bytes (8-byte columns only)
};

struct BPlan {
    BSide b, p;
    uint32_t log_slots;  // directory = 2^log_slots slots, partitions of 2^log_part
    uint64_t dir_base;   // first directory entry of the plan
    // build side in slot order (written by the build phase)
    uint8_t *bslot_car[kBatchMaxCarry];
    uint8_t *bslot_vld[kBatchMaxCarry];
};

struct BatchArgs {
    const BPlan *plans;
    uint32_t nplans;
    const uint32_t *bpage_start;  // [nplans + 1] prefix of build key pages
    const uint32_t *ppage_start;  // [nplans + 1] prefix of probe key pages
    // the same prefixes counting only the pages of plans with at most / more
    // than kStagedMaxParts partitions (the two scatter kernels' work lists)
    const uint32_t *bpage_start_s, *bpage_start_d;
    const uint32_t *ppage_start_s, *ppage_start_d;
    const uint32_t *part_start;   // [nplans + 1] prefix of partitions
    uint32_t max_parts;           // largest partition count of a plan (LDS histogram)
    uint32_t staged_parts;        // largest partition count of a plan scattered through LDS staging
    uint32_t direct_parts;        // largest partition count of a plan scattered directly
    uint32_t nparts;
    // build side
    uint64_t *bcount, *bcur;      // [nparts + 1]
    const uint64_t *boff;         // exclusive scan of bcount
    int32_t *bstage_keys;         // partition order
    int32_t *bkeys;               // slot order
    uint32_t *dir;                // partition-local end of each slot
    // probe side
    uint64_t *pcount, *pcur;
    const uint64_t *poff;
    int32_t *pkeys;               // partition order
    // join units
    uint64_t *ucount;             // [nparts + 1] units per partition
    const uint64_t *uoff;         // exclusive scan (uoff[nparts] = units)
    uint32_t max_units;
    uint64_t *mcount;             // [max_units + 1] matches per unit
    const uint64_t *moff;         // exclusive scan
    uint32_t lds_cap;             // build tuples a join workgroup holds in LDS
    uint32_t tuple_bytes;         // LDS bytes of a staged build tuple (key + carried values)
    uint32_t log_part;            // slots per partition = 2^log_part
    uint32_t chunk;               // probe tuples per join unit
    uint32_t join_queue;          // 1: join units from the work queue, 0: strided by workgroup
    uint32_t *queue;              // [kBatchQueues] join work queue counters (zeroed before each join launch)
};

// Output columns of one plan (known after the read-back).  Row o of output
// column c goes straight to its page: page o / R, slot o % R of the value
// area (R = 1984 INT32 or 1007 8-byte rows a page; the column's last page is
// shorter).  The page-init pass wrote every page's header and an all-ones
// validity bitmap before; a NULL clears its bit and counts on its page, and
// the page-fix pass packs the values of the pages that hold NULLs.
struct BOut {
    uint32_t nout;
    uint32_t rows;              // result rows of the plan (< 2^32)
    uint8_t src[kBatchMaxOut];  // 0 key, 1 build carried, 2 probe carried
    uint8_t j[kBatchMaxOut];    // carried column index
    uint8_t w[kBatchMaxOut];    // 4 or 8
    uint8_t *out[kBatchMaxOut];  // the column's pages
    uint32_t pg0[kBatchMaxOut];  // global index of its first page (NULL counters)
};

// Dense decode of a base column (values + validity bitmap).
struct BDecode {
    const uint8_t *pages;
    const uint64_t *rs;
    int32_t width;
    void *values;
    uint32_t *valid;  // zeroed
};

// One output column of one plan: its pages.
struct BPageCol {
    uint8_t *out;
    uint64_t rows;
    int32_t width;
};

// Workgroups of each persistent kernel that the device holds at once
// (occupancy at the kernel's registers and LDS x compute units): a launch
// of that many takes its units in one round.  A larger grid runs in rounds
// and, with units split evenly by stride, leaves the machine part-idle in
// the last one.
struct BatchGrids {
    uint32_t decode, part_count, part_scatter, part_scatter_direct, build, build_big, join_count, join_write, page_init,
        page_fix;
};
// Also checks that every kernel's LDS fits a workgroup (UnsupportedError if
// not) and raises the dynamic-LDS limit where a kernel needs more than 64 KiB;
// called when a batch template is made, outside graph capture.
BatchGrids batch_resident_grids(const BatchArgs &a, int device);

// Each launcher takes the persistent grid size (workgroups).
void batch_part_count(const BatchArgs &a, bool probe, uint32_t units, uint32_t grid, hipStream_t s);
// Plans with at most kStagedMaxParts partitions scatter through LDS staging
// (a page's tuples ordered by partition, written out in runs); plans with
// more scatter each tuple directly.  pstart / units: the kernel's work list.
constexpr uint32_t kStagedMaxParts = 256;
void batch_part_scatter(const BatchArgs &a, bool probe, const uint32_t *pstart, uint32_t units, bool staged,
                        uint32_t grid, hipStream_t s);
void batch_build(const BatchArgs &a, uint32_t grid, uint32_t grid_big, hipStream_t s);
void batch_units(const BatchArgs &a, hipStream_t s);
void batch_join_count(const BatchArgs &a, uint32_t grid, hipStream_t s);
void batch_join_write(const BatchArgs &a, const BOut *outs, uint32_t *page_nulls, uint32_t grid, hipStream_t s);
void batch_decode(const BDecode *jobs, const uint32_t *page_start, uint32_t njobs, uint32_t units, uint32_t grid,
                  hipStream_t s);
// Output pages: headers, all-ones bitmaps and zero NULL counters of all
// `pages` pages (page_start: prefix over the columns), then, after the
// write pass, the values of the pages that counted NULLs packed.
void batch_page_init(const BPageCol *cols, const uint32_t *page_start, uint32_t ncols, uint32_t pages,
                     uint32_t *page_nulls, uint32_t grid, hipStream_t s);
void batch_page_fix(const BPageCol *cols, const uint32_t *page_start, uint32_t ncols, uint32_t pages,
                    const uint32_t *page_nulls, uint32_t grid, hipStream_t s);
// totals[k] = matches of plan k's units
void batch_totals(const BatchArgs &a, uint64_t *totals, hipStream_t s);
void batch_copy_u64(const uint64_t *src, uint64_t *dst, uint64_t n, hipStream_t s);

}  // namespace cj
