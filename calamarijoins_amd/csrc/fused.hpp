// fused.hpp -- the root-join fast path for fixed-width outputs whose probe
// side is a base table: probe counting and pair writing over key pages, and
// multi-column output page writers reading the build side from its dense
// columns or direct-map records and the probe side straight from its pages.
#pragma once
#include "cj_internal.hpp"
#include "direct.hpp"

namespace cj {

// How the build side is looked up.
enum ProbeMode : int { kModeHash = 0, kModeNL = 1, kModeDirect = 2 };

// The probe side's key pages and the build structure they are looked up in.
struct ProbeArgs {
    KeySource key;  // columnar probe key column
    int32_t mode;   // ProbeMode
    HashTableDev ht;
    const NLKeys *nl;
    DirectTable dt;
};

// Matches per work unit of the probe side (key page, or kInterChunk rows of
// an intermediate), 8 rows per thread.
void probe_count_columnar(const ProbeArgs &a, uint64_t *counts, hipStream_t s);
// Writes, in probe-row order, the (build id, probe row) of every match at
// off[unit] onwards; build id = record slot in direct mode (the build row id
// when the records carry it), build row id otherwise.  Either output may be
// null.
void probe_write_columnar(const ProbeArgs &a, const uint64_t *off, uint32_t *build_ids, uint32_t *probe_rows,
                          hipStream_t s);

constexpr uint32_t kMatMaxCols = 8;

struct MatCol {
    int32_t width;     // 4 or 8
    int32_t is_int32;  // INT32_MIN -> NULL
    // build side: dense decoded column (+ row-id chain) or direct-map record field
    const void *values;
    const uint32_t *valid_bits;
    const uint32_t *rids;
    uint32_t rec_offset, rec_bit;
    // probe side: pages + page index + per output page first window page
    const uint8_t *pages;
    uint32_t num_pages;
    const uint64_t *row_start;   // [num_pages + 1]
    const uint64_t *byte_start;  // [num_pages + 1]
    const uint16_t *byte_prefix;
    const uint32_t *win_first;   // [output pages]
    uint8_t *out;                // output page block of this column
};

// Output columns of one side and one width.  Result rows are written in
// blocks of kBlockRows: block k fills INT32 page k and 8-byte pages 2k and
// 2k + 1.
struct MatArgs {
    int32_t from_build;
    int32_t direct;
    DirectTable dt;
    const uint32_t *ids;  // per output row: build id or probe row
    uint64_t rows;
    uint32_t ncols;
    uint32_t cols_per_wg;  // columns one workgroup writes (1..ncols)
    MatCol cols[kMatMaxCols];
};

constexpr uint32_t kBlockRows = kRowsI32;      // 1984: a full INT32 page
constexpr uint32_t kRowsPage8 = kBlockRows / 2;  // 992 rows per 8-byte page

void materialize_side(const MatArgs &a, hipStream_t s);

// Page index of a fixed-width column: per bitmap byte, the non-NULL rows of
// its page before it.
void page_byte_prefix(const uint8_t *pages, uint32_t num_pages, const uint64_t *byte_start, uint16_t *prefix,
                      hipStream_t s);
// For each block of R result rows, the payload page holding its first probe
// row: first[k] = upper_bound(col_row_start, probe_rows[k * R]) - 1.
void window_first(const uint32_t *probe_rows, uint64_t rows, uint32_t rows_per_block, const uint64_t *col_row_start,
                  uint32_t col_pages, uint32_t *first, hipStream_t s);

}  // namespace cj
