// pagemat.hpp -- page-aligned root materialization for joins in which every
// probe row matches at most one build row (a direct-mapped build side).
//
// A result column may choose its own page boundaries: the join result is a
// multiset of rows, and row i of the result is row i of every column.  Both
// sides here list the matched probe rows in ascending order, so
//   * each probe-side output page is one payload page of the probe table with
//     its unmatched rows removed (read once, coalesced, never searched);
//   * each key page of the probe side yields the build-side output page(s) of
//     its own matched rows, straight from the lookup that found them.
// No (build id, probe row) lists are written or read.
#pragma once
#include "fused.hpp"

namespace cj {

// Counting pass that also sets bit r of `match_bits` (zeroed by the caller)
// for every probe row r with a match.  dense_rows: the key column's pages
// hold no NULLs and at most this many rows each (0: unknown).
void probe_count_match(const ProbeArgs &a, uint64_t *counts, uint32_t *match_bits, uint32_t dense_rows,
                       hipStream_t s);

// Pair lists from the match bits of probe_count_match (direct mode, key
// pages): the matched rows of key page pg, in row order, as (build id, probe
// row) at off[pg] onwards; the build id is the build row id the direct map's
// records carry (dt.rid_offset >= 0), else the record slot.  Either output
// may be null.
void direct_pairs(const ProbeArgs &a, const uint32_t *match_bits, const uint64_t *off, uint32_t *ob, uint32_t *op,
                  hipStream_t s);
// The same over a row-id list (KEY_INTER probe side, direct map): match bit
// r = row r of the intermediate; counts per kInterChunk-row unit.
void inter_count_match(const ProbeArgs &a, uint64_t *counts, uint32_t *match_bits, hipStream_t s);
void inter_pairs(const ProbeArgs &a, const uint32_t *match_bits, const uint64_t *off, uint32_t *ob, uint32_t *op,
                 hipStream_t s);

// Output pages each key page produces for the build side:
// p4[pg] = ceil(m / 1984), p8[pg] = ceil(m / 1007) for its m matches.
void key_page_out_pages(const uint64_t *counts, uint32_t pages, uint64_t *p4, uint64_t *p8, hipStream_t s);

constexpr uint32_t kCap4 = kRowsI32;  // 1984 rows: a full INT32 page
constexpr uint32_t kCap8 = kRowsI64;  // 1007 rows: a full 8-byte page

struct BuildPageCol {
    int32_t width;     // 4 or 8
    int32_t is_int32;  // INT32_MIN -> NULL
    uint32_t rec_offset, rec_bit;
    uint8_t *out;      // page block of this column
    uint64_t cap_pages;  // pages the block holds: a key page whose pages would pass it writes nothing
};

struct BuildPagesArgs {
    KeySource key;  // columnar probe key column
    DirectTable dt;
    const uint32_t *match_bits;  // from probe_count_match
    const uint64_t *p4_off;  // first output page per key page (exclusive scans)
    const uint64_t *p8_off;
    uint32_t ncols;
    BuildPageCol cols[kMatMaxCols];
};
// Build-side output pages, one workgroup per key page.  dt.stride must be a
// multiple of 16 (at most 80).
void build_pages(const BuildPagesArgs &a, hipStream_t s);

// nonempty[q] = 1 if payload page q holds a matched row, else 0.
void page_match_flags(const uint8_t *pages, uint32_t num_pages, const uint64_t *row_start,
                      const uint32_t *match_bits, uint64_t *nonempty, hipStream_t s);

struct FilterArgs {
    const uint8_t *pages;  // payload column pages
    uint32_t num_pages;
    const uint64_t *row_start;  // [num_pages + 1]
    const uint32_t *match_bits;
    // output page of each payload page (exclusive scan of nonempty); null:
    // payload page q writes output page q and a page without a matched row
    // is left unwritten and counted in *empty_pages (see compact_pages)
    const uint64_t *pos;
    uint64_t *empty_pages;
    int32_t width;
    int32_t is_int32;
    uint32_t max_page_rows;  // largest n_rows of a page of the column
    uint8_t *out;
};
// Probe-side output pages, one per payload page with a match.
void probe_filter(const FilterArgs &a, hipStream_t s);
// dst page pos[q] = src page q for every page q with pos[q + 1] > pos[q]
// (closes the gaps the empty pages of a fixed-slot probe_filter left).
void compact_pages(const uint8_t *src, uint32_t num_pages, const uint64_t *pos, uint8_t *dst, hipStream_t s);

}  // namespace cj

namespace cj {
// dst[i] = *l.p[i]: collects scattered device counters for one read-back.
struct U64List {
    const uint64_t *p[16];
    uint32_t n;
};
void gather_u64s(const U64List &l, uint64_t *dst, hipStream_t s);
}  // namespace cj
