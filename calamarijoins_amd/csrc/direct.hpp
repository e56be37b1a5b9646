// direct.hpp -- direct-mapped build side (unique, dense INT32 keys).
#pragma once
#include "cj_internal.hpp"

namespace cj {

constexpr uint32_t kDirectMaxFields = 8;
constexpr uint32_t kDirectPresentBit = 31;  // flags bit: the slot holds a build row
constexpr uint32_t kDirectParts = 4096;     // workgroup partials of each build pass
constexpr uint32_t kMinmaxParts = 1024;     // workgroup partials of key_minmax

struct DirectField {
    int32_t width;  // 4 or 8
    uint32_t offset;  // byte offset inside the record
    const void *values;  // dense decoded build column
    const uint32_t *valid_bits;
    const uint8_t *pages;  // values == nullptr: an INT32 column read in place (dense_page_value)
    uint32_t dense_rows;
};

struct DirectBuildArgs {
    const int32_t *keys;  // dense decoded build key column
    const uint32_t *key_valid;
    const uint8_t *key_pages;  // keys == nullptr: the key column read in place (dense_page_value)
    uint32_t key_rows;
    uint64_t rows;
    const int32_t *minmax;  // device [min, max] of the non-NULL keys (key_minmax)
    uint64_t max_range;     // capacity of presence / records, in slots
    uint32_t *presence;     // max_range / 32 + 2 words
    uint8_t *records;       // max_range * stride bytes
    uint32_t stride;        // multiple of 16
    uint32_t flags_offset;  // u32: bit f = field f non-NULL, bit 31 = present
    uint32_t *counters;     // [0] build rows with a key, [1] slots present
    uint32_t *partials;     // 2 * kDirectParts scratch words
    uint32_t nfields;
    DirectField fields[kDirectMaxFields];
    int32_t rid_offset;  // >= 0: the build row id is stored at this record offset; with no
                         // fields the records are the row ids alone (stride 4, offset 0)
    // Statistics of an earlier build over the same key column.  known != 0:
    // the key range is [lo, hi] (minmax is not read, key_minmax is not run)
    // and the keys are unique.  full != 0 as well (every key of the range
    // present): the build is the scatter alone, since it writes every slot
    // of the range and the probes skip the bitmap.
    uint32_t known, full;
    int32_t lo, hi;
};

// The table as the probe sees it.
struct DirectTable {
    int32_t min_key;
    uint64_t range;
    const uint32_t *presence;
    const uint8_t *records;
    uint32_t stride;
    uint32_t flags_offset;
    uint32_t full;  // every key of the range is present: no presence lookup
    int32_t rid_offset;  // >= 0: record offset of the build row id (joins whose output lists row ids)
};

// mm[0] = min, mm[1] = max of the non-NULL keys (INT32_MAX, INT32_MIN when
// there is none); `part` holds 2 * kMinmaxParts scratch ints.
void key_minmax(const int32_t *keys, const uint32_t *valid, const uint8_t *key_pages, uint32_t key_rows, uint64_t n,
                int32_t *mm, int32_t *part, hipStream_t s);
// Zero the records of the key range, scatter one record per build row, then
// derive the presence bitmap from the records (for a map of row ids alone,
// from the slots a row id filled).  Nothing is written when the
// key range exceeds max_range.  Keys are unique iff counters[0] == counters[1].
// With known statistics the counters are not written, and a full range is
// neither zeroed nor given a bitmap.
void direct_build(const DirectBuildArgs &a, hipStream_t s);

}  // namespace cj
