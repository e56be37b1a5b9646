// direct.hpp -- direct-mapped build side (unique, dense INT32 keys).
#pragma once
#include "cj_internal.hpp"

namespace cj {

constexpr uint32_t kDirectMaxFields = 8;

struct DirectField {
    int32_t width;  // 4 or 8
    uint32_t offset;  // byte offset inside the record
    const void *values;  // dense decoded build column
    const uint32_t *valid_bits;
};

struct DirectBuildArgs {
    const int32_t *keys;  // dense decoded build key column
    const uint32_t *key_valid;
    uint64_t rows;
    int32_t min_key;
    uint32_t *presence;   // ceil(range / 32) words, zeroed
    uint8_t *records;     // range * stride bytes
    uint32_t stride;
    uint32_t flags_offset;  // u32 of per-field non-NULL bits
    uint32_t *dup_flag;
    uint32_t nfields;
    DirectField fields[kDirectMaxFields];
};

// The table as the probe sees it.
struct DirectTable {
    int32_t min_key;
    uint64_t range;
    const uint32_t *presence;
    const uint8_t *records;
    uint32_t stride;
    uint32_t flags_offset;
};

void key_minmax(const int32_t *keys, const uint32_t *valid, uint64_t n, int32_t *mm, hipStream_t s);
void direct_build(const DirectBuildArgs &a, hipStream_t s);

}  // namespace cj
