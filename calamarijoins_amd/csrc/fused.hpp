// fused.hpp -- arguments of the fused probe + materialization kernel.
#pragma once
#include "cj_internal.hpp"

namespace cj {

constexpr uint32_t kFusedBatch = 1984;  // output rows per batch = one INT32 page
constexpr uint32_t kFusedHalf = 992;    // rows of each of the two INT64/FP64 pages of a batch
constexpr uint32_t kFusedMaxCols = 8;

struct FusedCol {
    int32_t from_build;  // 1: build side (dense decoded), 0: probe side (pages)
    int32_t width;       // 4 or 8
    int32_t is_int32;    // INT32_MIN -> NULL
    // build side
    const void *values;
    const uint32_t *valid_bits;
    const uint32_t *rids;  // intermediate build side: base row ids (nullptr = identity)
    // probe side
    const uint8_t *pages;
    uint32_t num_pages;
    const uint64_t *row_start;   // [num_pages + 1]
    const uint64_t *byte_start;  // [num_pages + 1] first bitmap byte of each page in byte_prefix
    const uint16_t *byte_prefix;
    // output
    uint8_t *out;  // output page block of this column
};

struct FusedArgs {
    KeySource key;  // probe key column (pages)
    HashTableDev ht;
    const NLKeys *nl;
    const uint64_t *off_rows;  // [units + 1]
    const uint64_t *off_p32;   // [units + 1] first INT32 page of each workgroup
    const uint64_t *off_p64;   // [units + 1] first INT64/FP64 page of each workgroup
    uint32_t ncols;
    FusedCol cols[kFusedMaxCols];
};

void page_byte_prefix(const uint8_t *pages, uint32_t num_pages, const uint64_t *byte_start, uint16_t *prefix,
                      hipStream_t s);
void fused_batch_pages(const uint64_t *cnt, uint32_t n, uint64_t *p32, uint64_t *p64, hipStream_t s);
void fused_write(const FusedArgs &a, uint32_t units, bool nl, hipStream_t s);

}  // namespace cj
