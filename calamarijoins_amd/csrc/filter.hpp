// filter.hpp -- row filters on the GPU (cj_filter): a postfix program of
// comparisons and AND / OR / NOT evaluated per row over decoded columns,
// then the passing row ids in order.
#pragma once
#include "../../include/cj_api.h"
#include "cj_internal.hpp"

namespace cj {

constexpr uint32_t kFilterMaxNodes = 64;  // program length
constexpr uint32_t kFilterMaxDepth = 32;  // evaluation stack (bits of a register)

struct FNode {
    int32_t kind;      // CJ_F_*
    int32_t op;        // CJ_OP_*
    uint32_t col;      // index into the FCol array
    int32_t lit_kind;  // CJ_LIT_*
    int64_t ival;
    double fval;
    uint32_t soff, slen;  // string literal in the literal buffer
};

struct FCol {
    int32_t type;
    const void *values;      // INT32 / INT64 / FP64: dense decoded values
    const uint32_t *valid;   // validity bitmap by row
    const uint8_t *pages;    // VARCHAR: the column's pages
    const uint64_t *refs;    // VARCHAR: decoded refs (page, offset, length)
};

// sel: filter_sel_words(rows) words, written whole (no clearing needed)
inline uint64_t filter_sel_words(uint64_t rows) { return (rows + 255) / 256 * 8; }
void filter_eval(const FNode *prog, uint32_t n, const FCol *cols, const uint8_t *lits, uint64_t rows,
                 uint32_t *sel, hipStream_t s);
// rows set in sel -> their ids in ascending order; counts/offs: chunks + 1
constexpr uint32_t kSelChunk = 2048;
void sel_count(const uint32_t *sel, uint64_t rows, uint64_t *counts, hipStream_t s);
void sel_rids(const uint32_t *sel, uint64_t rows, const uint64_t *offs, uint32_t *rids, hipStream_t s);

}  // namespace cj
