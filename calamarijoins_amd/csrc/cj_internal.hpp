// cj_internal.hpp -- declarations shared by the HIP kernels and the host
// executor of the MI355X join engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace cj {

constexpr uint32_t kPage = 8192;
constexpr int kWG = 256;          // threads per workgroup (4 wave64)
constexpr int32_t kI32Null = INT32_MIN;

// Output page row capacity for fixed-width columns: the largest row count
// whose all-non-NULL page fits 8192 bytes (4 + 4*1984 + 248 = 8188 for
// INT32, 8 + 8*1007 + 126 = 8190 for INT64/FP64).  Same capacities the
// canonical encoder reaches on dense data (reference build_table.cpp:706-716).
constexpr uint32_t kRowsI32 = 1984;
constexpr uint32_t kRowsI64 = 1007;

// Packed VARCHAR reference to a string inside a resident page:
//   bits 63..32 page index, 31..16 byte offset of the chars inside the
//   page, 15..0 length.  Long strings (0xffff first page): byte offset
//   0xffff and the low 16 bits hold the number of pages of the run.
constexpr uint32_t kLongMarker = 0xffff;

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct UnsupportedError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct InvalidError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct OomError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct BusyError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// The calling thread's message for cj_last_error() (api.cpp).
void set_last_error(const std::string &msg);

#define CJ_HIP(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            throw ::cj::DeviceError(std::string(#expr " failed: ") + hipGetErrorString(e_)); \
    } while (0)

// ---- hash: 64-bit mix of the INT32 key; the directory slot is the top
//      log2(slots) bits, the Bloom tag index the low 11 bits. -------------
__host__ __device__ inline uint64_t hash_key(int32_t key) {
    uint64_t h = (uint64_t)(uint32_t)key * 0x9E3779B97F4A7C15ull;
    h ^= h >> 32;
    h *= 0xD6E8FEB86659FD93ull;
    h ^= h >> 32;
    return h;
}

// Hash table geometry (unchained table with Bloom tags in the directory).
struct HashGeom {
    int log_slots;   // directory = 2^log_slots entries of u64 (end<<16 | tag)
    int log_part;    // slots per radix partition = 2^log_part
    int log_sub;     // slots a build workgroup counts in LDS at a time (<= log_part)
    uint32_t num_partitions;
};

struct HashTableDev {
    const uint64_t *dir;
    const int32_t *keys;
    const uint32_t *rids;
    int log_slots;
    uint64_t size;
};

// How a join side's key is read.
enum KeyMode : int {
    KEY_COLUMNAR = 0,  // key pages of a base column; NULL by bitmap
    KEY_INTER = 1      // gathered through a row-id list; NULL or INT32_MIN skip
};

// ---- launch wrappers (defined in the .hip files) -------------------------
// scan.hip
void scan_exclusive_u64(const uint64_t *in, uint64_t *out, uint64_t n, hipStream_t s,
                        uint64_t *tmp, size_t tmp_elems);
size_t scan_tmp_elems(uint64_t n);
// Up to kMultiScanJobs exclusive scans of u64 arrays in two launches for
// all of them: out[i] = sum of f(in[j]) for j < i, out[n] = the total, with
// f(x) = x, or ceil(x / div) when div > 0 (pages of div rows for x rows;
// x below 2^32).  sums: multi_scan_blocks(jobs, njobs) words of scratch.
constexpr uint32_t kMultiScanJobs = 8;
constexpr uint64_t kMultiScanMax = uint64_t(1) << 20;  // longest array it takes
struct ScanJob {
    const uint64_t *in;
    uint64_t *out;
    uint64_t n;
    uint64_t div;
};
void multi_scan_u64(const ScanJob *jobs, uint32_t njobs, uint64_t *sums, hipStream_t s);
uint32_t multi_scan_blocks(const ScanJob *jobs, uint32_t njobs);
void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s);
void fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s);
void page_row_counts(const uint8_t *pages, uint32_t num_pages, uint64_t *counts, hipStream_t s);
void decode_fixed(const uint8_t *pages, uint32_t num_pages, const uint64_t *row_start, int width,
                  void *values, uint32_t *valid_bits, hipStream_t s);
void decode_varchar(const uint8_t *pages, uint32_t num_pages, const uint64_t *row_start,
                    uint64_t *refs, uint32_t *valid_bits, hipStream_t s);
void gather_u32(const uint32_t *src, const uint32_t *idx, uint32_t *dst, uint64_t n, hipStream_t s);
constexpr uint32_t kGatherSet = 8;
struct GatherSet {  // dst[k][i] = src[k][idx[i]] for k < count
    const uint32_t *src[kGatherSet];
    uint32_t *dst[kGatherSet];
    uint32_t count;
};
void gather_u32_set(const GatherSet &g, const uint32_t *idx, uint64_t n, hipStream_t s);

// join.hip
struct KeySource {
    int mode;                    // KeyMode
    // columnar
    const uint8_t *pages;
    uint32_t num_pages;
    const uint64_t *row_start;   // [num_pages + 1]
    // intermediate
    const int32_t *values;       // decoded dense key column of the base table
    const uint32_t *valid_bits;
    const uint32_t *rids;        // row-id list of the intermediate (nullptr = identity)
    uint64_t rows;               // rows of the intermediate
    // values == nullptr: the base column is read in place from its pages
    // (INT32, no NULLs, dense_rows rows on every page but the last)
    const uint8_t *dense_pages;
    uint32_t dense_rows;
};

// Value of row `rid` of an INT32 column without NULLs whose pages all hold
// `rows_per_page` rows (the last may hold fewer): value rid % R of page
// rid / R, at its aligned offset.
__device__ __forceinline__ int32_t dense_page_value(const uint8_t *pages, uint32_t rows_per_page, uint32_t rid) {
    const uint32_t pg = rid / rows_per_page;
    const uint32_t i = rid - pg * rows_per_page;
    return *reinterpret_cast<const int32_t *>(pages + (size_t)pg * kPage + 4 + 4 * (size_t)i);
}
constexpr uint32_t kInterChunk = 2048;  // rows of a row-id list per work unit
// work units (pages / row chunks)
__host__ __device__ inline uint32_t key_source_units(const KeySource &k) {
    if (k.mode == KEY_COLUMNAR) return k.num_pages;
    return (uint32_t)((k.rows + kInterChunk - 1) / kInterChunk);
}

void build_partition_hist(const KeySource &k, const HashGeom &g, uint64_t *part_count, hipStream_t s);
void build_partition_scatter(const KeySource &k, const HashGeom &g, uint64_t *part_cursor,
                             int32_t *tup_keys, uint32_t *tup_rids, hipStream_t s);
void build_partitions(const HashGeom &g, const uint64_t *part_off, const int32_t *tup_keys,
                      const uint32_t *tup_rids, uint64_t *dir, int32_t *keys, uint32_t *rids,
                      hipStream_t s);

struct NLKeys {  // tiny build side (< 8 rows): keys broadcast to every lane
    int32_t count;
    int32_t keys[8];
    uint32_t rids[8];
};
void nl_collect(const KeySource &k, NLKeys *out, hipStream_t s);

// materialize.hip
struct FixedSource {
    int width;                 // 4 or 8
    int is_int32;              // INT32_MIN -> NULL quirk
    const void *values;        // dense decoded base column
    const uint32_t *valid_bits;
    const uint32_t *rids;      // final row ids into the base column (nullptr = identity)
    const uint8_t *dense_pages;  // values == nullptr: an INT32 column read in place (KeySource)
    uint32_t dense_rows;
};
void materialize_fixed(const FixedSource &src, uint64_t rows, uint32_t rows_per_page,
                       uint8_t *out_pages, hipStream_t s);

struct VarcharSource {
    const uint8_t *pages;      // base column pages (for the bytes)
    const uint64_t *refs;      // decoded per-row refs of the base column
    const uint32_t *valid_bits;
    const uint32_t *rids;
};
constexpr uint32_t kVarcharChunk = 2048;  // most output rows planned per workgroup
// Rows per chunk for a VARCHAR result of `rows` rows: 2048, halved (down to
// 256) while that leaves fewer than 256 chunks, so small results still
// spread over the machine (each chunk starts a new page).
inline uint32_t varchar_chunk_rows(uint64_t rows) {
    uint32_t c = kVarcharChunk;
    while (c > 256 && (rows + c - 1) / c < 256) c /= 2;
    return c;
}
void varchar_plan(const VarcharSource &src, uint64_t rows, uint32_t chunk, uint64_t *chunk_pages, hipStream_t s);
void varchar_write(const VarcharSource &src, uint64_t rows, uint32_t chunk, const uint64_t *chunk_page_off,
                   uint8_t *out_pages, hipStream_t s);
// Adds a result column's algorithmic bytes (non-NULL value bytes, VARCHAR
// chars + 2 per value, bitmap bytes) to *acc.
void page_payload_bytes(const uint8_t *pages, uint32_t num_pages, int32_t type, uint64_t *acc, hipStream_t s);

}  // namespace cj
