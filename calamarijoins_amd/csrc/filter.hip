// filter.hip -- row filters (filter.hpp).
//
// Statement::eval over a table's columns (reference src/statement.cpp:47-249,
// include/data_model/inner_column.h): a comparison is valid(row) AND
// (value op literal); IS NULL is !valid(row); NOT inverts; AND / OR combine.
// One thread evaluates one row with the program's results on a bit stack in
// a register; a wave ballots its 64 results into the selection bitmap.
#include "filter.hpp"
#include "device_util.hpp"

namespace cj {

namespace {

// A VARCHAR value of a row: one run of chars in a page, or a long string
// spread over a 0xffff page and its 0xfffe continuations (chars at +4, count
// at +2 of each page).
struct StrView {
    const CJ_G uint8_t *pages;
    uint32_t pg, off, len, npages;
    bool lng;

    __device__ StrView(const CJ_G uint8_t *p, uint64_t ref) : pages(p) {
        pg = (uint32_t)(ref >> 32);
        const uint32_t mid = (uint32_t)(ref >> 16) & 0xffffu;
        lng = mid == kLongMarker;
        if (lng) {
            npages = (uint32_t)(ref & 0xffffu);
            off = 4;
            len = 0;
            for (uint32_t q = 0; q < npages; ++q) len += ld16(pages + (size_t)(pg + q) * kPage + 2);
        } else {
            npages = 1;
            off = mid;
            len = (uint32_t)(ref & 0xffffu);
        }
    }
    __device__ uint8_t at(uint32_t i) const {
        if (!lng) return pages[(size_t)pg * kPage + off + i];
        uint32_t q = 0;
        for (;;) {
            const uint32_t c = ld16(pages + (size_t)(pg + q) * kPage + 2);
            if (i < c || q + 1 == npages) return pages[(size_t)(pg + q) * kPage + 4 + i];
            i -= c;
            ++q;
        }
    }
};

// Bytes of the UTF-8 character at s[i] (1 for an invalid sequence), within n.
__device__ uint32_t utf8_len(const StrView &s, uint32_t i, uint32_t n) {
    const uint8_t c = s.at(i);
    uint32_t k = c < 0x80 ? 1 : (c >= 0xC2 && c < 0xE0) ? 2 : (c >= 0xE0 && c < 0xF0) ? 3 : (c >= 0xF0 && c < 0xF5) ? 4 : 1;
    if (k > n - i) return 1;
    for (uint32_t j = 1; j < k; ++j)
        if ((s.at(i + j) & 0xC0) != 0x80) return 1;
    return k;
}

// SQL LIKE as the reference's regular expression reads it (statement.h:
// like_match): % = any run of characters, _ = one character, neither
// crossing a newline; the pattern holds no newline (checked on the host), so
// a string with one cannot match.  Greedy match with one backtrack point
// (the last %), advancing by whole characters.
__device__ bool like(const StrView &s, const CJ_G uint8_t *p, uint32_t m) {
    const uint32_t n = s.len;
    uint32_t si = 0, pi = 0, mark = 0;
    int32_t star = -1;
    while (si < n) {
        const uint8_t c = s.at(si);
        if (c == '\n') return false;
        const uint8_t pc = pi < m ? p[pi] : 0;
        if (pi < m && pc == '%') {
            star = (int32_t)pi++;
            mark = si;
        } else if (pi < m && pc == '_') {
            si += utf8_len(s, si, n);
            ++pi;
        } else if (pi < m && pc == c) {
            ++si;
            ++pi;
        } else if (star >= 0) {
            pi = (uint32_t)star + 1;
            mark += utf8_len(s, mark, n);
            si = mark;
        } else {
            return false;
        }
    }
    while (pi < m && p[pi] == '%') ++pi;
    return pi == m;
}

// Byte-wise three-way comparison with a literal.
__device__ int strcmp3(const StrView &s, const CJ_G uint8_t *p, uint32_t m) {
    const uint32_t k = s.len < m ? s.len : m;
    for (uint32_t i = 0; i < k; ++i) {
        const uint8_t a = s.at(i), b = p[i];
        if (a != b) return a < b ? -1 : 1;
    }
    return s.len < m ? -1 : (s.len > m ? 1 : 0);
}

template <class T>
__device__ __forceinline__ bool cmp_op(int32_t op, T a, T b) {
    switch (op) {
    case CJ_OP_EQ: return a == b;
    case CJ_OP_NEQ: return a != b;
    case CJ_OP_LT: return a < b;
    case CJ_OP_GT: return a > b;
    case CJ_OP_LEQ: return a <= b;
    case CJ_OP_GEQ: return a >= b;
    default: return false;
    }
}

__device__ bool eval_cmp(const FNode &nd, const FCol &c, const CJ_G uint8_t *lits, uint64_t r) {
    const bool valid = valid_bit(glb(c.valid), r);
    if (nd.op == CJ_OP_IS_NULL) return !valid;
    if (nd.op == CJ_OP_IS_NOT_NULL) return valid;
    if (!valid) return false;
    switch (c.type) {
    case CJ_INT32:
        return cmp_op<int32_t>(nd.op, glb(static_cast<const int32_t *>(c.values))[r], (int32_t)nd.ival);
    case CJ_INT64:
        return cmp_op<int64_t>(nd.op, glb(static_cast<const int64_t *>(c.values))[r], nd.ival);
    case CJ_FP64:
        return cmp_op<double>(nd.op, glb(static_cast<const double *>(c.values))[r], nd.fval);
    default: {
        const StrView s(glb(c.pages), glb(c.refs)[r]);
        const CJ_G uint8_t *lit = lits + nd.soff;
        if (nd.op == CJ_OP_LIKE) return like(s, lit, nd.slen);
        if (nd.op == CJ_OP_NOT_LIKE) return !like(s, lit, nd.slen);
        return cmp_op<int>(nd.op, strcmp3(s, lit, nd.slen), 0);
    }
    }
}

}  // namespace

__global__ __launch_bounds__(kWG) void k_filter_eval(const FNode *__restrict__ prog, uint32_t n,
                                                     const FCol *__restrict__ cols, const uint8_t *__restrict__ lits,
                                                     uint64_t rows, uint32_t *sel) {
    for (uint64_t base = (uint64_t)blockIdx.x * kWG; base < rows; base += (uint64_t)gridDim.x * kWG) {
        const uint64_t r = base + threadIdx.x;
        bool res = false;
        if (r < rows) {
            uint32_t stk = 0, sp = 0;
            for (uint32_t i = 0; i < n; ++i) {
                const FNode nd = prog[i];
                if (nd.kind == CJ_F_CMP) {
                    const bool b = eval_cmp(nd, cols[nd.col], glb(lits), r);
                    stk = (stk & ~(1u << sp)) | ((uint32_t)b << sp);
                    ++sp;
                } else if (nd.kind == CJ_F_NOT) {
                    stk ^= 1u << (sp - 1);
                } else {
                    const uint32_t a = (stk >> (sp - 1)) & 1u, b = (stk >> (sp - 2)) & 1u;
                    const uint32_t v = nd.kind == CJ_F_AND ? (a & b) : (a | b);
                    --sp;
                    stk = (stk & ~(1u << (sp - 1))) | (v << (sp - 1));
                }
            }
            res = stk & 1u;
        }
        const uint64_t m = __ballot(res);
        if (lane_id() == 0) {
            const uint64_t w = (base + 64 * wave_id()) >> 5;
            sel[w] = (uint32_t)m;
            sel[w + 1] = (uint32_t)(m >> 32);
        }
    }
}

// Rows passing in each chunk of kSelChunk rows.
__global__ void k_sel_count(const uint32_t *sel, uint64_t rows, uint64_t *counts) {
    const uint64_t ch = (uint64_t)blockIdx.x * kWG + threadIdx.x;
    const uint64_t chunks = (rows + kSelChunk - 1) / kSelChunk;
    if (ch >= chunks) return;
    const uint64_t w0 = ch * (kSelChunk / 32);
    const uint64_t w1 = (rows + 31) / 32 < w0 + kSelChunk / 32 ? (rows + 31) / 32 : w0 + kSelChunk / 32;
    uint64_t c = 0;
    for (uint64_t w = w0; w < w1; ++w) {
        uint32_t bits = sel[w];
        if (w == (rows - 1) / 32 && (rows & 31)) bits &= (1u << (rows & 31)) - 1;  // rows past the end
        c += __popc(bits);
    }
    counts[ch] = c;
}

// One wave per chunk: lane l owns word l (32 rows).
__global__ __launch_bounds__(kWG) void k_sel_rids(const uint32_t *sel, uint64_t rows, const uint64_t *offs,
                                                  uint32_t *rids) {
    const uint64_t ch = (uint64_t)blockIdx.x * (kWG / 64) + wave_id();
    const uint64_t chunks = (rows + kSelChunk - 1) / kSelChunk;
    if (ch >= chunks) return;
    const uint64_t w = ch * (kSelChunk / 32) + lane_id();
    uint32_t bits = 0;
    if (w * 32 < rows) {
        bits = sel[w];
        if (w == (rows - 1) / 32 && (rows & 31)) bits &= (1u << (rows & 31)) - 1;
    }
    const uint32_t cnt = __popc(bits);
    uint64_t o = offs[ch] + wave_incl_scan<uint32_t>(cnt) - cnt;
    while (bits) {
        const uint32_t b = __ffs(bits) - 1;
        rids[o++] = (uint32_t)(w * 32 + b);
        bits &= bits - 1;
    }
}

void filter_eval(const FNode *prog, uint32_t n, const FCol *cols, const uint8_t *lits, uint64_t rows, uint32_t *sel,
                 hipStream_t s) {
    if (!rows) return;
    const uint64_t blocks = (rows + kWG - 1) / kWG;
    k_filter_eval<<<(unsigned)(blocks < 16384 ? blocks : 16384), kWG, 0, s>>>(prog, n, cols, lits, rows, sel);
    CJ_HIP(hipGetLastError());
}

void sel_count(const uint32_t *sel, uint64_t rows, uint64_t *counts, hipStream_t s) {
    const uint64_t chunks = (rows + kSelChunk - 1) / kSelChunk;
    if (!chunks) return;
    k_sel_count<<<(unsigned)((chunks + kWG - 1) / kWG), kWG, 0, s>>>(sel, rows, counts);
    CJ_HIP(hipGetLastError());
}

void sel_rids(const uint32_t *sel, uint64_t rows, const uint64_t *offs, uint32_t *rids, hipStream_t s) {
    const uint64_t chunks = (rows + kSelChunk - 1) / kSelChunk;
    if (!chunks) return;
    k_sel_rids<<<(unsigned)((chunks + 3) / 4), kWG, 0, s>>>(sel, rows, offs, rids);
    CJ_HIP(hipGetLastError());
}

}  // namespace cj
