// materialize.hip -- two-pass (plan, then write) columnar materialization of
// join results back into 8 KiB pages.
//
// Reference: materialization/materialize.h:85-321 and page_builders.h:92-391
// (Int32PageBuilder / VarcharPageBuilder).  Output page boundaries are not
// part of the result contract (rows are compared as a multiset), so the
// fixed-width writer uses a constant rows-per-page, which makes the page of
// output row j simply j / R: one workgroup per output page, a wave ballot
// ranks the non-NULL rows, values are packed at the aligned offset and the
// NULL bitmap is written at the page end.  VARCHAR pages are packed greedily
// with the canonical capacity rule inside independent chunks of up to 2048
// rows (plan pass counts pages, a scan places them, the write pass fills
// them); long strings are emitted as 0xffff/0xfffe page runs copied from the source,
// as copy_long_string_pages does (page_builders.h:344-360).
#include "cj_internal.hpp"
#include "device_util.hpp"

namespace cj {

// ------------------------------------------------------------ fixed width
template <int W>
__global__ __launch_bounds__(kWG) void k_mat_fixed(FixedSource src, uint64_t rows, uint32_t R, uint8_t *out) {
    __shared__ uint32_t lds4[4];
    const uint64_t pg = blockIdx.x;
    const uint64_t r0 = pg * R;
    const uint32_t cnt = (uint32_t)(rows - r0 < R ? rows - r0 : R);
    uint8_t *p = out + pg * kPage;
    const uint32_t nbytes = (cnt + 7) >> 3;
    uint8_t *bm = p + kPage - nbytes;
    uint8_t *vals = p + (W == 4 ? 4 : 8);
    uint32_t running = 0;
    for (uint32_t t0 = 0; t0 < cnt; t0 += kWG) {
        const uint32_t i = t0 + threadIdx.x;
        bool ok = false;
        uint64_t v = 0;
        if (i < cnt) {
            uint64_t rid = src.rids ? src.rids[r0 + i] : r0 + i;
            ok = !src.values || valid_bit(src.valid_bits, rid);
            if (W == 4) {
                int32_t x = src.values ? static_cast<const int32_t *>(src.values)[rid]
                                       : dense_page_value(src.dense_pages, src.dense_rows, (uint32_t)rid);
                // value_t NULL sentinel: an INT32 value equal to INT32_MIN is
                // emitted as NULL (page_builders.h:123, intermediate.h:58-62).
                if (src.is_int32 && x == kI32Null) ok = false;
                v = (uint32_t)x;
            } else {
                v = static_cast<const uint64_t *>(src.values)[rid];
            }
        }
        uint64_t m;
        uint32_t tot;
        const uint32_t vi = running + tile_rank(ok, &m, &tot, lds4);
        if (ok) {
            if (W == 4) *reinterpret_cast<uint32_t *>(vals + 4 * (size_t)vi) = (uint32_t)v;
            else *reinterpret_cast<uint64_t *>(vals + 8 * (size_t)vi) = v;
        }
        const uint32_t lane = lane_id();
        if (lane < 8) {
            uint32_t b = ((t0 + 64 * wave_id()) >> 3) + lane;
            if (b < nbytes) bm[b] = (uint8_t)(m >> (8 * lane));
        }
        running += tot;
    }
    if (threadIdx.x == 0) {
        *reinterpret_cast<uint16_t *>(p) = (uint16_t)cnt;
        *reinterpret_cast<uint16_t *>(p + 2) = (uint16_t)running;
    }
}

void materialize_fixed(const FixedSource &src, uint64_t rows, uint32_t R, uint8_t *out_pages, hipStream_t s) {
    if (!rows) return;
    uint64_t pages = (rows + R - 1) / R;
    if (src.width == 4) k_mat_fixed<4><<<(unsigned)pages, kWG, 0, s>>>(src, rows, R, out_pages);
    else k_mat_fixed<8><<<(unsigned)pages, kWG, 0, s>>>(src, rows, R, out_pages);
    CJ_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- VARCHAR
// Row descriptor: kind in bits 31..30 (0 NULL, 1 short string, 2 long
// string), length (short) or page-run length (long) in the low bits.
constexpr uint32_t kKindNull = 0, kKindShort = 1u << 30, kKindLong = 2u << 30;

// info[i] (and, when refs is given, refs[i]) of rows r0 + i, i < cnt (>= 1):
// each thread's rows have their row ids, then validity words, then refs
// loaded in one batch per phase.
__device__ __forceinline__ void load_rows(const VarcharSource &src, uint64_t r0, uint32_t cnt, uint32_t *info,
                                          uint64_t *refs) {
    constexpr uint32_t K = kVarcharChunk / kWG;
    uint64_t rid[K], ref[K];
    uint32_t vw[K];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        const uint32_t i = threadIdx.x + kWG * k;
        const uint64_t row = r0 + (i < cnt ? i : cnt - 1);
        rid[k] = src.rids ? src.rids[row] : row;
    }
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) vw[k] = src.valid_bits[rid[k] >> 5];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) ref[k] = src.refs[rid[k]];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        const uint32_t i = threadIdx.x + kWG * k;
        if (i >= cnt) continue;
        uint32_t in = kKindNull;
        if ((vw[k] >> (rid[k] & 31)) & 1u) {
            const uint32_t off = (uint32_t)(ref[k] >> 16) & 0xffffu;
            in = (off == kLongMarker ? kKindLong : kKindShort) | (uint32_t)(ref[k] & 0xffffu);
        }
        info[i] = in;
        if (refs) refs[i] = ref[k];
    }
}

struct PageEntry {
    uint32_t prel;       // output page, relative to the chunk's first page
    uint16_t first_row;  // chunk-relative first row
    uint16_t num_rows;
    uint16_t nvals;
    uint16_t nchars;
};

// Greedy page packing of a chunk's rows (info[0 .. cnt)) by the canonical
// capacity rule (Table::to_columnar VARCHAR branch, build_table.cpp:846-906):
// a short string goes to the open page unless 4 + 2 (values + 1) + chars +
// its length + (rows / 8 + 1) exceeds the page, a NULL unless 4 + 2 values +
// chars + (rows / 8 + 1) does, a long string closes the open page and takes
// its own page run.  Run by one wave: 64 rows at a time, the page state each
// row would see comes from prefix sums over the lanes, and the first lane
// that overflows (or holds a long string) ends the step.  WRITE records, per
// row, its page-table entry (short and NULL rows) or first output page (long
// rows) and, for short rows, its value index and char offset, and lists the
// long rows; returns the chunk's output pages.
template <bool WRITE>
__device__ uint32_t pack_chunk(const uint32_t *info, uint32_t cnt, uint32_t *row_page, uint16_t *row_vi,
                               uint16_t *row_cstart, PageEntry *ptab, uint32_t *n_ptab, uint16_t *long_rows,
                               uint32_t *n_long) {
    const uint32_t lane = lane_id();
    uint32_t rows = 0, nv = 0, nc = 0;  // the open page (uniform)
    uint32_t first = 0, np = 0, prel = 0, nl = 0;
    auto close_page = [&]() {
        if (WRITE && lane == 0) ptab[np] = PageEntry{prel, (uint16_t)first, (uint16_t)rows, (uint16_t)nv, (uint16_t)nc};
        ++np;
        ++prel;
        rows = nv = nc = 0;
    };
    uint32_t i = 0;
    while (i < cnt) {
        const uint32_t r = i + lane;
        const uint32_t in = r < cnt ? info[r] : kKindLong;
        const uint32_t kind = in & (3u << 30), v = in & 0x3fffffffu;
        const uint32_t dvr = kind == kKindShort ? 0x10001u : (kind == kKindNull ? 0x10000u : 0u);  // rows << 16 | values
        const uint32_t dc = kind == kKindShort ? v : 0u;
        const uint32_t ivr = wave_incl_scan<uint32_t>(dvr), ic = wave_incl_scan<uint32_t>(dc);
        const uint32_t bv = nv + ((ivr - dvr) & 0xffffu), br = rows + ((ivr - dvr) >> 16), bc = nc + ic - dc;
        bool brk = r >= cnt || kind == kKindLong;
        if (kind == kKindShort) brk = 4 + (bv + 1) * 2 + bc + v + (br / 8 + 1) > kPage;
        if (kind == kKindNull) brk = 4 + bv * 2 + bc + (br / 8 + 1) > kPage;
        const uint64_t bm = __ballot(brk);
        const uint32_t f = bm ? (uint32_t)__builtin_ctzll(bm) : 64u;  // rows [i, i + f) join the open page
        if (f) {
            if (rows == 0) first = i;
            if (WRITE && lane < f) {
                row_page[r] = np;
                row_vi[r] = (uint16_t)bv;
                row_cstart[r] = (uint16_t)bc;
            }
            const uint32_t tv = __shfl(ivr, f - 1, 64), tc = __shfl(ic, f - 1, 64);
            nv += tv & 0xffffu;
            rows += tv >> 16;
            nc += tc;
            i += f;
        }
        if (f == 64 || i >= cnt) continue;
        const uint32_t fin = __shfl(in, f, 64);
        const uint32_t fk = fin & (3u << 30), fv = fin & 0x3fffffffu;
        if (rows) close_page();
        if (fk == kKindLong) {
            if (WRITE && lane == 0) {
                row_page[i] = prel;
                long_rows[nl] = (uint16_t)i;
            }
            ++nl;
            prel += fv;
        } else {
            // the row opens the next page (it fits an empty one: a short
            // string of a source page is at most 8185 bytes)
            first = i;
            if (WRITE && lane == 0) {
                row_page[i] = np;
                row_vi[i] = 0;
                row_cstart[i] = 0;
            }
            rows = 1;
            nv = fk == kKindShort ? 1u : 0u;
            nc = fk == kKindShort ? fv : 0u;
        }
        ++i;
    }
    if (rows) close_page();
    if (WRITE && lane == 0) {
        *n_ptab = np;
        *n_long = nl;
    }
    return prel;
}

__global__ __launch_bounds__(kWG) void k_varchar_plan(VarcharSource src, uint64_t rows, uint32_t chunk,
                                                      uint64_t *chunk_pages) {
    __shared__ uint32_t info[kVarcharChunk];
    const uint64_t r0 = (uint64_t)blockIdx.x * chunk;
    const uint32_t cnt = (uint32_t)(rows - r0 < chunk ? rows - r0 : chunk);
    load_rows(src, r0, cnt, info, nullptr);
    __syncthreads();
    if (threadIdx.x >= 64) return;
    const uint32_t pages = pack_chunk<false>(info, cnt, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (threadIdx.x == 0) chunk_pages[blockIdx.x] = pages;
}

void varchar_plan(const VarcharSource &src, uint64_t rows, uint32_t chunk, uint64_t *chunk_pages, hipStream_t s) {
    if (!rows) return;
    if (chunk == 0 || chunk > kVarcharChunk) throw InvalidError("bad VARCHAR chunk size");
    unsigned chunks = (unsigned)((rows + chunk - 1) / chunk);
    k_varchar_plan<<<chunks, kWG, 0, s>>>(src, rows, chunk, chunk_pages);
    CJ_HIP(hipGetLastError());
}

__global__ __launch_bounds__(kWG) void k_varchar_write(VarcharSource src, uint64_t rows, uint32_t chunk,
                                                       const uint64_t *chunk_page_off, uint8_t *out) {
    __shared__ uint32_t info[kVarcharChunk];
    __shared__ uint32_t row_page[kVarcharChunk];  // short rows: page-table index; long: output page (rel)
    __shared__ uint16_t row_vi[kVarcharChunk];
    __shared__ uint16_t row_cstart[kVarcharChunk];
    __shared__ PageEntry ptab[kVarcharChunk];
    __shared__ uint64_t rref[kVarcharChunk];
    __shared__ uint16_t long_rows[kVarcharChunk];
    __shared__ uint32_t n_ptab, n_long;
    const uint64_t r0 = (uint64_t)blockIdx.x * chunk;
    const uint32_t cnt = (uint32_t)(rows - r0 < chunk ? rows - r0 : chunk);
    uint8_t *const base = out + chunk_page_off[blockIdx.x] * (uint64_t)kPage;
    load_rows(src, r0, cnt, info, rref);
    __syncthreads();
    if (threadIdx.x < 64) pack_chunk<true>(info, cnt, row_page, row_vi, row_cstart, ptab, &n_ptab, long_rows, &n_long);
    __syncthreads();
    // short strings: offset entry + chars
    for (uint32_t i = threadIdx.x; i < cnt; i += kWG) {
        uint32_t in = info[i];
        if ((in & (3u << 30)) != kKindShort) continue;
        const PageEntry pe = ptab[row_page[i]];
        uint8_t *dp = base + (uint64_t)pe.prel * kPage;
        uint32_t len = in & 0x3fffffffu, cs = row_cstart[i];
        *reinterpret_cast<uint16_t *>(dp + 4 + 2 * (uint32_t)row_vi[i]) = (uint16_t)(cs + len);
        const uint64_t ref = rref[i];
        const uint8_t *sp = src.pages + (ref >> 32) * (uint64_t)kPage + ((ref >> 16) & 0xffffu);
        uint8_t *cp = dp + 4 + 2 * (uint32_t)pe.nvals + cs;
        // 16 bytes per round, every load issued before the stores (a byte
        // loop waits on each load before its store)
        for (uint32_t b0 = 0; b0 < len; b0 += 16) {
            uint8_t t[16];
#pragma unroll
            for (uint32_t k = 0; k < 16; ++k) t[k] = sp[b0 + k < len ? b0 + k : len - 1];
#pragma unroll
            for (uint32_t k = 0; k < 16; ++k)
                if (b0 + k < len) cp[b0 + k] = t[k];
        }
    }
    // page headers + bitmaps: one byte (8 rows) per thread
    const uint32_t npt = n_ptab;
    for (uint32_t q = 0; q < npt; ++q) {
        const PageEntry pe = ptab[q];
        uint8_t *dp = base + (uint64_t)pe.prel * kPage;
        if (threadIdx.x == 0) {
            *reinterpret_cast<uint16_t *>(dp) = pe.num_rows;
            *reinterpret_cast<uint16_t *>(dp + 2) = pe.nvals;
        }
        const uint32_t nb = ((uint32_t)pe.num_rows + 7) >> 3;
        uint8_t *bm = dp + kPage - nb;
        for (uint32_t b = threadIdx.x; b < nb; b += kWG) {
            uint32_t byte = 0;
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k) {
                const uint32_t r = b * 8 + k;
                if (r < pe.num_rows && (info[pe.first_row + r] & (3u << 30)) == kKindShort) byte |= 1u << k;
            }
            bm[b] = (uint8_t)byte;
        }
    }
    // long strings: copy the source page run page-for-page
    const uint32_t nlong = n_long;
    for (uint32_t q = 0; q < nlong; ++q) {
        const uint32_t i = long_rows[q];
        const uint32_t in = info[i];
        uint32_t npg = in & 0x3fffffffu;
        uint64_t first = rref[i] >> 32;
        const uint4 *sp = reinterpret_cast<const uint4 *>(src.pages + first * (uint64_t)kPage);
        uint4 *dp = reinterpret_cast<uint4 *>(base + (uint64_t)row_page[i] * kPage);
        for (uint32_t w = threadIdx.x; w < npg * (kPage / 16); w += kWG) dp[w] = sp[w];
    }
}

void varchar_write(const VarcharSource &src, uint64_t rows, uint32_t chunk, const uint64_t *chunk_page_off,
                   uint8_t *out_pages, hipStream_t s) {
    if (!rows) return;
    if (chunk == 0 || chunk > kVarcharChunk) throw InvalidError("bad VARCHAR chunk size");
    unsigned chunks = (unsigned)((rows + chunk - 1) / chunk);
    k_varchar_write<<<chunks, kWG, 0, s>>>(src, rows, chunk, chunk_page_off, out_pages);
    CJ_HIP(hipGetLastError());
}

}  // namespace cj

namespace cj {
__global__ void k_page_payload_bytes(const uint8_t *pages, uint32_t num_pages, int32_t type, uint64_t *acc) {
    __shared__ uint64_t red[4];
    uint64_t b = 0;
    for (uint32_t i = blockIdx.x * kWG + threadIdx.x; i < num_pages; i += gridDim.x * kWG) {
        const uint8_t *p = pages + (size_t)i * kPage;
        uint32_t n = ld16(p), nv = ld16(p + 2);
        if (type == 3 && (n == 0xffffu || n == 0xfffeu)) {
            b += nv;
            continue;
        }
        b += (n + 7) >> 3;
        if (type == 3) b += (nv ? ld16(p + 4 + 2 * (nv - 1)) : 0u) + 2ull * nv;
        else b += (uint64_t)nv * (type == 0 ? 4u : 8u);
    }
    uint64_t tot;
    block_excl_scan<uint64_t>(b, &tot, red);
    if (threadIdx.x == 0 && tot) atomicAdd((unsigned long long *)acc, (unsigned long long)tot);
}
void page_payload_bytes(const uint8_t *pages, uint32_t num_pages, int32_t type, uint64_t *acc, hipStream_t s) {
    if (!num_pages) return;
    unsigned g = (num_pages + kWG - 1) / kWG;
    k_page_payload_bytes<<<g < 1024 ? g : 1024, kWG, 0, s>>>(pages, num_pages, type, acc);
    CJ_HIP(hipGetLastError());
}
}  // namespace cj
