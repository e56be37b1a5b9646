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
// with the canonical capacity rule inside independent 2048-row chunks (plan
// pass counts pages, a scan places them, the write pass fills them); long
// strings are emitted as 0xffff/0xfffe page runs copied from the source,
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
            ok = valid_bit(src.valid_bits, rid);
            if (W == 4) {
                int32_t x = static_cast<const int32_t *>(src.values)[rid];
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

__device__ __forceinline__ uint32_t row_info(const VarcharSource &src, uint64_t row, uint64_t *ref_out) {
    uint64_t rid = src.rids ? src.rids[row] : row;
    if (!valid_bit(src.valid_bits, rid)) return kKindNull;
    uint64_t ref = src.refs[rid];
    *ref_out = ref;
    uint32_t off = (uint32_t)(ref >> 16) & 0xffffu;
    if (off == kLongMarker) return kKindLong | (uint32_t)(ref & 0xffffu);
    return kKindShort | (uint32_t)(ref & 0xffffu);
}

// Canonical capacity rule (Table::to_columnar VARCHAR branch,
// build_table.cpp:846-906).
struct PackState {
    uint32_t num_rows, nvals, nchars;
};
__device__ __forceinline__ bool need_flush_short(const PackState &st, uint32_t len) {
    return 4 + (st.nvals + 1) * 2 + st.nchars + len + (st.num_rows / 8 + 1) > kPage;
}
__device__ __forceinline__ bool need_flush_null(const PackState &st) {
    return 4 + st.nvals * 2 + st.nchars + (st.num_rows / 8 + 1) > kPage;
}

__global__ __launch_bounds__(kWG) void k_varchar_plan(VarcharSource src, uint64_t rows, uint64_t *chunk_pages) {
    __shared__ uint32_t info[kVarcharChunk];
    const uint64_t r0 = (uint64_t)blockIdx.x * kVarcharChunk;
    const uint32_t cnt = (uint32_t)(rows - r0 < kVarcharChunk ? rows - r0 : kVarcharChunk);
    for (uint32_t i = threadIdx.x; i < cnt; i += kWG) {
        uint64_t ref;
        info[i] = row_info(src, r0 + i, &ref);
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    uint64_t pages = 0;
    PackState st{0, 0, 0};
    for (uint32_t i = 0; i < cnt; ++i) {
        uint32_t in = info[i], kind = in & (3u << 30), v = in & 0x3fffffffu;
        if (kind == kKindLong) {
            if (st.num_rows) { ++pages; st = PackState{0, 0, 0}; }
            pages += v;
        } else if (kind == kKindShort) {
            if (need_flush_short(st, v)) { ++pages; st = PackState{0, 0, 0}; }
            st.nvals++; st.nchars += v; st.num_rows++;
        } else {
            if (need_flush_null(st)) { ++pages; st = PackState{0, 0, 0}; }
            st.num_rows++;
        }
    }
    if (st.num_rows) ++pages;
    chunk_pages[blockIdx.x] = pages;
}

void varchar_plan(const VarcharSource &src, uint64_t rows, uint64_t *chunk_pages, hipStream_t s) {
    if (!rows) return;
    unsigned chunks = (unsigned)((rows + kVarcharChunk - 1) / kVarcharChunk);
    k_varchar_plan<<<chunks, kWG, 0, s>>>(src, rows, chunk_pages);
    CJ_HIP(hipGetLastError());
}

struct PageEntry {
    uint32_t prel;       // output page, relative to the chunk's first page
    uint16_t first_row;  // chunk-relative first row
    uint16_t num_rows;
    uint16_t nvals;
    uint16_t nchars;
};

__global__ __launch_bounds__(kWG) void k_varchar_write(VarcharSource src, uint64_t rows,
                                                       const uint64_t *chunk_page_off, uint8_t *out) {
    __shared__ uint32_t info[kVarcharChunk];
    __shared__ uint32_t row_page[kVarcharChunk];  // short rows: page-table index; long: output page (rel)
    __shared__ uint16_t row_vi[kVarcharChunk];
    __shared__ uint16_t row_cstart[kVarcharChunk];
    __shared__ PageEntry ptab[kVarcharChunk];
    __shared__ uint32_t n_ptab;
    const uint64_t r0 = (uint64_t)blockIdx.x * kVarcharChunk;
    const uint32_t cnt = (uint32_t)(rows - r0 < kVarcharChunk ? rows - r0 : kVarcharChunk);
    uint8_t *const base = out + chunk_page_off[blockIdx.x] * (uint64_t)kPage;
    for (uint32_t i = threadIdx.x; i < cnt; i += kWG) {
        uint64_t ref;
        info[i] = row_info(src, r0 + i, &ref);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t prel = 0, np = 0;
        PackState st{0, 0, 0};
        auto open_page = [&](uint32_t first) {
            ptab[np] = PageEntry{prel, (uint16_t)first, 0, 0, 0};
        };
        auto close_page = [&]() {
            ptab[np].num_rows = (uint16_t)st.num_rows;
            ptab[np].nvals = (uint16_t)st.nvals;
            ptab[np].nchars = (uint16_t)st.nchars;
            ++np;
            ++prel;
            st = PackState{0, 0, 0};
        };
        for (uint32_t i = 0; i < cnt; ++i) {
            uint32_t in = info[i], kind = in & (3u << 30), v = in & 0x3fffffffu;
            if (kind == kKindLong) {
                if (st.num_rows) close_page();
                row_page[i] = prel;
                prel += v;
                continue;
            }
            if (kind == kKindShort ? need_flush_short(st, v) : need_flush_null(st)) close_page();
            if (st.num_rows == 0) open_page(i);
            row_page[i] = np;
            row_vi[i] = (uint16_t)st.nvals;
            row_cstart[i] = (uint16_t)st.nchars;
            if (kind == kKindShort) { st.nvals++; st.nchars += v; }
            st.num_rows++;
        }
        if (st.num_rows) close_page();
        n_ptab = np;
    }
    __syncthreads();
    // short strings: offset entry + chars
    for (uint32_t i = threadIdx.x; i < cnt; i += kWG) {
        uint32_t in = info[i];
        if ((in & (3u << 30)) != kKindShort) continue;
        const PageEntry pe = ptab[row_page[i]];
        uint8_t *dp = base + (uint64_t)pe.prel * kPage;
        uint32_t len = in & 0x3fffffffu, cs = row_cstart[i];
        *reinterpret_cast<uint16_t *>(dp + 4 + 2 * (uint32_t)row_vi[i]) = (uint16_t)(cs + len);
        uint64_t rid = src.rids ? src.rids[r0 + i] : r0 + i;
        uint64_t ref = src.refs[rid];
        const uint8_t *sp = src.pages + (ref >> 32) * (uint64_t)kPage + ((ref >> 16) & 0xffffu);
        uint8_t *cp = dp + 4 + 2 * (uint32_t)pe.nvals + cs;
        for (uint32_t b = 0; b < len; ++b) cp[b] = sp[b];
    }
    // page headers + bitmaps
    for (uint32_t q = threadIdx.x; q < n_ptab; q += kWG) {
        const PageEntry pe = ptab[q];
        uint8_t *dp = base + (uint64_t)pe.prel * kPage;
        *reinterpret_cast<uint16_t *>(dp) = pe.num_rows;
        *reinterpret_cast<uint16_t *>(dp + 2) = pe.nvals;
        uint32_t nb = ((uint32_t)pe.num_rows + 7) >> 3;
        uint8_t *bm = dp + kPage - nb;
        for (uint32_t b = 0; b < nb; ++b) {
            uint32_t byte = 0;
            for (uint32_t k = 0; k < 8; ++k) {
                uint32_t r = b * 8 + k;
                if (r < pe.num_rows && (info[pe.first_row + r] & (3u << 30)) == kKindShort) byte |= 1u << k;
            }
            bm[b] = (uint8_t)byte;
        }
    }
    // long strings: copy the source page run page-for-page
    for (uint32_t i = 0; i < cnt; ++i) {
        uint32_t in = info[i];
        if ((in & (3u << 30)) != kKindLong) continue;
        uint32_t npg = in & 0x3fffffffu;
        uint64_t rid = src.rids ? src.rids[r0 + i] : r0 + i;
        uint64_t first = src.refs[rid] >> 32;
        const uint4 *sp = reinterpret_cast<const uint4 *>(src.pages + first * (uint64_t)kPage);
        uint4 *dp = reinterpret_cast<uint4 *>(base + (uint64_t)row_page[i] * kPage);
        for (uint32_t w = threadIdx.x; w < npg * (kPage / 16); w += kWG) dp[w] = sp[w];
    }
}

void varchar_write(const VarcharSource &src, uint64_t rows, const uint64_t *chunk_page_off, uint8_t *out_pages,
                   hipStream_t s) {
    if (!rows) return;
    unsigned chunks = (unsigned)((rows + kVarcharChunk - 1) / kVarcharChunk);
    k_varchar_write<<<chunks, kWG, 0, s>>>(src, rows, chunk_page_off, out_pages);
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
