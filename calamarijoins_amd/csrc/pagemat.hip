// pagemat.hip -- page-aligned root materialization (see pagemat.hpp).
//
// The reference builds every output column from per-thread (build_rid,
// probe_rid) chunk lists (materialize.h:85-186) and reads each value through
// ColumnarReader's page search (columnar_reader.h:300-474).  When every probe
// row matches at most once, the matched probe rows in ascending order are the
// result rows, and the work splits into streams that need no lists:
//   k_probe_count_match  keys -> per-key-page match counts + match bitmap
//   k_build_pages        keys -> direct-map records -> build-side pages
//   k_probe_filter       payload page + match bits -> probe-side page
// Page format: values at +4 (INT32) or +8, NULL bitmap in the last
// ceil(rows / 8) bytes (README.md:94-104; page_builders.h:92-151).
#include <type_traits>

#include "pagemat.hpp"
#include "probe8.hpp"

namespace cj {

constexpr uint32_t kPagesGrid = 4096;  // workgroups of the key-page passes

// ------------------------------------------------------------ counting pass
// One workgroup per key page (grid-stride).  The tile's match bits are
// gathered in LDS aligned to global bitmap words; words wholly inside the
// tile are stored, the two edge words (shared with the neighbouring tile or
// page) are OR-ed with device atomics.
template <int MODE>
__global__ __launch_bounds__(kWG) void k_probe_count_match(ProbeArgs a, uint64_t *counts, uint32_t *match_bits) {
    __shared__ uint32_t lds32[32];
    __shared__ uint64_t red[4];
    __shared__ uint32_t bits[kTile / 32 + 2];
    NLKeys nl;
    if (MODE == kModeNL) nl = *a.nl;
    const KeySource &k = a.key;
    for (uint32_t pg = blockIdx.x; pg < k.num_pages; pg += gridDim.x) {
        const uint8_t *p = k.pages + (size_t)pg * kPage;
        const uint32_t n = ld16(p), nv = ld16(p + 2);
        const bool dense = n == nv;
        const uint64_t r0 = k.row_start[pg];
        uint32_t running = 0;
        uint64_t mine = 0;
        for (uint32_t t0 = 0; t0 < n; t0 += kTile) {
            int32_t key[kRPT];
            uint32_t valid = load_keys8(p, n, nv, dense, t0, &running, key, lds32);
            uint32_t m[kRPT], s[kRPT], e[kRPT];
            lookup8<MODE>(valid, key, a.ht, nl, a.dt, m, s, e);
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) mine += m[j];
            const uint64_t g0 = r0 + t0;
            const uint32_t sh = (uint32_t)(g0 & 31);
            const uint32_t nt = n - t0 < kTile ? n - t0 : kTile;
            const uint32_t nwords = (sh + nt + 31) >> 5;
            for (uint32_t i = threadIdx.x; i < kTile / 32 + 2; i += kWG) bits[i] = 0;
            __syncthreads();
            const uint32_t lane = lane_id();
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) {
                const uint64_t mk = __ballot(m[j] != 0);
                if (lane == 0 && mk) {
                    const uint32_t b = sh + kWG * j + 64 * wave_id();  // bit of the chunk's first row
                    const uint32_t w = b >> 5, o = b & 31;
                    const uint64_t lo = mk << o;
                    atomicOr(&bits[w], (uint32_t)lo);
                    atomicOr(&bits[w + 1], (uint32_t)(lo >> 32));
                    if (o) atomicOr(&bits[w + 2], (uint32_t)(mk >> (64 - o)));
                }
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < nwords; i += kWG) {
                const uint32_t word = bits[i];
                const bool edge = (i == 0 && sh) || (i == nwords - 1 && ((sh + nt) & 31));
                uint32_t *dst = match_bits + (g0 >> 5) + i;
                if (edge) {
                    if (word) atomicOr(dst, word);
                } else {
                    *dst = word;
                }
            }
            __syncthreads();
        }
        uint64_t tot;
        block_excl_scan<uint64_t>(mine, &tot, red);
        if (threadIdx.x == 0) counts[pg] = tot;
    }
}

void probe_count_match(const ProbeArgs &a, uint64_t *counts, uint32_t *match_bits, hipStream_t s) {
    const uint32_t units = a.key.num_pages;
    if (!units) return;
    const uint32_t grid = units < kPagesGrid ? units : kPagesGrid;
    if (a.mode == kModeNL) k_probe_count_match<kModeNL><<<grid, kWG, 0, s>>>(a, counts, match_bits);
    else if (a.mode == kModeDirect) k_probe_count_match<kModeDirect><<<grid, kWG, 0, s>>>(a, counts, match_bits);
    else k_probe_count_match<kModeHash><<<grid, kWG, 0, s>>>(a, counts, match_bits);
    CJ_HIP(hipGetLastError());
}

__global__ void k_key_page_out_pages(const uint64_t *counts, uint32_t pages, uint64_t *p4, uint64_t *p8) {
    const uint32_t i = blockIdx.x * kWG + threadIdx.x;
    if (i >= pages) return;
    const uint64_t m = counts[i];
    p4[i] = (m + kCap4 - 1) / kCap4;
    p8[i] = (m + kCap8 - 1) / kCap8;
}

void key_page_out_pages(const uint64_t *counts, uint32_t pages, uint64_t *p4, uint64_t *p8, hipStream_t s) {
    if (!pages) return;
    k_key_page_out_pages<<<(pages + kWG - 1) / kWG, kWG, 0, s>>>(counts, pages, p4, p8);
    CJ_HIP(hipGetLastError());
}

// --------------------------------------------------------- build-side pages
constexpr uint32_t kMaxPagesPerKeyPage = 3;  // <= 2047 matches: 2 INT32 or 3 8-byte pages

// Dword `d` of a record held as NQ 16-byte chunks (d is uniform).
template <int NQ>
__device__ __forceinline__ uint32_t rec_dword(const uint4 (&r)[NQ], uint32_t d) {
    uint32_t x = 0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        x = d == 4u * q + 0 ? r[q].x : x;
        x = d == 4u * q + 1 ? r[q].y : x;
        x = d == 4u * q + 2 ? r[q].z : x;
        x = d == 4u * q + 3 ? r[q].w : x;
    }
    return x;
}

template <int NQ>
__global__ __launch_bounds__(kWG) void k_build_pages(BuildPagesArgs a) {
    __shared__ uint32_t lds32[32];
    __shared__ uint32_t obm[kMatMaxCols][kMaxPagesPerKeyPage][kCap4 / 32 + 1];  // output bitmaps
    __shared__ uint32_t pstart[kMatMaxCols][kMaxPagesPerKeyPage + 1];  // values before each page
    __shared__ uint32_t nn_run[kMatMaxCols];
    NLKeys nl{};
    const KeySource &k = a.key;
    const uint32_t words = kMaxPagesPerKeyPage * (kCap4 / 32 + 1);
    for (uint32_t pg = blockIdx.x; pg < k.num_pages; pg += gridDim.x) {
        const uint8_t *p = k.pages + (size_t)pg * kPage;
        const uint32_t n = ld16(p), nv = ld16(p + 2);
        const bool dense = n == nv;
        for (uint32_t i = threadIdx.x; i < a.ncols * words; i += kWG) (&obm[0][0][0])[i] = 0;
        if (threadIdx.x < kMatMaxCols) nn_run[threadIdx.x] = 0;
        __syncthreads();
        uint32_t running = 0, run_m = 0;
        for (uint32_t t0 = 0; t0 < n; t0 += kTile) {
            int32_t key[kRPT];
            const uint32_t valid = load_keys8(p, n, nv, dense, t0, &running, key, lds32);
            uint32_t m[kRPT], slot[kRPT], e[kRPT];
            HashTableDev ht{};
            lookup8<kModeDirect>(valid, key, ht, nl, a.dt, m, slot, e);
            uint4 rec[kRPT][NQ];
            uint32_t mm = 0;
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) {
                mm |= (m[j] ? 1u : 0u) << j;
                // unconditional (slot 0 when unmatched): all record loads in flight together
                const uint4 *r = reinterpret_cast<const uint4 *>(a.dt.records + (uint64_t)slot[j] * a.dt.stride);
#pragma unroll
                for (int q = 0; q < NQ; ++q) rec[j][q] = r[q];
            }
            uint32_t ko[kRPT];
            uint64_t mask[kRPT];
            const uint32_t tot_m = rank_slots<kRPT>(mm, ko, mask, lds32);
            const uint32_t fl_dw = a.dt.flags_offset >> 2;
            for (uint32_t c = 0; c < a.ncols; ++c) {
                const BuildPageCol &col = a.cols[c];
                const uint32_t cap = col.width == 4 ? kCap4 : kCap8;
                const uint32_t dw = col.rec_offset >> 2;
                uint64_t v[kRPT];
                uint32_t ok = 0;
#pragma unroll
                for (uint32_t j = 0; j < kRPT; ++j) {
                    const uint32_t lo = rec_dword<NQ>(rec[j], dw);
                    const uint32_t hi = col.width == 8 ? rec_dword<NQ>(rec[j], dw + 1) : 0u;
                    const uint32_t fl = rec_dword<NQ>(rec[j], fl_dw);
                    v[j] = ((uint64_t)hi << 32) | lo;
                    bool o = ((mm >> j) & 1u) && ((fl >> col.rec_bit) & 1u);
                    if (col.is_int32 && (int32_t)lo == kI32Null) o = false;  // page_builders.h:123
                    ok |= (o ? 1u : 0u) << j;
                }
                uint32_t so[kRPT];
                uint64_t omask[kRPT];
                const uint32_t base_nn = nn_run[c];
                const uint32_t tot_ok = rank_slots<kRPT>(ok, so, omask, lds32);
                // the first row of each output page records the values before it
#pragma unroll
                for (uint32_t j = 0; j < kRPT; ++j)
                    if ((mm >> j) & 1u) {
                        const uint32_t r = run_m + ko[j];
                        if (r % cap == 0 && r / cap < kMaxPagesPerKeyPage) pstart[c][r / cap] = base_nn + so[j];
                    }
                __syncthreads();
                const uint64_t first = col.width == 4 ? a.p4_off[pg] : a.p8_off[pg];
#pragma unroll
                for (uint32_t j = 0; j < kRPT; ++j)
                    if ((ok >> j) & 1u) {
                        const uint32_t r = run_m + ko[j];
                        const uint32_t pp = r / cap;
                        if (pp >= kMaxPagesPerKeyPage) continue;
                        const uint32_t vi = base_nn + so[j] - pstart[c][pp];
                        uint8_t *page = col.out + (first + pp) * kPage;
                        if (col.width == 4) *reinterpret_cast<uint32_t *>(page + 4 + 4 * (size_t)vi) = (uint32_t)v[j];
                        else *reinterpret_cast<uint64_t *>(page + 8 + 8 * (size_t)vi) = v[j];
                        const uint32_t rr = r - pp * cap;
                        atomicOr(&obm[c][pp][rr >> 5], 1u << (rr & 31));
                    }
                __syncthreads();
                if (threadIdx.x == 0) nn_run[c] = base_nn + tot_ok;
            }
            run_m += tot_m;
        }
        __syncthreads();
        // headers and bitmaps
        for (uint32_t c = 0; c < a.ncols; ++c) {
            const BuildPageCol &col = a.cols[c];
            const uint32_t cap = col.width == 4 ? kCap4 : kCap8;
            const uint32_t np = (run_m + cap - 1) / cap;
            const uint64_t first = col.width == 4 ? a.p4_off[pg] : a.p8_off[pg];
            for (uint32_t pp = 0; pp < np && pp < kMaxPagesPerKeyPage; ++pp) {
                uint8_t *page = col.out + (first + pp) * kPage;
                const uint32_t rows = run_m - pp * cap < cap ? run_m - pp * cap : cap;
                const uint32_t nb = (rows + 7) >> 3;
                const uint8_t *src = reinterpret_cast<const uint8_t *>(obm[c][pp]);
                for (uint32_t b = threadIdx.x; b < nb; b += kWG) page[kPage - nb + b] = src[b];
                if (threadIdx.x == 0) {
                    const uint32_t end = pp + 1 < np ? pstart[c][pp + 1] : nn_run[c];
                    *reinterpret_cast<uint16_t *>(page) = (uint16_t)rows;
                    *reinterpret_cast<uint16_t *>(page + 2) = (uint16_t)(end - pstart[c][pp]);
                }
            }
        }
        __syncthreads();
    }
}

void build_pages(const BuildPagesArgs &a, hipStream_t s) {
    const uint32_t units = a.key.num_pages;
    if (!units || !a.ncols) return;
    if (a.dt.stride % 16 || a.dt.stride == 0 || a.dt.stride > 80) throw InvalidError("bad direct-map record stride");
    const uint32_t grid = units < kPagesGrid ? units : kPagesGrid;
    switch (a.dt.stride / 16) {
        case 1: k_build_pages<1><<<grid, kWG, 0, s>>>(a); break;
        case 2: k_build_pages<2><<<grid, kWG, 0, s>>>(a); break;
        case 3: k_build_pages<3><<<grid, kWG, 0, s>>>(a); break;
        case 4: k_build_pages<4><<<grid, kWG, 0, s>>>(a); break;
        default: k_build_pages<5><<<grid, kWG, 0, s>>>(a); break;
    }
    CJ_HIP(hipGetLastError());
}

// --------------------------------------------------------- probe-side pages
// One wave per payload page: does any row of [rs, rs + n) have its match bit?
__global__ __launch_bounds__(kWG) void k_page_match_flags(const uint8_t *pages, uint32_t num_pages,
                                                          const uint64_t *row_start, const uint32_t *match_bits,
                                                          uint64_t *nonempty) {
    const uint32_t q = blockIdx.x * (kWG / 64) + wave_id();
    if (q >= num_pages) return;
    const uint64_t rs = row_start[q], re = row_start[q + 1];
    uint32_t any = 0;
    if (re > rs) {
        const uint64_t w0 = rs >> 5, w1 = (re - 1) >> 5;
        for (uint64_t w = w0 + lane_id(); w <= w1; w += 64) {
            uint32_t x = match_bits[w];
            if (w == w0) x &= ~0u << (rs & 31);
            if (w == w1 && (re & 31)) x &= (1u << (re & 31)) - 1u;
            any |= x;
        }
    }
    const uint64_t b = __ballot(any != 0);
    if (lane_id() == 0) nonempty[q] = b ? 1 : 0;
}

void page_match_flags(const uint8_t *pages, uint32_t num_pages, const uint64_t *row_start,
                      const uint32_t *match_bits, uint64_t *nonempty, hipStream_t s) {
    if (!num_pages) return;
    const uint32_t per = kWG / 64;
    k_page_match_flags<<<(num_pages + per - 1) / per, kWG, 0, s>>>(pages, num_pages, row_start, match_bits,
                                                                   nonempty);
    CJ_HIP(hipGetLastError());
}

constexpr uint32_t kWaveMaxRows = 4096;  // pages k_probe_filter_wave handles
constexpr uint32_t kWaveBatch = 16;      // 64-row chunks whose loads are issued together

// One wave per payload page of at most kWaveMaxRows rows, no barriers: for
// each batch of 16 chunks of 64 rows, the page-bitmap bytes and match words
// are loaded, value indices come from ballots of the validity bits, all
// value loads are issued, then matched values are stored compacted and the
// output bitmap is assembled in the wave's LDS slice.
template <int W>
__global__ __launch_bounds__(kWG) void k_probe_filter_wave(FilterArgs a) {
    using T = typename std::conditional<W == 4, uint32_t, uint64_t>::type;
    __shared__ uint32_t obm_all[kWG / 64][kWaveMaxRows / 32];
    const uint32_t lane = lane_id();
    const uint32_t q = blockIdx.x * (kWG / 64) + wave_id();
    if (q >= a.num_pages) return;
    const uint64_t o = a.pos[q];
    if (a.pos[q + 1] == o) return;  // no matched row
    const uint8_t *p = a.pages + (size_t)q * kPage;
    const uint32_t n = ld16(p), nv = ld16(p + 2);
    if (n > kWaveMaxRows) return;  // k_probe_filter's page
    const uint64_t rs = a.row_start[q];
    const bool dense = n == nv;
    const uint8_t *bm = p + kPage - ((n + 7) >> 3);
    const T *vals = reinterpret_cast<const T *>(p + (W == 4 ? 4 : 8));
    uint8_t *out = a.out + o * kPage;
    T *ovals = reinterpret_cast<T *>(out + (W == 4 ? 4 : 8));
    const uint32_t *mw = a.match_bits + (rs >> 5);
    const uint32_t msh = (uint32_t)(rs & 31);
    uint32_t *obm = obm_all[wave_id()];
    for (uint32_t i = lane; i < (n + 31) / 32; i += 64) obm[i] = 0;
    __builtin_amdgcn_wave_barrier();
    const uint64_t lt = lanemask_lt();
    uint32_t vbase = 0, mbase = 0, obase = 0;
    // every load below is unconditional with a clamped in-page address, so
    // the compiler issues a whole batch before the first wait
    const uint32_t nlast = n - 1, vlast = nv ? nv - 1 : 0;
    for (uint32_t c0 = 0; c0 < n; c0 += 64 * kWaveBatch) {
        uint32_t vbits = 0, mbits = 0;
        uint32_t byte[kWaveBatch], word[kWaveBatch];
#pragma unroll
        for (uint32_t c = 0; c < kWaveBatch; ++c) {
            const uint32_t lr = c0 + 64 * c + lane;
            const uint32_t lc = lr < n ? lr : nlast;
            byte[c] = bm[lc >> 3];
            word[c] = mw[(msh + lc) >> 5];
        }
#pragma unroll
        for (uint32_t c = 0; c < kWaveBatch; ++c) {
            const uint32_t lr = c0 + 64 * c + lane;
            const uint32_t lc = lr < n ? lr : nlast;
            const bool in = lr < n;
            if (in && (dense || ((byte[c] >> (lc & 7)) & 1u))) vbits |= 1u << c;
            if (in && ((word[c] >> ((msh + lc) & 31)) & 1u)) mbits |= 1u << c;
        }
        T v[kWaveBatch];
        uint32_t okbits = 0;
#pragma unroll
        for (uint32_t c = 0; c < kWaveBatch; ++c) {
            const uint64_t vm = __ballot((vbits >> c) & 1u);
            const uint32_t vi = vbase + (uint32_t)__popcll(vm & lt);
            vbase += (uint32_t)__popcll(vm);
            v[c] = vals[vi < nv ? vi : vlast];
            if ((((vbits & mbits) >> c) & 1u) && vi < nv) okbits |= 1u << c;
        }
#pragma unroll
        for (uint32_t c = 0; c < kWaveBatch; ++c) {
            bool ok = (okbits >> c) & 1u;
            if (W == 4 && a.is_int32 && (int32_t)(uint32_t)v[c] == kI32Null) ok = false;  // page_builders.h:123
            const uint64_t om = __ballot(ok);
            const uint64_t mm = __ballot((mbits >> c) & 1u);
            if (ok) {
                const uint32_t so = obase + (uint32_t)__popcll(om & lt);
                const uint32_t ko = mbase + (uint32_t)__popcll(mm & lt);
                ovals[so] = v[c];
                atomicOr(&obm[ko >> 5], 1u << (ko & 31));
            }
            obase += (uint32_t)__popcll(om);
            mbase += (uint32_t)__popcll(mm);
        }
    }
    __builtin_amdgcn_wave_barrier();
    const uint32_t nbo = (mbase + 7) >> 3;
    const uint8_t *ob = reinterpret_cast<const uint8_t *>(obm);
    for (uint32_t b = lane; b < nbo; b += 64) out[kPage - nbo + b] = ob[b];
    if (lane == 0) {
        *reinterpret_cast<uint16_t *>(out) = (uint16_t)mbase;
        *reinterpret_cast<uint16_t *>(out + 2) = (uint16_t)obase;
    }
}

// One workgroup per payload page of more than kWaveMaxRows rows (sparse
// pages): the page is staged in LDS with 16-byte loads, the value index of each row comes from a rank over the page bitmap,
// and matched rows are compacted into the output page.
template <int W>
__global__ __launch_bounds__(kWG) void k_probe_filter(FilterArgs a) {
    __shared__ uint4 stage[kPage / 16];
    __shared__ uint32_t obm[kPage / 4];  // output bitmap (up to 65535 rows)
    __shared__ uint32_t lds32[32];
    const uint32_t q = blockIdx.x;
    const uint64_t out_pg = a.pos[q];
    if (a.pos[q + 1] == out_pg) return;  // no matched row
    if (ld16(a.pages + (size_t)q * kPage) <= kWaveMaxRows) return;  // k_probe_filter_wave's page
    const uint4 *src = reinterpret_cast<const uint4 *>(a.pages + (size_t)q * kPage);
    for (uint32_t i = threadIdx.x; i < kPage / 16; i += kWG) stage[i] = src[i];
    const uint64_t rs = a.row_start[q];
    __syncthreads();
    const uint8_t *sp = reinterpret_cast<const uint8_t *>(stage);
    const uint32_t n = ld16(sp), nv = ld16(sp + 2);
    const bool dense = n == nv;
    const uint8_t *bm = sp + kPage - ((n + 7) >> 3);
    const uint32_t obw = (n + 31) >> 5;
    for (uint32_t i = threadIdx.x; i < obw; i += kWG) obm[i] = 0;
    __syncthreads();
    uint8_t *out = a.out + out_pg * kPage;
    uint32_t run_v = 0, run_m = 0, run_ok = 0;
    for (uint32_t t0 = 0; t0 < n; t0 += kTile) {
        uint32_t valid = 0, matched = 0;
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j) {
            const uint32_t lr = t0 + kWG * j + threadIdx.x;
            if (lr < n) {
                if (dense || ((bm[lr >> 3] >> (lr & 7)) & 1u)) valid |= 1u << j;
                const uint64_t g = rs + lr;
                if ((a.match_bits[g >> 5] >> (g & 31)) & 1u) matched |= 1u << j;
            }
        }
        uint32_t vi[kRPT], ko[kRPT], so[kRPT];
        uint64_t mask[kRPT];
        const uint32_t tv = rank_slots<kRPT>(valid, vi, mask, lds32);
        uint64_t v[kRPT];
        uint32_t ok = 0;
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j) {
            v[j] = 0;
            const uint32_t x = run_v + vi[j];
            if (((valid & matched) >> j) & 1u && x < nv) {
                v[j] = W == 4 ? (uint64_t)(uint32_t)ld32(sp + 4 + 4 * (size_t)x) : ld64(sp + 8 + 8 * (size_t)x);
                bool o = true;
                if (W == 4 && a.is_int32 && (int32_t)(uint32_t)v[j] == kI32Null) o = false;  // page_builders.h:123
                ok |= (o ? 1u : 0u) << j;
            }
        }
        const uint32_t tm = rank_slots<kRPT>(matched, ko, mask, lds32);
        const uint32_t tok = rank_slots<kRPT>(ok, so, mask, lds32);
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j)
            if ((ok >> j) & 1u) {
                const uint32_t x = run_ok + so[j], r = run_m + ko[j];
                if (W == 4) *reinterpret_cast<uint32_t *>(out + 4 + 4 * (size_t)x) = (uint32_t)v[j];
                else *reinterpret_cast<uint64_t *>(out + 8 + 8 * (size_t)x) = v[j];
                atomicOr(&obm[r >> 5], 1u << (r & 31));
            }
        run_v += tv;
        run_m += tm;
        run_ok += tok;
    }
    __syncthreads();
    const uint32_t nb = (run_m + 7) >> 3;
    const uint8_t *ob = reinterpret_cast<const uint8_t *>(obm);
    for (uint32_t b = threadIdx.x; b < nb; b += kWG) out[kPage - nb + b] = ob[b];
    if (threadIdx.x == 0) {
        *reinterpret_cast<uint16_t *>(out) = (uint16_t)run_m;
        *reinterpret_cast<uint16_t *>(out + 2) = (uint16_t)run_ok;
    }
}

void probe_filter(const FilterArgs &a, hipStream_t s) {
    if (!a.num_pages) return;
    const uint32_t per = kWG / 64;
    const uint32_t grid = (a.num_pages + per - 1) / per;
    if (a.width == 4) k_probe_filter_wave<4><<<grid, kWG, 0, s>>>(a);
    else k_probe_filter_wave<8><<<grid, kWG, 0, s>>>(a);
    if (a.max_page_rows > kWaveMaxRows) {
        if (a.width == 4) k_probe_filter<4><<<a.num_pages, kWG, 0, s>>>(a);
        else k_probe_filter<8><<<a.num_pages, kWG, 0, s>>>(a);
    }
    CJ_HIP(hipGetLastError());
}

__global__ void k_gather_u64s(U64List l, uint64_t *dst) {
    if (threadIdx.x < l.n) dst[threadIdx.x] = *l.p[threadIdx.x];
}

void gather_u64s(const U64List &l, uint64_t *dst, hipStream_t s) {
    if (!l.n) return;
    k_gather_u64s<<<1, 64, 0, s>>>(l, dst);
    CJ_HIP(hipGetLastError());
}

}  // namespace cj
