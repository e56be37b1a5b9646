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
#include <algorithm>
#include <mutex>
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
            uint32_t hit[kRPT];
            lookup8<MODE>(valid, key, a.ht, nl, a.dt, m, s, e, hit);
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

// Direct mode: one wave per key page, no barriers.  Per segment of 2048 rows
// the 32 chunk masks (64 rows each) go to the wave's LDS slice, then lane i
// assembles global bitmap word i of the segment from them: interior words
// are stored, the two edge words OR-ed with device atomics.
constexpr uint32_t kCountBatch = 32;

// Segment-local row mask msk (32-bit word k: rows 32k .. 32k + 31 of the
// nseg rows from global row g0) into the global bitmap: interior words are
// stored, the two edge words (shared with a neighbouring page) OR-ed.
__device__ __forceinline__ void store_segment_bits(const uint32_t *msk, uint32_t *match_bits, uint64_t g0,
                                                   uint32_t nseg, uint32_t lane) {
    const uint32_t sh = (uint32_t)(g0 & 31);
    const uint32_t nwords = (sh + nseg + 31) >> 5;
    for (uint32_t i = lane; i < nwords; i += 64) {
        // bits [32 i - sh, 32 i - sh + 32) of the segment's row mask
        uint32_t w;
        if (i == 0) {
            w = msk[0] << sh;
        } else {
            const uint32_t st = 32 * i - sh, d = st >> 5, o = st & 31;
            w = o ? (msk[d] >> o) | (msk[d + 1] << (32 - o)) : msk[d];
        }
        const bool edge = (i == 0 && sh) || (i == nwords - 1 && ((sh + nseg) & 31));
        uint32_t *dst = match_bits + (g0 >> 5) + i;
        if (edge) {
            if (w) atomicOr(dst, w);
        } else {
            *dst = w;
        }
    }
}

// A key page without NULL keys and of at most kTile rows, whole: lane l
// takes rows 4l .. 4l + 3 of each 256-row block with one 16-byte load (all
// eight blocks in flight together); the four match bits of a lane are
// merged into 32-bit row-mask words across groups of eight lanes.  Returns
// the page's matches; msk receives the row mask.
constexpr uint32_t kDenseBlk = kTile / 256;  // 256-row blocks of a page count_dense_page takes

// The 16-byte key loads of count_dense_page, issued before the page header
// is known: rows past the page's last are read from inside the page (the 16
// bytes at +4 + 4 r stay below 8192 for r <= 2043) and dropped.
__device__ __forceinline__ void load_dense_keys(const uint8_t *p, uint32_t lane, u32x4u (&kv)[kDenseBlk]) {
#pragma unroll
    for (uint32_t b = 0; b < kDenseBlk; ++b) {
        const uint32_t r = 256 * b + 4 * lane;
        kv[b] = *reinterpret_cast<const u32x4u *>(p + 4 + 4 * (r < 2040 ? r : 2040));
    }
}

__device__ __forceinline__ uint32_t count_dense_page(const DirectTable &dt, const u32x4u (&kv)[kDenseBlk], uint32_t n,
                                                     uint32_t *msk, uint32_t lane) {
    constexpr uint32_t kBlk = kDenseBlk;
    uint32_t slot[kBlk][4], in[kBlk];
#pragma unroll
    for (uint32_t b = 0; b < kBlk; ++b) {
        in[b] = 0;
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t r = 256 * b + 4 * lane + j;
            const int64_t off = (int64_t)(int32_t)kv[b][j] - (int64_t)dt.min_key;
            const bool ok = (r < n) & (off >= 0) & ((uint64_t)off < dt.range);
            slot[b][j] = ok ? (uint32_t)off : 0u;
            in[b] |= (uint32_t)ok << j;
        }
    }
    if (!dt.full) {
        uint32_t word[kBlk][4];
#pragma unroll
        for (uint32_t b = 0; b < kBlk; ++b)
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) word[b][j] = dt.presence[slot[b][j] >> 5];
#pragma unroll
        for (uint32_t b = 0; b < kBlk; ++b)
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j)
                if (!((word[b][j] >> (slot[b][j] & 31)) & 1u)) in[b] &= ~(1u << j);
    }
    uint32_t total = 0;
#pragma unroll
    for (uint32_t b = 0; b < kBlk; ++b) {
        total += (uint32_t)__popc(in[b]);
        uint32_t v = in[b] << (4 * (lane & 7));
        v |= __shfl_xor(v, 1, 64);
        v |= __shfl_xor(v, 2, 64);
        v |= __shfl_xor(v, 4, 64);
        if ((lane & 7) == 0) msk[8 * b + (lane >> 3)] = v;
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) total += __shfl_xor(total, d, 64);
    return total;
}

// DENSE: every key page holds n == nv <= kTile rows (known on the host from
// the column), so only count_dense_page is compiled in: the general path's
// batched registers would cap the occupancy at 3 waves a SIMD.
template <bool DENSE>
__global__ __launch_bounds__(kWG) void k_count_match_direct(ProbeArgs a, uint64_t *counts, uint32_t *match_bits) {
    __shared__ uint32_t msk_all[kWG / 64][2 * (kTile / 64) + 2];
    const uint32_t lane = lane_id();
    const KeySource &k = a.key;
    const DirectTable &dt = a.dt;
    const uint32_t pg = blockIdx.x * (kWG / 64) + wave_id();
    if (pg >= k.num_pages) return;
    const uint8_t *p = k.pages + (size_t)pg * kPage;
    // the keys of a dense page are loaded together with its header (one
    // round trip instead of two); the empty asm keeps the loads here
    u32x4u kv[kDenseBlk];
    load_dense_keys(p, lane, kv);
    const uint32_t n = ld16(p), nv = ld16(p + 2);
    const uint64_t r0 = k.row_start[pg];
#pragma unroll
    for (uint32_t b = 0; b < kDenseBlk; ++b) asm volatile("" : "+v"(kv[b]));
    const bool dense = n == nv;
    const uint8_t *bm = p + kPage - ((n + 7) >> 3);
    const int32_t *keys = reinterpret_cast<const int32_t *>(p + 4);
    uint32_t *msk = msk_all[wave_id()];
    const uint64_t lt = lanemask_lt();
    const uint32_t nlast = n ? n - 1 : 0, vlast = nv ? nv - 1 : 0;
    if (DENSE || (dense && n <= kTile)) {
        const uint32_t total = count_dense_page(dt, kv, n, msk, lane);
        if (lane == 0) msk[2 * (kTile / 64)] = msk[2 * (kTile / 64) + 1] = 0;
        __builtin_amdgcn_wave_barrier();
        store_segment_bits(msk, match_bits, r0, n, lane);
        if (lane == 0) counts[pg] = total;
        return;
    }
    if constexpr (!DENSE) {  // sparse pages, or pages of more than kTile rows
        uint32_t vbase = 0;
        uint64_t total = 0;
        for (uint32_t s0 = 0; s0 < n; s0 += kTile) {
            const uint32_t nseg = n - s0 < kTile ? n - s0 : kTile;
            for (uint32_t b0 = 0; b0 < kTile / 64; b0 += kCountBatch) {
                uint32_t vbits = 0, vi[kCountBatch];
                if (dense) {
#pragma unroll
                    for (uint32_t c = 0; c < kCountBatch; ++c) {
                        const uint32_t lr = s0 + 64 * (b0 + c) + lane;
                        vi[c] = lr;
                        if (lr < n) vbits |= 1u << c;
                    }
                } else {
                    uint32_t byte[kCountBatch];
#pragma unroll
                    for (uint32_t c = 0; c < kCountBatch; ++c) {
                        const uint32_t lr = s0 + 64 * (b0 + c) + lane;
                        byte[c] = bm[(lr < n ? lr : nlast) >> 3];
                    }
#pragma unroll
                    for (uint32_t c = 0; c < kCountBatch; ++c) {
                        const uint32_t lr = s0 + 64 * (b0 + c) + lane;
                        const bool v = lr < n && ((byte[c] >> (lr & 7)) & 1u);
                        const uint64_t vm = __ballot(v);
                        vi[c] = vbase + (uint32_t)__popcll(vm & lt);
                        vbase += (uint32_t)__popcll(vm);
                        if (v) vbits |= 1u << c;
                    }
                }
                // all key loads first, unconditionally (a load under a
                // short-circuit condition is issued and awaited one at a time)
                int32_t key[kCountBatch];
#pragma unroll
                for (uint32_t c = 0; c < kCountBatch; ++c) key[c] = keys[vi[c] < nv ? vi[c] : vlast];
                uint32_t slot[kCountBatch], word[kCountBatch], in = 0;
#pragma unroll
                for (uint32_t c = 0; c < kCountBatch; ++c) {
                    const int64_t off = (int64_t)key[c] - (int64_t)dt.min_key;
                    const bool ok = ((vbits >> c) & 1u) & (vi[c] < nv) & (off >= 0) & ((uint64_t)off < dt.range);
                    slot[c] = ok ? (uint32_t)off : 0u;
                    in |= (uint32_t)ok << c;
                }
                if (dt.full) {
#pragma unroll
                    for (uint32_t c = 0; c < kCountBatch; ++c) word[c] = ~0u;
                } else {
#pragma unroll
                    for (uint32_t c = 0; c < kCountBatch; ++c) word[c] = dt.presence[slot[c] >> 5];
                }
#pragma unroll
                for (uint32_t c = 0; c < kCountBatch; ++c) {
                    const bool m = ((in >> c) & 1u) && ((word[c] >> (slot[c] & 31)) & 1u);
                    const uint64_t mm = __ballot(m);
                    total += (uint32_t)__popcll(mm);
                    if (lane == 0) {
                        msk[2 * (b0 + c)] = (uint32_t)mm;
                        msk[2 * (b0 + c) + 1] = (uint32_t)(mm >> 32);
                    }
                }
            }
            if (lane == 0) msk[2 * (kTile / 64)] = msk[2 * (kTile / 64) + 1] = 0;
            __builtin_amdgcn_wave_barrier();
            store_segment_bits(msk, match_bits, r0 + s0, nseg, lane);
            __builtin_amdgcn_wave_barrier();
        }
        if (lane == 0) counts[pg] = total;
    }
}

// The counting pass over dense key pages for a map whose keys do not fill
// its range, with the presence bitmap in LDS: every probe key costs a
// random presence lookup, and from L2 those run at its request rate (~240 G
// lookups/s on MI355X); from LDS they do not.  A bitmap larger than a
// workgroup's LDS is taken in windows of kLpWinWords words, one pass over
// the key pages per window (each key is looked up in the pass whose window
// holds its word); a workgroup of kLpWG threads stays resident and keeps the
// same key pages in every pass, so a page's count is added up by the same
// lane and its match words are OR-ed into the zeroed bitmap.
constexpr uint32_t kLpWG = 512;
constexpr uint32_t kLpWinWords = 39 * 1024;  // 156 KiB of presence words a pass

__device__ __forceinline__ void or_segment_bits(const uint32_t *msk, uint32_t *match_bits, uint64_t g0, uint32_t nseg,
                                                uint32_t lane) {
    const uint32_t sh = (uint32_t)(g0 & 31);
    const uint32_t nwords = (sh + nseg + 31) >> 5;
    for (uint32_t i = lane; i < nwords; i += 64) {
        uint32_t w;
        if (i == 0) {
            w = msk[0] << sh;
        } else {
            const uint32_t st = 32 * i - sh, d = st >> 5, o = st & 31;
            w = o ? (msk[d] >> o) | (msk[d + 1] << (32 - o)) : msk[d];
        }
        if (w) atomicOr(match_bits + (g0 >> 5) + i, w);
    }
}

__global__ __launch_bounds__(kLpWG) void k_count_match_direct_lp(ProbeArgs a, uint64_t *counts, uint32_t *match_bits,
                                                                 uint32_t words) {
    extern __shared__ uint32_t pres[];
    __shared__ uint32_t msk_all[kLpWG / 64][2 * (kTile / 64) + 2];
    const uint32_t lane = lane_id(), wv = wave_id();
    const KeySource &k = a.key;
    const DirectTable &dt = a.dt;
    uint32_t *msk = msk_all[wv];
    for (uint32_t w0 = 0; w0 < words; w0 += kLpWinWords) {
        const uint32_t wn = words - w0 < kLpWinWords ? words - w0 : kLpWinWords;
        __syncthreads();  // the previous window is no longer read
        for (uint32_t i = threadIdx.x; i < wn; i += kLpWG) lds(pres)[i] = dt.presence[w0 + i];
        __syncthreads();
        // the next page's keys are loaded while this one is counted
        const uint32_t step = gridDim.x * (kLpWG / 64);
        uint32_t pg = blockIdx.x * (kLpWG / 64) + wv;
        u32x4u nkv[kDenseBlk];
        if (pg < k.num_pages) load_dense_keys(k.pages + (size_t)pg * kPage, lane, nkv);
        for (; pg < k.num_pages; pg += step) {
            const uint8_t *p = k.pages + (size_t)pg * kPage;
            u32x4u kv[kDenseBlk];
#pragma unroll
            for (uint32_t b = 0; b < kDenseBlk; ++b) kv[b] = nkv[b];
            const uint32_t nx = pg + step < k.num_pages ? pg + step : pg;
            load_dense_keys(k.pages + (size_t)nx * kPage, lane, nkv);
            const uint32_t n = ld16(p);  // dense: every row has its key (checked on the host)
            const uint64_t r0 = k.row_start[pg];
            uint32_t in[kDenseBlk], word[kDenseBlk][4], slot[kDenseBlk][4];
#pragma unroll
            for (uint32_t b = 0; b < kDenseBlk; ++b) {
                in[b] = 0;
#pragma unroll
                for (uint32_t j = 0; j < 4; ++j) {
                    const uint32_t r = 256 * b + 4 * lane + j;
                    const int64_t off = (int64_t)(int32_t)kv[b][j] - (int64_t)dt.min_key;
                    const bool ok = (r < n) & (off >= 0) & ((uint64_t)off < dt.range) &&
                                    (uint64_t)off >= 32ull * w0 && (uint64_t)off < 32ull * (w0 + wn);
                    slot[b][j] = ok ? (uint32_t)off : 32u * w0;
                    in[b] |= (uint32_t)ok << j;
                }
            }
#pragma unroll
            for (uint32_t b = 0; b < kDenseBlk; ++b)
#pragma unroll
                for (uint32_t j = 0; j < 4; ++j) word[b][j] = lds(pres)[(slot[b][j] >> 5) - w0];
            uint32_t total = 0;
#pragma unroll
            for (uint32_t b = 0; b < kDenseBlk; ++b) {
#pragma unroll
                for (uint32_t j = 0; j < 4; ++j)
                    if (!((word[b][j] >> (slot[b][j] & 31)) & 1u)) in[b] &= ~(1u << j);
                total += (uint32_t)__popc(in[b]);
                uint32_t v = in[b] << (4 * (lane & 7));
                v |= __shfl_xor(v, 1, 64);
                v |= __shfl_xor(v, 2, 64);
                v |= __shfl_xor(v, 4, 64);
                if ((lane & 7) == 0) msk[8 * b + (lane >> 3)] = v;
            }
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) total += __shfl_xor(total, d, 64);
            if (lane == 0) msk[2 * (kTile / 64)] = msk[2 * (kTile / 64) + 1] = 0;
            __builtin_amdgcn_wave_barrier();
            or_segment_bits(msk, match_bits, r0, n, lane);
            __builtin_amdgcn_wave_barrier();  // msk is rewritten by the next page
            if (lane == 0) counts[pg] = (w0 ? counts[pg] : 0ull) + total;
        }
    }
}

static void launch_count_lp(const ProbeArgs &a, uint64_t *counts, uint32_t *match_bits, uint32_t words,
                            uint32_t units, hipStream_t s) {
    static std::once_flag raised;  // dynamic LDS above 64 KiB needs the kernel attribute
    std::call_once(raised, [] {
        CJ_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_count_match_direct_lp),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kLpWinWords * 4)));
    });
    int dev = 0, cus = 0, per = 0;
    CJ_HIP(hipGetDevice(&dev));
    CJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t dyn = (size_t)std::min<uint32_t>(words, kLpWinWords) * 4;
    CJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_count_match_direct_lp, kLpWG, dyn));
    const uint32_t need = (units + kLpWG / 64 - 1) / (kLpWG / 64);
    const uint32_t g = std::min<uint32_t>(need, (uint32_t)std::max(per, 1) * (uint32_t)std::max(cus, 1));
    k_count_match_direct_lp<<<g, kLpWG, dyn, s>>>(a, counts, match_bits, words);
}

void probe_count_match(const ProbeArgs &a, uint64_t *counts, uint32_t *match_bits, uint32_t dense_rows,
                       hipStream_t s) {
    const uint32_t units = a.key.num_pages;
    if (!units) return;
    const uint32_t grid = units < kPagesGrid ? units : kPagesGrid;
    if (a.mode == kModeDirect) {
        const uint32_t per = kWG / 64, g = (units + per - 1) / per;
        const bool dense = dense_rows && dense_rows <= kTile;
        const uint64_t words = (a.dt.range + 31) / 32;
        // the LDS bitmap pays off for at most two windows over many pages
        if (dense && !a.dt.full && words <= 2ull * kLpWinWords && units >= 4096)
            launch_count_lp(a, counts, match_bits, (uint32_t)words, units, s);
        else if (dense)
            k_count_match_direct<true><<<g, kWG, 0, s>>>(a, counts, match_bits);
        else
            k_count_match_direct<false><<<g, kWG, 0, s>>>(a, counts, match_bits);
    } else if (a.mode == kModeNL) {
        k_probe_count_match<kModeNL><<<grid, kWG, 0, s>>>(a, counts, match_bits);
    } else {
        k_probe_count_match<kModeHash><<<grid, kWG, 0, s>>>(a, counts, match_bits);
    }
    CJ_HIP(hipGetLastError());
}

// One wave per key page with matches: bitmap bytes (sparse pages), keys and
// match words loaded in batches of 8 chunks of 64 rows, then the build ids
// of the matched rows (one record load each), then the pairs, each chunk's
// matched lanes writing a contiguous run.
constexpr uint32_t kPairBatch = 8;
__global__ __launch_bounds__(kWG) void k_direct_pairs(ProbeArgs a, const uint32_t *match_bits, const uint64_t *off,
                                                      uint32_t *ob, uint32_t *op) {
    const uint32_t lane = lane_id();
    const KeySource &k = a.key;
    const DirectTable &dt = a.dt;
    const uint32_t pg = blockIdx.x * (kWG / 64) + wave_id();
    if (pg >= k.num_pages) return;
    uint64_t o = off[pg];
    if (off[pg + 1] == o) return;
    const uint8_t *p = k.pages + (size_t)pg * kPage;
    const uint32_t n = ld16(p), nv = ld16(p + 2);
    const bool dense = n == nv;
    const uint8_t *bm = p + kPage - ((n + 7) >> 3);
    const int32_t *keys = reinterpret_cast<const int32_t *>(p + 4);
    const uint64_t r0 = k.row_start[pg];
    const uint32_t *mw = match_bits + (r0 >> 5);
    const uint32_t msh = (uint32_t)(r0 & 31);
    const uint64_t lt = lanemask_lt();
    const uint32_t nlast = n ? n - 1 : 0, vlast = nv ? nv - 1 : 0;
    uint32_t vbase = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += 64 * kPairBatch) {
        uint32_t vbits = 0, vi[kPairBatch];
        if (dense) {
#pragma unroll
            for (uint32_t c = 0; c < kPairBatch; ++c) {
                const uint32_t lr = c0 + 64 * c + lane;
                vi[c] = lr;
                vbits |= (uint32_t)(lr < n) << c;
            }
        } else {
            uint32_t byte[kPairBatch];
#pragma unroll
            for (uint32_t c = 0; c < kPairBatch; ++c) {
                const uint32_t lr = c0 + 64 * c + lane;
                byte[c] = bm[(lr < n ? lr : nlast) >> 3];
            }
#pragma unroll
            for (uint32_t c = 0; c < kPairBatch; ++c) {
                const uint32_t lr = c0 + 64 * c + lane;
                const bool v = (lr < n) & (((byte[c] >> (lr & 7)) & 1u) != 0);
                const uint64_t vm = __ballot(v);
                vi[c] = vbase + (uint32_t)__popcll(vm & lt);
                vbase += (uint32_t)__popcll(vm);
                vbits |= (uint32_t)v << c;
            }
        }
        int32_t key[kPairBatch];
        uint32_t word[kPairBatch];
#pragma unroll
        for (uint32_t c = 0; c < kPairBatch; ++c) {
            const uint32_t lr = c0 + 64 * c + lane;
            key[c] = keys[vi[c] < nv ? vi[c] : vlast];
            word[c] = mw[(msh + (lr < n ? lr : nlast)) >> 5];
        }
        uint32_t mbits = 0, bid[kPairBatch];
#pragma unroll
        for (uint32_t c = 0; c < kPairBatch; ++c) {
            const uint32_t lr = c0 + 64 * c + lane;
            const int64_t so = (int64_t)key[c] - (int64_t)dt.min_key;
            // a set match bit means the key is in range and present
            const bool m = ((vbits >> c) & 1u) & (vi[c] < nv) & (((word[c] >> ((msh + lr) & 31)) & 1u) != 0) &
                           (so >= 0) & ((uint64_t)so < dt.range);
            bid[c] = m ? (uint32_t)so : 0u;
            mbits |= (uint32_t)m << c;
        }
        if (dt.rid_offset >= 0) {
#pragma unroll
            for (uint32_t c = 0; c < kPairBatch; ++c)
                bid[c] = *reinterpret_cast<const uint32_t *>(dt.records + (uint64_t)bid[c] * dt.stride +
                                                             (uint32_t)dt.rid_offset);
        }
#pragma unroll
        for (uint32_t c = 0; c < kPairBatch; ++c) {
            const bool m = (mbits >> c) & 1u;
            const uint64_t mm = __ballot(m);
            if (m) {
                const uint64_t at = o + (uint64_t)__popcll(mm & lt);
                if (ob) ob[at] = bid[c];
                if (op) op[at] = (uint32_t)(r0 + c0 + 64 * c + lane);
            }
            o += (uint64_t)__popcll(mm);
        }
    }
}

void direct_pairs(const ProbeArgs &a, const uint32_t *match_bits, const uint64_t *off, uint32_t *ob, uint32_t *op,
                  hipStream_t s) {
    const uint32_t units = a.key.num_pages;
    if (!units) return;
    k_direct_pairs<<<(units + kWG / 64 - 1) / (kWG / 64), kWG, 0, s>>>(a, match_bits, off, ob, op);
    CJ_HIP(hipGetLastError());
}

// ------------------------------------------------ direct maps, row-id lists
// One wave per unit of kInterChunk rows of the intermediate; rows are taken
// 64 a chunk, 8 chunks at a time: row ids, then validity words and keys (a
// decoded column, or the base column's pages read in place).  Units and
// chunks start at multiples of 64 rows, so a chunk's match mask is two whole
// words of the match bitmap.
constexpr uint32_t kInterBatch = 8;

// Keys of rows r0 + c0 + 64 c + lane (c < kInterBatch) of the intermediate
// whose unit holds n rows; bit c of the result: the row exists, its key is
// non-NULL and is not INT32_MIN (a NULL once it went through an
// intermediate).  Rows outside use[] (bit clear) read the unit's first row.
__device__ __forceinline__ uint32_t inter_keys(const KeySource &k, uint64_t r0, uint32_t n, uint32_t c0, uint32_t lane,
                                               uint32_t use, int32_t (&key)[kInterBatch]) {
    uint64_t rid[kInterBatch];
    uint32_t vw[kInterBatch];
#pragma unroll
    for (uint32_t c = 0; c < kInterBatch; ++c) {
        const uint32_t r = c0 + 64 * c + lane;
        const uint64_t row = r0 + ((((use >> c) & 1u) != 0) & (r < n) ? r : 0u);
        rid[c] = k.rids ? k.rids[row] : row;
    }
    if (k.values) {
#pragma unroll
        for (uint32_t c = 0; c < kInterBatch; ++c) vw[c] = k.valid_bits[rid[c] >> 5];
#pragma unroll
        for (uint32_t c = 0; c < kInterBatch; ++c) key[c] = k.values[rid[c]];
    } else {
#pragma unroll
        for (uint32_t c = 0; c < kInterBatch; ++c) {
            vw[c] = ~0u;
            key[c] = dense_page_value(k.dense_pages, k.dense_rows, (uint32_t)rid[c]);
        }
    }
    uint32_t ok = 0;
#pragma unroll
    for (uint32_t c = 0; c < kInterBatch; ++c) {
        const uint32_t r = c0 + 64 * c + lane;
        ok |= (uint32_t)((r < n) & (((vw[c] >> (rid[c] & 31)) & 1u) != 0) & (key[c] != kI32Null)) << c;
    }
    return ok & use;
}

__global__ __launch_bounds__(kWG) void k_inter_count_match(ProbeArgs a, uint64_t *counts, uint32_t *match_bits) {
    const uint32_t lane = lane_id();
    const KeySource &k = a.key;
    const DirectTable &dt = a.dt;
    const uint32_t u = blockIdx.x * (kWG / 64) + wave_id();
    if ((uint64_t)u * kInterChunk >= k.rows) return;
    const uint64_t r0 = (uint64_t)u * kInterChunk;
    const uint32_t n = k.rows - r0 < kInterChunk ? (uint32_t)(k.rows - r0) : kInterChunk;
    uint64_t total = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += 64 * kInterBatch) {
        int32_t key[kInterBatch];
        const uint32_t ok = inter_keys(k, r0, n, c0, lane, ~0u, key);
        uint32_t slot[kInterBatch], word[kInterBatch], in = 0;
#pragma unroll
        for (uint32_t c = 0; c < kInterBatch; ++c) {
            const int64_t off = (int64_t)key[c] - (int64_t)dt.min_key;
            const bool m = ((ok >> c) & 1u) & (off >= 0) & ((uint64_t)off < dt.range);
            slot[c] = m ? (uint32_t)off : 0u;
            in |= (uint32_t)m << c;
        }
        if (dt.full) {
#pragma unroll
            for (uint32_t c = 0; c < kInterBatch; ++c) word[c] = ~0u;
        } else {
#pragma unroll
            for (uint32_t c = 0; c < kInterBatch; ++c) word[c] = dt.presence[slot[c] >> 5];
        }
#pragma unroll
        for (uint32_t c = 0; c < kInterBatch; ++c) {
            const bool m = ((in >> c) & 1u) && ((word[c] >> (slot[c] & 31)) & 1u);
            const uint64_t mm = __ballot(m);
            total += (uint32_t)__popcll(mm);
            const uint32_t s = c0 + 64 * c;
            if (s < n && lane < 2) match_bits[((r0 + s) >> 5) + lane] = (uint32_t)(mm >> (32 * lane));
        }
    }
    if (lane == 0) counts[u] = total;
}

void inter_count_match(const ProbeArgs &a, uint64_t *counts, uint32_t *match_bits, hipStream_t s) {
    const uint64_t units = (a.key.rows + kInterChunk - 1) / kInterChunk;
    if (!units) return;
    if (a.mode != kModeDirect || a.key.mode != KEY_INTER) throw InvalidError("inter_count_match needs a direct map and a row-id list");
    const uint32_t per = kWG / 64;
    k_inter_count_match<<<(unsigned)((units + per - 1) / per), kWG, 0, s>>>(a, counts, match_bits);
    CJ_HIP(hipGetLastError());
}

// Pairs of one unit from its match bits: keys only for the matched rows
// (the others read the unit's first row), then the build ids.
__global__ __launch_bounds__(kWG) void k_inter_pairs(ProbeArgs a, const uint32_t *match_bits, const uint64_t *off,
                                                     uint32_t *ob, uint32_t *op) {
    const uint32_t lane = lane_id();
    const KeySource &k = a.key;
    const DirectTable &dt = a.dt;
    const uint32_t u = blockIdx.x * (kWG / 64) + wave_id();
    if ((uint64_t)u * kInterChunk >= k.rows) return;
    uint64_t o = off[u];
    if (off[u + 1] == o) return;
    const uint64_t r0 = (uint64_t)u * kInterChunk;
    const uint32_t n = k.rows - r0 < kInterChunk ? (uint32_t)(k.rows - r0) : kInterChunk;
    const uint64_t lt = lanemask_lt();
    for (uint32_t c0 = 0; c0 < n; c0 += 64 * kInterBatch) {
        uint32_t use = 0;
#pragma unroll
        for (uint32_t c = 0; c < kInterBatch; ++c) {
            const uint32_t s = c0 + 64 * c;
            const uint32_t w = match_bits[((r0 + (s < n ? s : 0u)) >> 5) + (lane >> 5)];
            use |= (uint32_t)((s + lane < n) & (((w >> (lane & 31)) & 1u) != 0)) << c;
        }
        int32_t key[kInterBatch];
        inter_keys(k, r0, n, c0, lane, use, key);
        uint32_t bid[kInterBatch];
#pragma unroll
        for (uint32_t c = 0; c < kInterBatch; ++c) {
            // a set match bit means the key is in range and present
            const int64_t so = (int64_t)key[c] - (int64_t)dt.min_key;
            const bool m = ((use >> c) & 1u) & (so >= 0) & ((uint64_t)so < dt.range);
            bid[c] = m ? (uint32_t)so : 0u;
        }
        if (dt.rid_offset >= 0) {
#pragma unroll
            for (uint32_t c = 0; c < kInterBatch; ++c)
                bid[c] = *reinterpret_cast<const uint32_t *>(dt.records + (uint64_t)bid[c] * dt.stride +
                                                             (uint32_t)dt.rid_offset);
        }
#pragma unroll
        for (uint32_t c = 0; c < kInterBatch; ++c) {
            const bool m = (use >> c) & 1u;
            const uint64_t mm = __ballot(m);
            if (m) {
                const uint64_t at = o + (uint64_t)__popcll(mm & lt);
                if (ob) ob[at] = bid[c];
                if (op) op[at] = (uint32_t)(r0 + c0 + 64 * c + lane);
            }
            o += (uint64_t)__popcll(mm);
        }
    }
}

void inter_pairs(const ProbeArgs &a, const uint32_t *match_bits, const uint64_t *off, uint32_t *ob, uint32_t *op,
                 hipStream_t s) {
    const uint64_t units = (a.key.rows + kInterChunk - 1) / kInterChunk;
    if (!units) return;
    const uint32_t per = kWG / 64;
    k_inter_pairs<<<(unsigned)((units + per - 1) / per), kWG, 0, s>>>(a, match_bits, off, ob, op);
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
// Dword `d` of a record held as NQ 16-byte chunks (d is uniform): the chunk,
// then its 8-byte half, then a shift, so that few uniform lane masks stay
// live across the row loop.
template <int NQ>
__device__ __forceinline__ uint32_t rec_dword(const uint4 (&r)[NQ], uint32_t d) {
    uint32_t x = r[0].x, y = r[0].y, z = r[0].z, w = r[0].w;
#pragma unroll
    for (int q = 1; q < NQ; ++q) {
        const bool here = (d >> 2) == (uint32_t)q;
        x = here ? r[q].x : x;
        y = here ? r[q].y : y;
        z = here ? r[q].z : z;
        w = here ? r[q].w : w;
    }
    const uint64_t h = (d & 2) ? (((uint64_t)w << 32) | z) : (((uint64_t)y << 32) | x);
    return (uint32_t)(h >> (32 * (d & 1)));
}

constexpr uint32_t kBuildColsPerWave = 4;  // build columns one launch writes
constexpr uint32_t kBuildBatch = 8;        // 64-row chunks whose loads are issued together

// Output state of one column while a wave walks one key page: result rows
// fill the column's pages in order.  The page being filled starts qoff bytes
// after the column's first page and holds pos rows; the key page's non-NULL
// values so far number nn, pb of them on earlier pages (all wave-uniform).
// Lane j holds validity word j of the page (rows 32j .. 32j + 31; a page has
// at most 62 words) in bw.
struct OutCol {
    uint32_t nn, pb, qoff, pos;
    uint32_t bw;
};

// OR the validity bits pk (bit i: page row pos + i) into the page's words.
__device__ __forceinline__ void or_bits(uint32_t &bw, uint64_t pk, uint32_t pos, uint32_t lane) {
    const uint32_t w0 = pos >> 5, sh = pos & 31;
    const uint32_t lo = (uint32_t)(pk << sh);
    const uint32_t mid = (uint32_t)(pk >> (32 - sh));
    const uint32_t top = sh ? (uint32_t)(pk >> (64 - sh)) : 0u;
    const uint32_t e = lane - w0;
    bw |= e == 0 ? lo : (e == 1 ? mid : (e == 2 ? top : 0u));
}

// Finish the page being filled (rows result rows, nvals values): header and
// validity bitmap (at the page end, one byte per 8 rows); start the next.
__device__ __forceinline__ void close_page(OutCol &o, CJ_G uint8_t *base, uint32_t rows, uint32_t nvals,
                                           uint32_t lane) {
    CJ_G uint8_t *page = base + o.qoff;
    // an opaque row count keeps the compiler from hoisting this rare path's
    // lane masks out of the row loop, where they would crowd the scalar
    // registers
    asm volatile("" : "+v"(rows));
    const uint32_t nbytes = (rows + 7) >> 3;
    if (4 * lane < nbytes) {
        CJ_G uint8_t *dst = page + kPage - nbytes + 4 * lane;
        const uint32_t k = nbytes - 4 * lane;
        dst[0] = (uint8_t)o.bw;
        if (k > 1) dst[1] = (uint8_t)(o.bw >> 8);
        if (k > 2) dst[2] = (uint8_t)(o.bw >> 16);
        if (k > 3) dst[3] = (uint8_t)(o.bw >> 24);
    }
    if (lane == 0) *reinterpret_cast<CJ_G uint32_t *>(page) = rows | (nvals << 16);
    o.qoff += kPage;
    o.pos = 0;
    o.bw = 0;
}

// Where one column's fields sit in a direct-map record.
struct ColPick {
    uint32_t lo_dw;  // dword of the value (the low dword of an 8-byte value)
    uint32_t bit;    // validity bit in the flags dword
    bool w8;         // 8-byte value
    bool i32;        // INT32: INT32_MIN is NULL in a result
};

// One column, one 64-row chunk: the matched lanes (mask mm; lane rank among
// them `rank`, tm in all; dst: the lane's slot in the matched-first order)
// with a non-NULL value (ok; value words lo, hi) write it at the column's
// next value slot, and every matched lane appends its validity bit; a page
// that fills inside the chunk is finished.
__device__ __forceinline__ void emit_col(bool w8, CJ_G uint8_t *base, uint32_t lo, uint32_t hi, bool ok, uint64_t mm,
                                         bool matched, uint32_t rank, uint32_t tm, uint32_t dst, OutCol &o,
                                         uint32_t lane, uint64_t lt) {
    const uint32_t cap = w8 ? kCap8 : kCap4;
    const uint64_t om = __ballot(ok);
    const uint32_t room = cap - o.pos;  // >= 1: a full page is finished at once
    const bool cross = tm >= room;
    uint32_t nb = 0;  // values before the next page
    if (cross) nb = o.nn + (uint32_t)__popcll(om & __ballot(matched && rank < room));
    if (ok) {
        const bool next = cross && rank >= room;
        const uint32_t x = o.nn + (uint32_t)__popcll(om & lt) - (next ? nb : o.pb);
        const uint32_t off = o.qoff + (next ? kPage : 0u);
        if (w8) __builtin_nontemporal_store(((uint64_t)hi << 32) | lo, reinterpret_cast<CJ_G uint64_t *>(base + (off + 8 + 8 * x)));
        else __builtin_nontemporal_store(lo, reinterpret_cast<CJ_G uint32_t *>(base + (off + 4 + 4 * x)));
    }
    uint64_t pk;  // validity bits of the chunk's result rows
    if (om == mm) {
        pk = tm == 64 ? ~0ull : (1ull << tm) - 1ull;
    } else {
        const int got = __builtin_amdgcn_ds_permute((int)(dst << 2), ok ? 1 : 0);
        pk = __ballot(lane < tm && got != 0);
    }
    if (!cross) {
        or_bits(o.bw, pk, o.pos, lane);
        o.pos += tm;
    } else {
        or_bits(o.bw, room < 64 ? pk & ((1ull << room) - 1ull) : pk, o.pos, lane);
        close_page(o, base, cap, nb - o.pb, lane);
        o.pb = nb;
        if (tm > room) {
            or_bits(o.bw, pk >> room, 0, lane);
            o.pos = tm - room;
        }
    }
    o.nn += (uint32_t)__popcll(om);
}

// One wave per key page, no LDS.  For each batch of 8 chunks of 64 rows:
// bitmap bytes (sparse pages), keys, match words and records are loaded with
// clamped addresses so every load of a phase is in flight together; then,
// chunk by chunk, ballots give each matched row its result row within the
// key page and each column's non-NULL values their slot.  NC (2 or 4) bounds
// the columns of the launch.
template <int NQ, int NC>
__global__ __launch_bounds__(kWG) void k_build_pages(BuildPagesArgs a, uint32_t c_lo) {
    // chunks per batch: fewer for wide records and many columns, whose
    // registers would not fit
    constexpr uint32_t kB = NQ * NC <= 4 ? kBuildBatch : kBuildBatch / 2;
    const uint32_t lane = lane_id();
    const uint32_t pg = blockIdx.x * (kWG / 64) + wave_id();
    if (pg >= a.key.num_pages) return;
    const uint32_t nc = a.ncols - c_lo < (uint32_t)NC ? a.ncols - c_lo : (uint32_t)NC;
    const CJ_G uint8_t *p = glb(a.key.pages) + (size_t)pg * kPage;
    const uint32_t n = ld16(p), nv = ld16(p + 2);
    const bool dense = n == nv;
    const CJ_G uint8_t *bm = p + kPage - ((n + 7) >> 3);
    const CJ_G int32_t *keys = reinterpret_cast<const CJ_G int32_t *>(p + 4);
    const uint64_t r0 = a.key.row_start[pg];
    const uint32_t *mw = a.match_bits + (r0 >> 5);
    const uint32_t msh = (uint32_t)(r0 & 31);
    const uint64_t lt = lanemask_lt();
    const uint32_t nlast = n ? n - 1 : 0, vlast = nv ? nv - 1 : 0;
    const uint32_t fl_dw = a.dt.flags_offset >> 2;
    CJ_G uint8_t *base[NC];
    ColPick pick[NC];
    OutCol oc[NC];
    // a block sized from a match count that no longer holds must not be
    // overrun: a key page whose pages would pass a block's end writes
    // nothing (the host compares the page counts with the blocks afterwards)
    for (uint32_t k = 0; k < nc; ++k) {
        const BuildPageCol &col = a.cols[c_lo + k];
        if ((col.width == 8 ? a.p8_off : a.p4_off)[pg + 1] > col.cap_pages) return;
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        const BuildPageCol &col = a.cols[c_lo + (k < (int)nc ? k : 0)];
        base[k] = glb(col.out) + (col.width == 8 ? a.p8_off[pg] : a.p4_off[pg]) * kPage;
        pick[k] = ColPick{col.rec_offset >> 2, col.rec_bit, col.width == 8, col.is_int32 != 0};
        oc[k] = OutCol{0, 0, 0, 0, 0u};
    }
    uint32_t vbase = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += 64 * kB) {
        uint32_t vbits = 0;
        uint32_t vi[kB];
        if (dense) {
#pragma unroll
            for (uint32_t c = 0; c < kB; ++c) {
                const uint32_t lr = c0 + 64 * c + lane;
                vi[c] = lr;
                if (lr < n) vbits |= 1u << c;
            }
        } else {
            uint32_t byte[kB];
#pragma unroll
            for (uint32_t c = 0; c < kB; ++c) {
                const uint32_t lr = c0 + 64 * c + lane;
                byte[c] = bm[(lr < n ? lr : nlast) >> 3];
            }
#pragma unroll
            for (uint32_t c = 0; c < kB; ++c) {
                const uint32_t lr = c0 + 64 * c + lane;
                const bool v = lr < n && ((byte[c] >> (lr & 7)) & 1u);
                const uint64_t vm = __ballot(v);
                vi[c] = vbase + (uint32_t)__popcll(vm & lt);
                vbase += (uint32_t)__popcll(vm);
                if (v) vbits |= 1u << c;
            }
        }
        int32_t key[kB];
#pragma unroll
        for (uint32_t c = 0; c < kB; ++c) {
            // key pages, match words and output values stream through once:
            // non-temporal, so the L2 keeps the direct-map records
            key[c] = __builtin_nontemporal_load(&keys[vi[c] < nv ? vi[c] : vlast]);
            if (vi[c] >= nv) vbits &= ~(1u << c);
        }
        // matched rows come from the count pass's match bitmap (coalesced
        // words) instead of a second random presence lookup per row
        uint32_t mword[kB];
#pragma unroll
        for (uint32_t c = 0; c < kB; ++c) {
            const uint32_t lr = c0 + 64 * c + lane;
            mword[c] = __builtin_nontemporal_load(&mw[(msh + (lr < n ? lr : nlast)) >> 5]);
        }
        uint32_t mbits = 0;
        uint4 rec[kB][NQ];
#pragma unroll
        for (uint32_t c = 0; c < kB; ++c) {
            const uint32_t lr = c0 + 64 * c + lane;
            const int64_t off = (int64_t)key[c] - (int64_t)a.dt.min_key;
            const bool m = ((vbits >> c) & 1u) && ((mword[c] >> ((msh + lr) & 31)) & 1u) && off >= 0 &&
                           (uint64_t)off < a.dt.range;
            const uint32_t slot = m ? (uint32_t)off : 0u;
            if (m) mbits |= 1u << c;
            const uint4 *r = reinterpret_cast<const uint4 *>(a.dt.records + (uint64_t)slot * a.dt.stride);
#pragma unroll
            for (int q = 0; q < NQ; ++q) rec[c][q] = r[q];
        }
        // the columns' value words and validity out of the records, so the
        // records' registers are free before the emit loop
        uint32_t vlo[NC][kB], vhi[NC][kB], okb[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) okb[k] = 0;
#pragma unroll
        for (uint32_t c = 0; c < kB; ++c) {
            const uint32_t fl = rec_dword<NQ>(rec[c], fl_dw);
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                const uint32_t lo = rec_dword<NQ>(rec[c], pick[k].lo_dw);
                vlo[k][c] = lo;
                vhi[k][c] = pick[k].w8 ? rec_dword<NQ>(rec[c], pick[k].lo_dw + 1) : 0u;
                const bool ok = ((mbits >> c) & 1u) && ((fl >> pick[k].bit) & 1u) &&
                                !(pick[k].i32 && (int32_t)lo == kI32Null);  // page_builders.h:123
                okb[k] |= (uint32_t)ok << c;
            }
        }
#pragma unroll
        for (uint32_t c = 0; c < kB; ++c) {
            const bool matched = (mbits >> c) & 1u;
            const uint64_t mm = __ballot(matched);
            if (!mm) continue;
            const uint32_t rank = (uint32_t)__popcll(mm & lt);
            const uint32_t tm = (uint32_t)__popcll(mm);
            const uint32_t dst = matched ? rank : tm + (lane - rank);
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                if (k >= (int)nc) break;
                emit_col(pick[k].w8, base[k], vlo[k][c], vhi[k][c], (okb[k] >> c) & 1u, mm, matched, rank, tm, dst,
                         oc[k], lane, lt);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        if (k >= (int)nc) break;
        if (oc[k].pos) close_page(oc[k], base[k], oc[k].pos, oc[k].nn - oc[k].pb, lane);
    }
}

template <int NC>
static void launch_build_pages(const BuildPagesArgs &a, uint32_t c_lo, uint32_t grid, hipStream_t s) {
    switch (a.dt.stride / 16) {
        case 1: k_build_pages<1, NC><<<grid, kWG, 0, s>>>(a, c_lo); break;
        case 2: k_build_pages<2, NC><<<grid, kWG, 0, s>>>(a, c_lo); break;
        case 3: k_build_pages<3, NC><<<grid, kWG, 0, s>>>(a, c_lo); break;
        case 4: k_build_pages<4, NC><<<grid, kWG, 0, s>>>(a, c_lo); break;
        default: k_build_pages<5, NC><<<grid, kWG, 0, s>>>(a, c_lo); break;
    }
}

void build_pages(const BuildPagesArgs &a, hipStream_t s) {
    const uint32_t units = a.key.num_pages;
    if (!units || !a.ncols) return;
    if (a.dt.stride % 16 || a.dt.stride == 0 || a.dt.stride > 80) throw InvalidError("bad direct-map record stride");
    const uint32_t per = kWG / 64;
    const uint32_t grid = (units + per - 1) / per;
    for (uint32_t c_lo = 0; c_lo < a.ncols; c_lo += kBuildColsPerWave) {
        if (a.ncols - c_lo <= 2) launch_build_pages<2>(a, c_lo, grid, s);
        else launch_build_pages<kBuildColsPerWave>(a, c_lo, grid, s);
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

// Set bits of m below this lane.
__device__ __forceinline__ uint32_t rank_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// One wave per payload page of at most kWaveMaxRows rows, no barriers.  The
// page's validity words and match words are loaded once, lane j holding
// words j and 64 + j (a validity word from two aligned loads of the
// byte-aligned page bitmap, a match word from two words of the global
// bitmap: funnel shifts).  A 64-row chunk then takes its two words of each
// from those lanes into scalar registers (readlane): the value index of a
// row is the chunk's value base plus the valid rows below it (mbcnt), the
// chunk's value loads are issued for a batch of 16 (8-byte values: 8)
// chunks, and the matched values are stored compacted.  The output validity
// bits are the matched rows' bits in order: all ones up to the match count
// but for matched rows without a value (NULLs), whose ranks a scalar loop
// clears; they are assembled in registers (lane j holds output words j and
// 64 + j).  Vector work per chunk is a few lane ops; the bitmaps' bit
// arithmetic runs on the scalar unit.
template <int W>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(8))) void k_probe_filter_wave(FilterArgs a) {
    using T = typename std::conditional<W == 4, uint32_t, uint64_t>::type;
    // 8-byte values: half the batch, so that 8 waves fit a SIMD
    constexpr uint32_t kB = W == 4 ? kWaveBatch : kWaveBatch / 2;
    static_assert(32 % kB == 0, "a batch's chunks take their words from one register of the staged pair");
    const uint32_t lane = lane_id();
    const uint32_t q = blockIdx.x * (kWG / 64) + wave_id();
    if (q >= a.num_pages) return;
    const uint64_t o = a.pos ? a.pos[q] : q;
    if (a.pos && a.pos[q + 1] == o) return;  // no matched row
    const uint8_t *p = a.pages + (size_t)q * kPage;
    const uint32_t n = ld16(p), nv = ld16(p + 2);
    if (n > kWaveMaxRows) return;  // k_probe_filter's page
    const uint64_t rs = a.row_start[q];
    const T *vals = reinterpret_cast<const T *>(p + (W == 4 ? 4 : 8));
    uint8_t *out = a.out + o * kPage;
    T *ovals = reinterpret_cast<T *>(out + (W == 4 ? 4 : 8));
    // staged words: rows 32 w .. 32 w + 31 of the page, w = lane + 64 h (bits
    // of rows past n cleared; loads clamped into the page and the bitmap)
    const uint32_t bm0 = kPage - ((n + 7) >> 3);  // page offset of the bitmap
    const uint32_t *mw = a.match_bits + (rs >> 5);
    const uint32_t msh = (uint32_t)(rs & 31);
    uint32_t vword[2], mword[2];
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t w = lane + 64 * h, row0 = 32 * w;
        const uint32_t keep = row0 >= n ? 0u : (n - row0 >= 32 ? ~0u : (1u << (n - row0)) - 1u);
        const uint32_t wc = row0 < n ? w : 0u;
        const uint32_t off = bm0 + 4 * wc;
        const uint32_t a0 = (off & ~3u) < kPage - 4 ? (off & ~3u) : kPage - 4;
        const uint32_t a1 = a0 + 4 < kPage ? a0 + 4 : a0;  // bytes past the page end are rows past n
        const uint32_t d0 = *reinterpret_cast<const uint32_t *>(p + a0);
        const uint32_t d1 = *reinterpret_cast<const uint32_t *>(p + a1);
        const uint32_t m0 = mw[wc], m1 = mw[wc + 1];
        vword[h] = (n == nv ? ~0u : __builtin_amdgcn_alignbyte(d1, d0, off & 3u)) & keep;
        mword[h] = __builtin_amdgcn_alignbit(m1, m0, msh) & keep;
    }
    // lane j: words j and 64 + j of the output validity bitmap (a page of
    // this kernel has at most 4096 rows): all ones up to the result's row
    // count (cut at the end) but for the result rows of matched rows
    // without a value, cleared as they are found
    uint32_t bw = ~0u, bw1 = ~0u;
    uint32_t vbase = 0, mbase = 0, obase = 0;
    const uint32_t vlast = nv ? nv - 1 : 0, nch = (n + 63) >> 6;
    for (uint32_t c0 = 0; c0 < nch; c0 += kB) {
        // chunk c0 + c: words 2 (c0 + c) and 2 (c0 + c) + 1, all in register
        // c0 / 32 of the staged pair (chunks past the page's end read
        // cleared words)
        const uint32_t vsrc = c0 >= 32 ? vword[1] : vword[0], msrc = c0 >= 32 ? mword[1] : mword[0];
        T v[kB];
        uint32_t vi[kB];
#pragma unroll
        for (uint32_t c = 0; c < kB; ++c) {
            const uint32_t l = (2 * (c0 + c)) & 63;
            const uint32_t vlo = (uint32_t)__builtin_amdgcn_readlane(vsrc, l);
            const uint32_t vhi = (uint32_t)__builtin_amdgcn_readlane(vsrc, l + 1);
            vi[c] = vbase + __builtin_amdgcn_mbcnt_hi(vhi, __builtin_amdgcn_mbcnt_lo(vlo, 0u));
            v[c] = vals[vi[c] < nv ? vi[c] : vlast];
            vbase += (uint32_t)__popc(vlo) + (uint32_t)__popc(vhi);
        }
        uint32_t lb = (2 * c0) & 63;
#pragma unroll
        for (uint32_t c = 0; c < kB; ++c) {
            // (an opaque lane base: the chunk's word reads stay here instead
            // of all being hoisted into scalar registers)
            asm volatile("" : "+s"(lb));
            const uint32_t l = lb + 2 * c;
            const uint64_t vm = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(vsrc, l + 1) << 32) |
                                (uint32_t)__builtin_amdgcn_readlane(vsrc, l);
            const uint64_t mc = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(msrc, l + 1) << 32) |
                                (uint32_t)__builtin_amdgcn_readlane(msrc, l);
            // matched and valid (the lane masks are the scalar words), with
            // the value index inside the page's values
            bool ok = __builtin_amdgcn_inverse_ballot_w64(vm & mc) && vi[c] < nv;
            if (W == 4 && a.is_int32 && (int32_t)(uint32_t)v[c] == kI32Null) ok = false;  // page_builders.h:123
            const uint64_t om = __ballot(ok);
            if (ok) ovals[obase + rank_below(om)] = v[c];
            // the chunk's result rows are its matched rows in row order; one
            // without a value (a hole) leaves its bit clear
            for (uint64_t holes = mc & ~om; holes; holes &= holes - 1) {
                const uint32_t pos = mbase + (uint32_t)__popcll(mc & ((1ull << __builtin_ctzll(holes)) - 1ull));
                const uint32_t wd = pos >> 5, bit = 1u << (pos & 31);
                if (lane == (wd & 63)) {
                    if (wd < 64) bw &= ~bit;
                    else bw1 &= ~bit;
                }
            }
            obase += (uint32_t)__popcll(om);
            mbase += (uint32_t)__popcll(mc);
        }
    }
    // bits past the result's rows cleared
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t r0w = 32 * (lane + 64 * h);
        const uint32_t keep = r0w >= mbase ? 0u : (mbase - r0w >= 32 ? ~0u : (1u << (mbase - r0w)) - 1u);
        if (h) bw1 &= keep;
        else bw &= keep;
    }
    if (mbase == 0) {  // (fixed slots only) the slot stays unwritten
        if (lane == 0) atomicAdd(reinterpret_cast<unsigned long long *>(a.empty_pages), 1ull);
        return;
    }
    const uint32_t nbo = (mbase + 7) >> 3;
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t j = lane + 64 * h;
        const uint32_t word = h ? bw1 : bw;
        if (4 * j < nbo) {
            uint8_t *dst = out + kPage - nbo + 4 * j;
            const uint32_t k = nbo - 4 * j;
            dst[0] = (uint8_t)word;
            if (k > 1) dst[1] = (uint8_t)(word >> 8);
            if (k > 2) dst[2] = (uint8_t)(word >> 16);
            if (k > 3) dst[3] = (uint8_t)(word >> 24);
        }
    }
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
    const uint64_t out_pg = a.pos ? a.pos[q] : q;
    if (a.pos && a.pos[q + 1] == out_pg) return;  // no matched row
    if (a.row_start[q + 1] - a.row_start[q] <= kWaveMaxRows) return;  // k_probe_filter_wave's page
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
    if (run_m == 0) {  // (fixed slots only) the slot stays unwritten
        if (threadIdx.x == 0) atomicAdd(reinterpret_cast<unsigned long long *>(a.empty_pages), 1ull);
        return;
    }
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

__global__ __launch_bounds__(kWG) void k_compact_pages(const uint8_t *src, uint32_t num_pages, const uint64_t *pos,
                                                       uint8_t *dst) {
    const uint32_t q = blockIdx.x;
    if (q >= num_pages || pos[q + 1] == pos[q]) return;
    const uint4 *s = reinterpret_cast<const uint4 *>(src + (size_t)q * kPage);
    uint4 *d = reinterpret_cast<uint4 *>(dst + pos[q] * kPage);
    for (uint32_t i = threadIdx.x; i < kPage / 16; i += kWG) d[i] = s[i];
}

void compact_pages(const uint8_t *src, uint32_t num_pages, const uint64_t *pos, uint8_t *dst, hipStream_t s) {
    if (!num_pages) return;
    k_compact_pages<<<num_pages, kWG, 0, s>>>(src, num_pages, pos, dst);
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
