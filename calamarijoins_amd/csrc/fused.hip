// fused.hip -- root-join fast path for fixed-width outputs with a base-table
// probe side: page index, counting pass, pair writer and multi-column page
// writers.
//
// The reference collects (build_rid, probe_rid) pairs into per-thread chunk
// lists (match_collector.h:66-221), then each thread builds pages of each
// output column from its own list (materialize.h:85-186) with random page
// reads through ColumnarReader (columnar_reader.h:300-474).  Here:
//   1. k_probe_count8: matches per probe key page (one lookup per row);
//   2. k_probe_write8: after an exclusive scan of those counts, the
//      (build id, probe row) of every match, in probe-row order;
//   3. k_mat_block: one workgroup per block of 1984 result rows and per side
//      writes the block's pages of every output column of that side:
//      build-side values come from the dense decoded build columns (or the
//      direct-map records), probe-side values straight from the probe
//      table's pages.  Each id and each record is read once per side.
// Block k holds result rows [1984 k, 1984 (k + 1)): INT32 page k (1984 rows,
// the most an all-non-NULL INT32 page holds) and 8-byte pages 2k and 2k + 1
// (992 rows each).
//
// gfx950 mapping: 2048-row tiles, thread t owning rows t + 256 j, so each
// wave instruction loads or stores 64 consecutive rows; ranks come from wave
// ballots plus one LDS exchange.  Because probe rows arrive in ascending
// order, the rows of one output page live in a short window of payload
// pages: the window's row starts and page headers sit in LDS, and a value's
// slot in its page comes from the page index -- per bitmap byte, the
// non-NULL rows before it (the byte-granular form of the reference's
// PageIndex popcount prefixes, columnar_reader.h:160-190) -- plus one
// popcount.
#include "probe8.hpp"

namespace cj {

constexpr uint32_t kCountGrid = 4096;  // workgroups of the key-page passes (16 per CU)

// ---------------------------------------------------------------- page index
// byte_prefix[byte_start[pg] + b] = set bits of the page's bitmap before
// bitmap byte b.  One wave per page.
__global__ __launch_bounds__(kWG) void k_page_byte_prefix(const uint8_t *pages, uint32_t num_pages,
                                                          const uint64_t *byte_start, uint16_t *prefix) {
    const uint32_t pg = blockIdx.x * (kWG / 64) + wave_id();
    if (pg >= num_pages) return;
    const uint8_t *p = pages + (size_t)pg * kPage;
    const uint32_t n = ld16(p);
    if (n >= 0xfffeu) return;
    const uint32_t nb = (n + 7) >> 3;
    const uint8_t *bm = p + kPage - nb;
    uint16_t *out = prefix + byte_start[pg];
    uint32_t running = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += 64) {
        const uint32_t b = b0 + lane_id();
        const uint32_t c = b < nb ? (uint32_t)__popc(bm[b]) : 0u;
        const uint32_t inc = wave_incl_scan<uint32_t>(c);
        if (b < nb) out[b] = (uint16_t)(running + inc - c);
        running += __shfl(inc, 63, 64);
    }
}

void page_byte_prefix(const uint8_t *pages, uint32_t num_pages, const uint64_t *byte_start, uint16_t *prefix,
                      hipStream_t s) {
    if (!num_pages) return;
    const uint32_t per = kWG / 64;
    k_page_byte_prefix<<<(num_pages + per - 1) / per, kWG, 0, s>>>(pages, num_pages, byte_start, prefix);
    CJ_HIP(hipGetLastError());
}

// ---------------------------------------------------------- page windows
__global__ void k_window_first(const uint32_t *probe_rows, uint64_t rows, uint32_t R, const uint64_t *crs,
                               uint32_t col_pages, uint32_t *first) {
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k * R >= rows) return;
    const uint64_t r0 = probe_rows[k * R];
    uint32_t lo = 0, hi = col_pages;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (crs[mid + 1] <= r0) lo = mid + 1;
        else hi = mid;
    }
    first[k] = lo < col_pages ? lo : col_pages - 1;
}

void window_first(const uint32_t *probe_rows, uint64_t rows, uint32_t R, const uint64_t *col_row_start,
                  uint32_t col_pages, uint32_t *first, hipStream_t s) {
    if (!rows || !col_pages) return;
    const uint64_t pages = (rows + R - 1) / R;
    k_window_first<<<(unsigned)((pages + kWG - 1) / kWG), kWG, 0, s>>>(probe_rows, rows, R, col_row_start, col_pages,
                                                                      first);
    CJ_HIP(hipGetLastError());
}

// --------------------------------------------------------- counting pass
// Work unit pg of the probe side: a key page (columnar), or rows
// [pg kInterChunk, ...) of an intermediate's row-id list (INTER).
struct ProbeUnit {
    const uint8_t *p;
    uint32_t n, nv;
    bool dense;
    uint64_t r0;  // probe row of the unit's first row
};
template <bool INTER>
__device__ __forceinline__ ProbeUnit probe_unit(const KeySource &k, uint32_t pg) {
    ProbeUnit u{};
    if (INTER) {
        u.r0 = (uint64_t)pg * kInterChunk;
        u.n = (uint32_t)(k.rows - u.r0 < kInterChunk ? k.rows - u.r0 : kInterChunk);
        return u;
    }
    u.p = k.pages + (size_t)pg * kPage;
    u.n = ld16(u.p);
    u.nv = ld16(u.p + 2);
    u.dense = u.n == u.nv;
    u.r0 = k.row_start[pg];
    return u;
}
template <bool INTER>
__device__ __forceinline__ uint32_t unit_keys8(const KeySource &k, const ProbeUnit &u, uint32_t t0, uint32_t *running,
                                               int32_t key[kRPT], uint32_t *lds32) {
    if (INTER) return load_keys8_inter(k, u.r0, u.n, t0, key);
    return load_keys8(u.p, u.n, u.nv, u.dense, t0, running, key, lds32);
}

template <int MODE, bool INTER>
__global__ __launch_bounds__(kWG) void k_probe_count8(ProbeArgs a, uint64_t *counts) {
    __shared__ uint32_t lds32[32];
    __shared__ uint64_t red[4];
    NLKeys nl;
    if (MODE == kModeNL) nl = *a.nl;
    const KeySource &k = a.key;
    const uint32_t units = key_source_units(k);
    // a workgroup handles several units (grid-stride), so short pages do not
    // leave the machine waiting on workgroup dispatch
    for (uint32_t pg = blockIdx.x; pg < units; pg += gridDim.x) {
        const ProbeUnit u = probe_unit<INTER>(k, pg);
        const uint32_t n = u.n;
        uint32_t running = 0;
        uint64_t mine = 0;
        for (uint32_t t0 = 0; t0 < n; t0 += kTile) {
            int32_t key[kRPT];
            uint32_t valid = unit_keys8<INTER>(k, u, t0, &running, key, lds32);
            uint32_t m[kRPT], s[kRPT], e[kRPT], hit[kRPT];
            lookup8<MODE>(valid, key, a.ht, nl, a.dt, m, s, e, hit);
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) mine += m[j];
        }
        uint64_t tot;
        block_excl_scan<uint64_t>(mine, &tot, red);
        if (threadIdx.x == 0) counts[pg] = tot;
    }
}

template <bool INTER>
static void launch_count8(const ProbeArgs &a, uint64_t *counts, uint32_t grid, hipStream_t s) {
    if (a.mode == kModeNL) k_probe_count8<kModeNL, INTER><<<grid, kWG, 0, s>>>(a, counts);
    else if (a.mode == kModeDirect) k_probe_count8<kModeDirect, INTER><<<grid, kWG, 0, s>>>(a, counts);
    else k_probe_count8<kModeHash, INTER><<<grid, kWG, 0, s>>>(a, counts);
}

void probe_count_columnar(const ProbeArgs &a, uint64_t *counts, hipStream_t s) {
    const uint32_t units = key_source_units(a.key);
    if (!units) return;
    const uint32_t grid = units < kCountGrid ? units : kCountGrid;
    if (a.key.mode == KEY_INTER) launch_count8<true>(a, counts, grid, s);
    else launch_count8<false>(a, counts, grid, s);
    CJ_HIP(hipGetLastError());
}

// ------------------------------------------------------------ pair writer
// Exclusive prefix, in row order (slot-major: row t0 + 256 j + t), of the
// 8 per-row match counts of every thread.  `lds32` holds 32 u32.
__device__ __forceinline__ uint32_t scan_slots8(const uint32_t cnt[kRPT], uint32_t pre[kRPT], uint32_t *lds32) {
    const uint32_t w = wave_id(), lane = lane_id();
    uint32_t inc[kRPT];
#pragma unroll
    for (uint32_t j = 0; j < kRPT; ++j) {
        inc[j] = wave_incl_scan<uint32_t>(cnt[j]);
        if (lane == 63) lds32[4 * j + w] = inc[j];
    }
    __syncthreads();
    uint32_t run = 0;
#pragma unroll
    for (uint32_t j = 0; j < kRPT; ++j) {
#pragma unroll
        for (uint32_t ww = 0; ww < 4; ++ww) {
            if (ww == w) pre[j] = run + inc[j] - cnt[j];
            run += lds32[4 * j + ww];
        }
    }
    __syncthreads();
    return run;
}

// Pairs of the candidates [from, to) of one key per lane that equal its key,
// written at out onwards in candidate order: by its own lane for a short
// range, else by the whole wave, one lane's range at a time (see
// count_rest).  Called by every lane.
__device__ __forceinline__ void write_rest(const HashTableDev &ht, uint32_t from, uint32_t to, int32_t key,
                                           uint64_t out, uint64_t end, uint32_t prid, uint32_t *ob, uint32_t *op) {
    const uint32_t lane = lane_id();
    const uint64_t lt = lanemask_lt();
    if (to <= from + kShortRest) {
        uint64_t o = out;
        for (uint32_t i = from; i < to; ++i)
            if (ht.keys[i] == key) {
                if (o < end) {
                    if (ob) ob[o] = ht.rids[i];
                    if (op) op[o] = prid;
                }
                ++o;
            }
    }
    uint64_t pend = __ballot(to > from + kShortRest);
    while (pend) {
        const int l = (int)__builtin_ctzll(pend);
        const uint32_t lf = (uint32_t)__shfl((int)from, l, 64), lt_ = (uint32_t)__shfl((int)to, l, 64);
        const int32_t lk = __shfl(key, l, 64);
        const uint32_t lp = (uint32_t)__shfl((int)prid, l, 64);
        uint64_t lo = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(out >> 32), l, 64) << 32) |
                      (uint32_t)__shfl((int)(uint32_t)out, l, 64);
        for (uint32_t i0 = lf; i0 < lt_; i0 += 64) {
            const uint32_t i = i0 + lane;
            const uint32_t ic = i < lt_ ? i : lt_ - 1;
            const bool eq = (i < lt_) & (ht.keys[ic] == lk);
            const uint64_t em = __ballot(eq);
            if (eq) {
                const uint64_t at = lo + (uint64_t)__popcll(em & lt);
                if (at < end) {
                    if (ob) ob[at] = ht.rids[ic];
                    if (op) op[at] = lp;
                }
            }
            lo += (uint64_t)__popcll(em);
        }
        pend &= pend - 1;
    }
}

template <int MODE, bool INTER>
__global__ __launch_bounds__(kWG) void k_probe_write8(ProbeArgs a, const uint64_t *off, uint32_t *ob, uint32_t *op) {
    __shared__ uint32_t lds32[32];
    NLKeys nl;
    if (MODE == kModeNL) nl = *a.nl;
    const KeySource &k = a.key;
    const uint32_t units = key_source_units(k);
    for (uint32_t pg = blockIdx.x; pg < units; pg += gridDim.x) {
        uint64_t base = off[pg];
        const uint64_t end = off[pg + 1];
        if (base == end) continue;
        const ProbeUnit u = probe_unit<INTER>(k, pg);
        const uint32_t n = u.n;
        const uint64_t r0 = u.r0;
        uint32_t running = 0;
        for (uint32_t t0 = 0; t0 < n && base < end; t0 += kTile) {
            int32_t key[kRPT];
            uint32_t valid = unit_keys8<INTER>(k, u, t0, &running, key, lds32);
            uint32_t m[kRPT], s[kRPT], e[kRPT], pre[kRPT], hit[kRPT];
            lookup8<MODE>(valid, key, a.ht, nl, a.dt, m, s, e, hit);
            const uint32_t tot = scan_slots8(m, pre, lds32);
            // direct mode: the build id is the record slot, or the build row
            // id the record carries (all 8 loads issued together; slot 0 for
            // a key without a match)
            uint32_t bid[kRPT];
            if (MODE == kModeHash) {
                // row ids of the first candidates, in one batch (a match on
                // the second candidate is rarer and reads its own)
#pragma unroll
                for (uint32_t j = 0; j < kRPT; ++j) bid[j] = a.ht.rids[s[j]];
            }
            if (MODE == kModeDirect) {
#pragma unroll
                for (uint32_t j = 0; j < kRPT; ++j) bid[j] = s[j];
                if (a.dt.rid_offset >= 0) {
#pragma unroll
                    for (uint32_t j = 0; j < kRPT; ++j)
                        bid[j] = *reinterpret_cast<const uint32_t *>(a.dt.records + (uint64_t)s[j] * a.dt.stride +
                                                                     (uint32_t)a.dt.rid_offset);
                }
            }
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) {
                if (!m[j]) continue;
                const uint32_t prid = (uint32_t)(r0 + t0 + kWG * j + threadIdx.x);
                uint64_t o = base + pre[j];
                if (MODE == kModeDirect) {
                    if (o < end) {
                        if (ob) ob[o] = bid[j];
                        if (op) op[o] = prid;
                    }
                } else if (MODE == kModeNL) {
                    for (int i = 0; i < nl.count; ++i)
                        if (nl.keys[i] == key[j]) {
                            if (o < end) {
                                if (ob) ob[o] = nl.rids[i];
                                if (op) op[o] = prid;
                            }
                            ++o;
                        }
                } else {
                    if (hit[j] & 1u) {
                        if (o < end) {
                            if (ob) ob[o] = bid[j];
                            if (op) op[o] = prid;
                        }
                        ++o;
                    }
                    if (hit[j] & 2u) {
                        if (o < end) {
                            if (ob) ob[o] = a.ht.rids[s[j] + 1];
                            if (op) op[o] = prid;
                        }
                        ++o;
                    }
                }
            }
            // the rest of every longer candidate range, by the whole wave
            if (MODE == kModeHash) {
#pragma unroll
                for (uint32_t j = 0; j < kRPT; ++j)
                    write_rest(a.ht, s[j] + 2, e[j], key[j], base + pre[j] + (hit[j] & 1u) + (hit[j] >> 1), end,
                               (uint32_t)(r0 + t0 + kWG * j + threadIdx.x), ob, op);
            }
            base += tot;
        }
    }
}

template <bool INTER>
static void launch_write8(const ProbeArgs &a, const uint64_t *off, uint32_t *ob, uint32_t *op, uint32_t grid,
                          hipStream_t s) {
    if (a.mode == kModeNL) k_probe_write8<kModeNL, INTER><<<grid, kWG, 0, s>>>(a, off, ob, op);
    else if (a.mode == kModeDirect) k_probe_write8<kModeDirect, INTER><<<grid, kWG, 0, s>>>(a, off, ob, op);
    else k_probe_write8<kModeHash, INTER><<<grid, kWG, 0, s>>>(a, off, ob, op);
}

void probe_write_columnar(const ProbeArgs &a, const uint64_t *off, uint32_t *build_ids, uint32_t *probe_rows,
                          hipStream_t s) {
    const uint32_t units = key_source_units(a.key);
    if (!units) return;
    const uint32_t grid = units < kCountGrid ? units : kCountGrid;
    if (a.key.mode == KEY_INTER) launch_write8<true>(a, off, build_ids, probe_rows, grid, s);
    else launch_write8<false>(a, off, build_ids, probe_rows, grid, s);
    CJ_HIP(hipGetLastError());
}

// ----------------------------------------------------- output page writer
// Ranks of 8 predicates per thread in block-row order (row 256 j + t) and
// the number of true rows before kRowsPage8 (the first 8-byte page).
__device__ __forceinline__ uint32_t rank_block(uint32_t pred, uint32_t rank[kRPT], uint64_t mask[kRPT],
                                               uint32_t *lds32, uint32_t *before_split) {
    // row kRowsPage8 = 992 = slot 3, wave 3, lane 32
    constexpr uint32_t js = kRowsPage8 / kWG, ws = (kRowsPage8 % kWG) / 64, ls = kRowsPage8 % 64;
    const uint32_t w = wave_id(), lane = lane_id();
#pragma unroll
    for (uint32_t j = 0; j < kRPT; ++j) mask[j] = __ballot((pred >> j) & 1u);
    if (lane == 0) {
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j) lds32[4 * j + w] = (uint32_t)__popcll(mask[j]);
        if (w == ws) lds32[4 * kRPT] = (uint32_t)__popcll(mask[js] & ((1ull << ls) - 1ull));
    }
    __syncthreads();
    uint32_t run = 0, base[kRPT], split = 0;
#pragma unroll
    for (uint32_t j = 0; j < kRPT; ++j) {
#pragma unroll
        for (uint32_t ww = 0; ww < 4; ++ww) {
            if (ww == w) base[j] = run;
            if (j == js && ww == ws) split = run + lds32[4 * kRPT];
            run += lds32[4 * j + ww];
        }
    }
    __syncthreads();
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (uint32_t j = 0; j < kRPT; ++j) rank[j] = base[j] + (uint32_t)__popcll(mask[j] & lt);
    *before_split = split;
    return run;
}

// A workgroup writes, for `a.cols_per_wg` output columns of width W of one
// side, the pages of one block of kBlockRows result rows: one INT32 page
// (1984 rows) or two 8-byte pages (992 rows each).  Thread t owns block rows
// t + 256 j.  Workgroup i runs on XCD i % 8; the column groups of one block
// are given to consecutive workgroups of the same XCD so they share its L2
// for the block's ids.
template <int W, bool FROM_BUILD, bool DIRECT>
__global__ __launch_bounds__(kWG) void k_mat_block(MatArgs a) {
    __shared__ uint32_t lds32[4 * kRPT + 4];
    __shared__ WinState win[FROM_BUILD ? 1 : kMatMaxCols];
    const uint32_t groups = (a.ncols + a.cols_per_wg - 1) / a.cols_per_wg;
    const uint32_t xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
    const uint64_t b = (uint64_t)(slot / groups) * 8 + xcd;
    const uint32_t c0 = (slot % groups) * a.cols_per_wg;
    const uint32_t c1 = c0 + a.cols_per_wg < a.ncols ? c0 + a.cols_per_wg : a.ncols;
    const uint64_t r0 = b * kBlockRows;
    if (r0 >= a.rows) return;
    const uint32_t cnt = (uint32_t)(a.rows - r0 < kBlockRows ? a.rows - r0 : kBlockRows);
    const uint32_t cnt0 = cnt < kRowsPage8 ? cnt : kRowsPage8, cnt1 = cnt - cnt0;
    uint32_t id[kRPT];
#pragma unroll
    for (uint32_t j = 0; j < kRPT; ++j) {
        const uint32_t i = kWG * j + threadIdx.x;
        id[j] = i < cnt ? a.ids[r0 + i] : 0u;
    }
    if (!FROM_BUILD) {  // probe rows ascend: a page window per column, one wave each
        const uint64_t r_last = a.ids[r0 + cnt - 1];
        for (uint32_t c = c0 + wave_id(); c < c1; c += kWG / 64) {
            const MatCol &col = a.cols[c];
            window_setup(win[c - c0], col.pages, col.num_pages, col.row_start, col.win_first[b], r_last);
        }
        __syncthreads();
    }
    for (uint32_t c = c0; c < c1; ++c) {
        const MatCol &col = a.cols[c];
        constexpr bool w4 = W == 4;
        uint64_t v[kRPT];
        uint32_t ok = 0;
        if (FROM_BUILD && DIRECT) {  // ids are direct-map record slots
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) {
                v[j] = 0;
                if (kWG * j + threadIdx.x < cnt) {
                    const uint8_t *rec = a.dt.records + (uint64_t)id[j] * a.dt.stride;
                    const uint32_t fl = *reinterpret_cast<const uint32_t *>(rec + a.dt.flags_offset);
                    v[j] = w4 ? (uint64_t) * reinterpret_cast<const uint32_t *>(rec + col.rec_offset)
                              : *reinterpret_cast<const uint64_t *>(rec + col.rec_offset);
                    ok |= ((fl >> col.rec_bit) & 1u) << j;
                }
            }
        } else if (FROM_BUILD) {  // ids are build row ids (through a row-id chain)
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) {
                v[j] = 0;
                if (kWG * j + threadIdx.x < cnt) {
                    const uint32_t rid = col.rids ? col.rids[id[j]] : id[j];
                    v[j] = w4 ? (uint64_t) static_cast<const uint32_t *>(col.values)[rid]
                              : static_cast<const uint64_t *>(col.values)[rid];
                    ok |= (uint32_t)valid_bit(col.valid_bits, rid) << j;
                }
            }
        } else {  // ids are probe rows: read their pages
            uint32_t pgi[kRPT], lr[kRPT];
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) {
                pgi[j] = 0;
                lr[j] = 0;
                if (kWG * j + threadIdx.x < cnt) {
                    uint64_t rs;
                    uint32_t nn, nvv;
                    pgi[j] = find_page(col.pages, col.num_pages, col.row_start, win[c - c0], id[j], &rs, &nn, &nvv);
                    const uint8_t *pp = col.pages + (size_t)pgi[j] * kPage;
                    lr[j] = (uint32_t)(id[j] - rs);
                    if (lr[j] >= nn) {  // row outside its page: inconsistent row starts
                        lr[j] = 0;
                    } else if (nn != nvv) {  // sparse page: bitmap + page index
                        const uint32_t byte = pp[kPage - ((nn + 7) >> 3) + (lr[j] >> 3)];
                        const uint32_t pre = col.byte_prefix[col.byte_start[pgi[j]] + (lr[j] >> 3)];
                        if ((byte >> (lr[j] & 7)) & 1u) {
                            lr[j] = pre + (uint32_t)__popc(byte & ((1u << (lr[j] & 7)) - 1u));
                            ok |= (lr[j] < nvv ? 1u : 0u) << j;
                        }
                    } else {
                        ok |= 1u << j;
                    }
                }
            }
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) {
                v[j] = 0;
                if ((ok >> j) & 1u) {
                    const uint8_t *vals = col.pages + (size_t)pgi[j] * kPage + (w4 ? 4 : 8);
                    v[j] = w4 ? (uint64_t)(uint32_t)ld32(vals + 4 * (size_t)lr[j]) : ld64(vals + 8 * (size_t)lr[j]);
                }
            }
        }
        if (col.is_int32) {  // value_t NULL sentinel (page_builders.h:123)
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j)
                if ((int32_t)(uint32_t)v[j] == kI32Null) ok &= ~(1u << j);
        }
        uint32_t vi[kRPT], split;
        uint64_t mask[kRPT];
        const uint32_t tot = rank_block(ok, vi, mask, lds32, &split);
        const uint32_t lane = lane_id();
        if (w4) {  // W == 4
            uint8_t *page = col.out + b * kPage;
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j)
                if ((ok >> j) & 1u) *reinterpret_cast<uint32_t *>(page + 4 + 4 * (size_t)vi[j]) = (uint32_t)v[j];
            // bitmap: slot j of wave w covers rows 256 j + 64 w .. + 63 = 8 bytes
            const uint32_t nbytes = (cnt + 7) >> 3;
            uint8_t *bm = page + kPage - nbytes;
            if (lane < 8) {
#pragma unroll
                for (uint32_t j = 0; j < kRPT; ++j) {
                    const uint32_t by = ((kWG * j + 64 * wave_id()) >> 3) + lane;
                    if (by < nbytes) bm[by] = (uint8_t)(mask[j] >> (8 * lane));
                }
            }
            if (threadIdx.x == 0) {
                *reinterpret_cast<uint16_t *>(page) = (uint16_t)cnt;
                *reinterpret_cast<uint16_t *>(page + 2) = (uint16_t)tot;
            }
        } else {
            uint8_t *page0 = col.out + 2 * b * kPage, *page1 = page0 + kPage;
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j)
                if ((ok >> j) & 1u) {
                    const bool hi = kWG * j + threadIdx.x >= kRowsPage8;
                    uint8_t *dst = hi ? page1 + 8 + 8 * (size_t)(vi[j] - split) : page0 + 8 + 8 * (size_t)vi[j];
                    *reinterpret_cast<uint64_t *>(dst) = v[j];
                }
            const uint32_t nb0 = (cnt0 + 7) >> 3, nb1 = (cnt1 + 7) >> 3;
            constexpr uint32_t kSplitByte = kRowsPage8 / 8;
            if (lane < 8) {
#pragma unroll
                for (uint32_t j = 0; j < kRPT; ++j) {
                    const uint32_t by = ((kWG * j + 64 * wave_id()) >> 3) + lane;
                    const uint8_t bits = (uint8_t)(mask[j] >> (8 * lane));
                    if (by < kSplitByte) {
                        if (by < nb0) page0[kPage - nb0 + by] = bits;
                    } else if (by - kSplitByte < nb1) {
                        page1[kPage - nb1 + (by - kSplitByte)] = bits;
                    }
                }
            }
            if (threadIdx.x == 0) {
                *reinterpret_cast<uint16_t *>(page0) = (uint16_t)cnt0;
                *reinterpret_cast<uint16_t *>(page0 + 2) = (uint16_t)split;
                if (cnt1) {
                    *reinterpret_cast<uint16_t *>(page1) = (uint16_t)cnt1;
                    *reinterpret_cast<uint16_t *>(page1 + 2) = (uint16_t)(tot - split);
                }
            }
        }
    }
}

template <int W>
static void launch_block(const MatArgs &a, unsigned grid, hipStream_t s) {
    if (a.from_build && a.direct) k_mat_block<W, true, true><<<grid, kWG, 0, s>>>(a);
    else if (a.from_build) k_mat_block<W, true, false><<<grid, kWG, 0, s>>>(a);
    else k_mat_block<W, false, false><<<grid, kWG, 0, s>>>(a);
}

void materialize_side(const MatArgs &a, hipStream_t s) {
    if (!a.rows || !a.ncols) return;
    if (a.cols_per_wg < 1 || a.cols_per_wg > kMatMaxCols) throw InvalidError("bad columns per workgroup");
    const uint64_t blocks = (a.rows + kBlockRows - 1) / kBlockRows;
    const uint64_t groups = (a.ncols + a.cols_per_wg - 1) / a.cols_per_wg;
    const uint64_t grid = (blocks + 7) / 8 * 8 * groups;
    if (grid > 0x7fffffffull) throw UnsupportedError("result too large for one launch");
    if (a.cols[0].width == 4) launch_block<4>(a, (unsigned)grid, s);
    else launch_block<8>(a, (unsigned)grid, s);
    CJ_HIP(hipGetLastError());
}

}  // namespace cj
