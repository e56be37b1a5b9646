// probe8.hpp -- device helpers of the probe and root-materialization kernels:
// 2048-row tiles with 8 rows per thread, ranks from wave ballots, probe-key
// lookups against the three build structures, and payload page windows.
#pragma once
#include "bloom.hpp"
#include "cj_internal.hpp"
#include "device_util.hpp"
#include "fused.hpp"

namespace cj {

constexpr uint32_t kRPT = 8;           // rows per thread
constexpr uint32_t kTile = kWG * kRPT;  // rows per workgroup tile

// A tile gives thread t the rows t, t + 256, ..., t + 256 (S - 1), so every
// load or store instruction of a wave covers 64 consecutive rows.  rank_slots
// ranks S predicates per thread in row order (S wave ballots, one LDS
// exchange of the 4 S wave counts); `lds32` holds 32 u32; all threads must
// call it.  mask[j] is the wave's ballot of slot j.  Returns the total.
template <uint32_t S>
__device__ __forceinline__ uint32_t rank_slots(uint32_t pred, uint32_t rank[S], uint64_t mask[S], uint32_t *lds32) {
    const uint32_t w = wave_id(), lane = lane_id();
#pragma unroll
    for (uint32_t j = 0; j < S; ++j) mask[j] = __ballot((pred >> j) & 1u);
    if (lane == 0) {
#pragma unroll
        for (uint32_t j = 0; j < S; ++j) lds32[4 * j + w] = (uint32_t)__popcll(mask[j]);
    }
    __syncthreads();
    uint32_t run = 0, base[S];
#pragma unroll
    for (uint32_t j = 0; j < S; ++j) {
#pragma unroll
        for (uint32_t ww = 0; ww < 4; ++ww) {
            if (ww == w) base[j] = run;
            run += lds32[4 * j + ww];
        }
    }
    __syncthreads();
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (uint32_t j = 0; j < S; ++j) rank[j] = base[j] + (uint32_t)__popcll(mask[j] & lt);
    return run;
}

// Keys of this thread's 8 rows (t0 + 256 j + t) of a key page.  Bit j of
// the returned mask: the row exists and its key is not NULL.  Loads are
// unconditional with clamped in-page addresses (no branch around a load), so
// all 8 are in flight together.  P: a generic or a global (glb) page pointer.
template <class P>
__device__ __forceinline__ uint32_t load_keys8(P p, uint32_t n, uint32_t nv, bool dense, uint32_t t0,
                                               uint32_t *running, int32_t key[kRPT], uint32_t *lds32,
                                               uint32_t *vi_out = nullptr) {
    const uint32_t nlast = n ? n - 1 : 0, vlast = nv ? nv - 1 : 0;
    uint32_t valid = 0;
    uint32_t vi[kRPT];
    if (dense) {
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j) {
            vi[j] = t0 + kWG * j + threadIdx.x;
            if (vi[j] < n) valid |= 1u << j;
        }
    } else {
        const P bm = p + kPage - ((n + 7) >> 3);
        uint32_t byte[kRPT];
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j) {
            const uint32_t r = t0 + kWG * j + threadIdx.x;
            byte[j] = bm[(r < n ? r : nlast) >> 3];
        }
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j) {
            const uint32_t r = t0 + kWG * j + threadIdx.x;
            if (r < n && ((byte[j] >> (r & 7)) & 1u)) valid |= 1u << j;
        }
        uint64_t mask[kRPT];
        const uint32_t tot = rank_slots<kRPT>(valid, vi, mask, lds32);
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j) vi[j] += *running;
        *running += tot;
    }
#pragma unroll
    for (uint32_t j = 0; j < kRPT; ++j) {
        key[j] = ld32(p + 4 + 4 * (size_t)(vi[j] < nv ? vi[j] : vlast));
        if (vi[j] >= nv) valid &= ~(1u << j);  // malformed bitmap: stay inside the value area
    }
    if (vi_out) {  // the value index of each key (its position among the page's non-NULL values)
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j) vi_out[j] = vi[j];
    }
    return valid;
}

// Keys of this thread's 8 rows (t0 + 256 j + t) of rows [a, a + n) of an
// intermediate (n >= 1): row ids, validity words and values each loaded in
// one batch.  Through an intermediate a NULL key and INT32_MIN are alike
// (the reference's value_t), so both leave the row out of the mask.
__device__ __forceinline__ uint32_t load_keys8_inter(const KeySource &k, uint64_t a, uint32_t n, uint32_t t0,
                                                     int32_t key[kRPT]) {
    uint64_t rid[kRPT];
    uint32_t vw[kRPT];
#pragma unroll
    for (uint32_t j = 0; j < kRPT; ++j) {
        const uint32_t r = t0 + kWG * j + threadIdx.x;
        const uint64_t row = a + (r < n ? r : n - 1);
        rid[j] = k.rids ? k.rids[row] : row;
    }
    if (k.values) {
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j) vw[j] = k.valid_bits[rid[j] >> 5];
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j) key[j] = k.values[rid[j]];
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j) {
            vw[j] = ~0u;
            key[j] = dense_page_value(k.dense_pages, k.dense_rows, (uint32_t)rid[j]);
        }
    }
    uint32_t valid = 0;
#pragma unroll
    for (uint32_t j = 0; j < kRPT; ++j) {
        const uint32_t r = t0 + kWG * j + threadIdx.x;
        const bool ok = (r < n) & (((vw[j] >> (rid[j] & 31)) & 1u) != 0) & (key[j] != kI32Null);
        valid |= (uint32_t)ok << j;
    }
    return valid;
}

// Candidates [from, to) of one key per lane that equal its key.  A short
// range (at most kShortRest) is walked by its own lane; longer ones by the
// whole wave, one lane's range at a time, 64 candidates a step: a skewed
// build side gives a few keys thousands of candidates, which one lane would
// walk alone while the rest of the machine waits.  Called by every lane.
constexpr uint32_t kShortRest = 8;
__device__ __forceinline__ uint32_t count_rest(const HashTableDev &ht, uint32_t from, uint32_t to, int32_t key) {
    const uint32_t lane = lane_id();
    uint32_t c = 0;
    if (to <= from + kShortRest)
        for (uint32_t i = from; i < to; ++i) c += ht.keys[i] == key;
    uint64_t pend = __ballot(to > from + kShortRest);
    while (pend) {
        const int l = (int)__builtin_ctzll(pend);
        const uint32_t lf = (uint32_t)__shfl((int)from, l, 64), lt_ = (uint32_t)__shfl((int)to, l, 64);
        const int32_t lk = __shfl(key, l, 64);
        uint32_t cnt = 0;
        for (uint32_t i0 = lf; i0 < lt_; i0 += 64) {
            const uint32_t i = i0 + lane;
            const bool eq = (i < lt_) & (ht.keys[i < lt_ ? i : lt_ - 1] == lk);
            cnt += (uint32_t)__popcll(__ballot(eq));
        }
        if (lane == (uint32_t)l) c += cnt;
        pend &= pend - 1;
    }
    return c;
}

// Lookups of 8 keys, all loads issued before any is used.  m[j] = matches of
// key j; hash mode: [s[j], e[j]) is its directory range and bits 0 / 1 of
// hit[j] say whether candidates s[j] / s[j] + 1 hold the key (the first two
// candidates of every key are read in one batch; the rest of a longer range
// is counted by the whole wave, see count_rest); direct mode: s[j] is its
// record slot (0 when it has none).  Called by every lane of the wave.
template <int MODE>
__device__ __forceinline__ void lookup8(uint32_t valid, const int32_t key[kRPT], const HashTableDev &ht,
                                        const NLKeys &nl, const DirectTable &dt, uint32_t m[kRPT], uint32_t s[kRPT],
                                        uint32_t e[kRPT], uint32_t hit[kRPT]) {
    if (MODE == kModeDirect) {
        uint32_t w[kRPT];
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j) {
            const int64_t off = (int64_t)key[j] - (int64_t)dt.min_key;
            const bool in = ((valid >> j) & 1u) & (off >= 0) & ((uint64_t)off < dt.range);
            s[j] = in ? (uint32_t)off : 0u;
            e[j] = in;
        }
        // the presence loads in one batch (a load chosen per key by a
        // condition is issued and awaited one at a time)
        if (dt.full) {
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) w[j] = ~0u;
        } else {
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) w[j] = dt.presence[s[j] >> 5];
        }
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j) {
            m[j] = e[j] ? (w[j] >> (s[j] & 31)) & 1u : 0u;
            e[j] = 0;
        }
        return;
    }
    if (MODE == kModeNL) {
#pragma unroll
        for (uint32_t j = 0; j < kRPT; ++j) {
            m[j] = 0;
            s[j] = e[j] = 0;
            if ((valid >> j) & 1u)
                for (int i = 0; i < nl.count; ++i) m[j] += nl.keys[i] == key[j];
        }
        return;
    }
    uint64_t slot[kRPT], d[kRPT], dp[kRPT];
    uint32_t tag[kRPT];
#pragma unroll
    for (uint32_t j = 0; j < kRPT; ++j) {
        uint64_t h = hash_key(key[j]);
        slot[j] = h >> (64 - ht.log_slots);
        tag[j] = c_bloom.t[h & 0x7FF];
        d[j] = ht.dir[slot[j]];
        dp[j] = ht.dir[slot[j] ? slot[j] - 1 : 0];
    }
#pragma unroll
    for (uint32_t j = 0; j < kRPT; ++j) {
        const bool pass = ((valid >> j) & 1u) & ((d[j] & tag[j]) == tag[j]);
        e[j] = pass ? (uint32_t)(d[j] >> 16) : 0;
        s[j] = pass && slot[j] ? (uint32_t)(dp[j] >> 16) : 0;
        if (s[j] >= e[j]) s[j] = e[j] = 0;
    }
    // first two candidates (index 0 is always a tuple when the table is not
    // empty; the second index is clamped to the first)
    int32_t k0[kRPT], k1[kRPT];
#pragma unroll
    for (uint32_t j = 0; j < kRPT; ++j) {
        k0[j] = ht.keys[s[j]];
        k1[j] = ht.keys[s[j] + 1 < e[j] ? s[j] + 1 : s[j]];
    }
#pragma unroll
    for (uint32_t j = 0; j < kRPT; ++j) {
        hit[j] = (uint32_t)((s[j] < e[j]) & (k0[j] == key[j])) | ((uint32_t)((s[j] + 1 < e[j]) & (k1[j] == key[j])) << 1);
        m[j] = (hit[j] & 1u) + (hit[j] >> 1) + count_rest(ht, s[j] + 2, e[j], key[j]);
    }
}

// ---- payload page windows -------------------------------------------------
constexpr uint32_t kWin = 32;  // payload pages a window may hold

struct WinState {
    uint64_t rs[kWin + 1];  // row starts of the window pages (+ end)
    uint16_t nrows[kWin];   // page headers of the window pages
    uint16_t nvals[kWin];
    uint32_t first;         // first page of the window
    uint32_t count;         // pages in the window (0: window overflowed)
};

// Wave-cooperative set-up of the window of pages holding rows [r0, r1]
// starting at page `first` (the page holding r0).  Called by one wave.
__device__ __forceinline__ void window_setup(WinState &win, const uint8_t *pages, uint32_t num_pages,
                                             const uint64_t *row_start, uint32_t first, uint64_t r_last) {
    const uint32_t lane = lane_id(), p = first + lane;
    const uint64_t rs = (lane <= kWin && p <= num_pages) ? row_start[p] : ~0ull;
    const bool in = lane <= kWin && p < num_pages && rs <= r_last;  // pages are in row order
    const uint64_t bal = __ballot(in);
    const uint32_t cw = (uint32_t)__popcll(bal & 0xffffffffull);
    const bool overflow = (bal >> kWin) & 1ull;
    if (lane <= cw) win.rs[lane] = rs;
    if (lane < cw) {
        const uint8_t *wp = pages + (size_t)p * kPage;
        win.nrows[lane] = ld16(wp);
        win.nvals[lane] = ld16(wp + 2);
    }
    if (lane == 0) {
        win.first = first;
        win.count = overflow ? 0 : cw;
    }
}

// Page of `row` (window first, else a binary search), its first row and header.
__device__ __forceinline__ uint32_t find_page(const uint8_t *pages, uint32_t num_pages, const uint64_t *row_start,
                                              const WinState &w, uint64_t row, uint64_t *rs, uint32_t *n,
                                              uint32_t *nv) {
    if (w.count) {
        uint32_t k = 0;
        while (k + 1 < w.count && w.rs[k + 1] <= row) ++k;
        *rs = w.rs[k];
        *n = w.nrows[k];
        *nv = w.nvals[k];
        return w.first + k;
    }
    uint32_t lo = 0, hi = num_pages;  // upper_bound(row_start, row) - 1
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (row_start[mid + 1] <= row) lo = mid + 1;
        else hi = mid;
    }
    uint32_t pg = lo < num_pages ? lo : num_pages - 1;
    const uint8_t *p = pages + (size_t)pg * kPage;
    *rs = row_start[pg];
    *n = ld16(p);
    *nv = ld16(p + 2);
    return pg;
}

}  // namespace cj
