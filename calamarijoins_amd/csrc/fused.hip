// fused.hip -- fused probe + materialization for a root join whose probe side
// is a base table and whose outputs are fixed-width columns.
//
// The reference collects (build_rid, probe_rid) pairs into per-thread chunk
// lists (match_collector.h:66-221) and then builds each output column's pages
// from them (materialize.h:85-186).  Here the pairs never reach HBM: after a
// counting pass (k_probe_count) gives each probe key page its number of
// matches, one workgroup per probe key page re-probes its rows, stages the
// pairs of one output batch (1984 rows) in LDS and writes that batch's
// output pages for every column directly: one INT32 page (1984 rows) and two
// INT64/FP64 pages (992 rows each) per batch, so every output page is owned
// by exactly one workgroup.  The workgroup's pages follow each other, and
// the workgroups' page ranges follow probe page order (exclusive scans of
// ceil(matches / 1984) and ceil(matches / 992)).
//
// Probe-side payloads are read straight from their pages: the workgroup's
// probe rows [r0, r1) map to a short window of payload pages (binary search
// once, window row starts kept in LDS), and a value's slot inside its page
// comes from the page index -- per 8 rows, the number of non-NULL rows
// before them in the page (the byte-granular form of the reference's
// PageIndex popcount prefixes, columnar_reader.h:160-190) -- plus one
// popcount.  Build-side payloads (the smaller side) are read from their
// dense decoded columns.
#include "cj_internal.hpp"
#include "device_util.hpp"
#include "bloom.hpp"
#include "fused.hpp"

namespace cj {

// ---------------------------------------------------------------- page index
// byte_prefix[byte_start[pg] + b] = set bits of the page's bitmap before
// bitmap byte b.  One workgroup per page.
__global__ __launch_bounds__(kWG) void k_page_byte_prefix(const uint8_t *pages, const uint64_t *byte_start,
                                                          uint16_t *prefix) {
    __shared__ uint32_t lds4[4];
    const uint32_t pg = blockIdx.x;
    const uint8_t *p = pages + (size_t)pg * kPage;
    const uint32_t n = ld16(p);
    if (n >= 0xfffeu) return;
    const uint32_t nb = (n + 7) >> 3;
    const uint8_t *bm = p + kPage - nb;
    uint16_t *out = prefix + byte_start[pg];
    uint32_t running = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += kWG) {
        uint32_t b = b0 + threadIdx.x;
        uint32_t c = b < nb ? (uint32_t)__popc(bm[b]) : 0u;
        uint32_t tot;
        uint32_t pre = block_excl_scan<uint32_t>(c, &tot, lds4);
        if (b < nb) out[b] = (uint16_t)(running + pre);
        running += tot;
    }
}

void page_byte_prefix(const uint8_t *pages, uint32_t num_pages, const uint64_t *byte_start, uint16_t *prefix,
                      hipStream_t s) {
    if (!num_pages) return;
    k_page_byte_prefix<<<num_pages, kWG, 0, s>>>(pages, byte_start, prefix);
    CJ_HIP(hipGetLastError());
}

// ------------------------------------------------------- batch page counts
__global__ void k_batch_pages(const uint64_t *cnt, uint32_t n, uint64_t *p32, uint64_t *p64) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t c = cnt[i];
    p32[i] = (c + kFusedBatch - 1) / kFusedBatch;
    p64[i] = (c + kFusedHalf - 1) / kFusedHalf;
}
void fused_batch_pages(const uint64_t *cnt, uint32_t n, uint64_t *p32, uint64_t *p64, hipStream_t s) {
    if (!n) return;
    k_batch_pages<<<(n + kWG - 1) / kWG, kWG, 0, s>>>(cnt, n, p32, p64);
    CJ_HIP(hipGetLastError());
}

// ----------------------------------------------------------- fused write
constexpr uint32_t kWin = 32;  // payload pages a workgroup's window may hold

// Matches of one probe key: their number, and for the hash table the
// directory range [s, e) that holds them (one lookup per probe row).
template <bool NL>
__device__ __forceinline__ uint32_t lookup(int32_t key, const HashTableDev &ht, const NLKeys &nl, uint64_t *s,
                                           uint64_t *e) {
    uint32_t m = 0;
    *s = *e = 0;
    if (NL) {
        for (int i = 0; i < nl.count; ++i) m += nl.keys[i] == key;
        return m;
    }
    uint64_t h = hash_key(key);
    uint64_t slot = h >> (64 - ht.log_slots);
    uint64_t d = ht.dir[slot];
    uint32_t tag = c_bloom.t[h & 0x7FF];
    if ((d & tag) != tag) return 0;
    uint64_t end = d >> 16, start = slot ? (ht.dir[slot - 1] >> 16) : 0;
    for (uint64_t i = start; i < end; ++i) m += ht.keys[i] == key;
    *s = start;
    *e = end;
    return m;
}

// Stage the key's matches whose per-row ordinal lies in [lo, hi) into the
// LDS pair buffer at (ordinal - lo + at).
template <bool NL>
__device__ __forceinline__ void emit_range(int32_t key, uint32_t prid, const HashTableDev &ht, const NLKeys &nl,
                                           uint64_t s, uint64_t e, uint32_t lo, uint32_t hi, uint32_t at,
                                           uint32_t *pb, uint32_t *pp) {
    uint32_t j = 0;
    if (NL) {
        for (int i = 0; i < nl.count && j < hi; ++i)
            if (nl.keys[i] == key) {
                if (j >= lo) {
                    pb[at + j - lo] = nl.rids[i];
                    pp[at + j - lo] = prid;
                }
                ++j;
            }
        return;
    }
    for (uint64_t i = s; i < e && j < hi; ++i)
        if (ht.keys[i] == key) {
            if (j >= lo) {
                pb[at + j - lo] = ht.rids[i];
                pp[at + j - lo] = prid;
            }
            ++j;
        }
}

struct WinState {
    uint64_t rs[kWin + 1];  // row starts of the window pages (+ end)
    uint32_t first;         // first page of the window
    uint32_t count;         // pages in the window (0: window overflowed)
};

// Page of probe row `row` in column `c`.
__device__ __forceinline__ uint32_t find_page(const FusedCol &c, const WinState &w, uint64_t row) {
    if (w.count) {
        uint32_t k = 0;
        while (k + 1 < w.count && w.rs[k + 1] <= row) ++k;
        return w.first + k;
    }
    uint32_t lo = 0, hi = c.num_pages;  // upper_bound(row_start, row) - 1
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (c.row_start[mid + 1] <= row) lo = mid + 1;
        else hi = mid;
    }
    return lo < c.num_pages ? lo : c.num_pages - 1;
}

template <int W>
__device__ __forceinline__ bool probe_value(const FusedCol &c, const WinState &w, uint64_t row, uint64_t *v) {
    uint32_t pg = find_page(c, w, row);
    const uint8_t *p = c.pages + (size_t)pg * kPage;
    uint32_t n = ld16(p), nv = ld16(p + 2);
    uint32_t lr = (uint32_t)(row - c.row_start[pg]);
    uint32_t vi = lr;
    if (n != nv) {
        const uint8_t *bm = p + kPage - ((n + 7) >> 3);
        uint32_t byte = bm[lr >> 3];
        if (!((byte >> (lr & 7)) & 1u)) return false;
        vi = c.byte_prefix[c.byte_start[pg] + (lr >> 3)] + (uint32_t)__popc(byte & ((1u << (lr & 7)) - 1u));
        if (vi >= nv) return false;
    }
    const uint8_t *vals = p + (W == 4 ? 4 : 8);
    *v = W == 4 ? (uint64_t)(uint32_t)ld32(vals + 4 * (size_t)vi) : ld64(vals + 8 * (size_t)vi);
    return true;
}

// Write rows [i0, i0 + cnt) of the staged batch as one output page of column c.
template <int W>
__device__ __forceinline__ void write_page(const FusedCol &c, const WinState &w, const uint32_t *pb, const uint32_t *pp,
                                           uint32_t i0, uint32_t cnt, uint8_t *page, uint32_t *lds4) {
    const uint32_t nbytes = (cnt + 7) >> 3;
    uint8_t *bm = page + kPage - nbytes;
    uint8_t *vals = page + (W == 4 ? 4 : 8);
    uint32_t running = 0;
    for (uint32_t t0 = 0; t0 < cnt; t0 += kWG) {
        const uint32_t i = t0 + threadIdx.x;
        bool ok = false;
        uint64_t v = 0;
        if (i < cnt) {
            if (c.from_build) {
                uint32_t rid = pb[i0 + i];
                if (c.rids) rid = c.rids[rid];
                ok = valid_bit(c.valid_bits, rid);
                v = W == 4 ? (uint64_t)(uint32_t) static_cast<const int32_t *>(c.values)[rid]
                           : static_cast<const uint64_t *>(c.values)[rid];
            } else {
                ok = probe_value<W>(c, w, pp[i0 + i], &v);
            }
            if (W == 4 && c.is_int32 && (int32_t)(uint32_t)v == kI32Null) ok = false;  // value_t NULL sentinel
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
        *reinterpret_cast<uint16_t *>(page) = (uint16_t)cnt;
        *reinterpret_cast<uint16_t *>(page + 2) = (uint16_t)running;
    }
}

template <bool NL>
__global__ __launch_bounds__(kWG) void k_fused_write(FusedArgs a) {
    __shared__ uint32_t pb[kFusedBatch];
    __shared__ uint32_t pp[kFusedBatch];
    __shared__ WinState win[kFusedMaxCols];
    __shared__ uint32_t lds4[4];
    __shared__ uint64_t red[4];
    __shared__ uint32_t s_fill;
    const uint32_t pg = blockIdx.x;
    const uint64_t out_rows = a.off_rows[pg + 1] - a.off_rows[pg];
    if (out_rows == 0) return;
    NLKeys nl;
    if (NL) nl = *a.nl;
    const KeySource &k = a.key;
    const uint8_t *kp = k.pages + (size_t)pg * kPage;
    const uint32_t n = ld16(kp), nv = ld16(kp + 2);
    const uint64_t r0 = k.row_start[pg], r1 = r0 + n;

    // page windows of the probe-side payload columns
    for (uint32_t c = 0; c < a.ncols; ++c) {
        const FusedCol &col = a.cols[c];
        if (col.from_build) continue;
        if (threadIdx.x == 0) {
            uint32_t lo = 0, hi = col.num_pages;
            while (lo < hi) {
                uint32_t mid = (lo + hi) >> 1;
                if (col.row_start[mid + 1] <= r0) lo = mid + 1;
                else hi = mid;
            }
            win[c].first = lo;
            uint32_t cntw = 0;
            while (lo + cntw < col.num_pages && cntw < kWin && col.row_start[lo + cntw] < r1) ++cntw;
            bool fits = lo + cntw >= col.num_pages || col.row_start[lo + cntw] >= r1;
            win[c].count = fits ? cntw : 0;
        }
        __syncthreads();
        uint32_t cw = win[c].count;
        for (uint32_t i = threadIdx.x; i <= cw && cw; i += kWG) win[c].rs[i] = col.row_start[win[c].first + i];
        __syncthreads();
    }

    const uint8_t *bm = kp + kPage - ((n + 7) >> 3);
    const bool dense = n == nv;
    uint32_t running = 0;      // non-NULL keys seen so far in the page
    uint32_t batch = 0;
    uint64_t staged = 0;  // output rows of this workgroup staged so far
    if (threadIdx.x == 0) s_fill = 0;
    __syncthreads();

    auto flush = [&](uint32_t rows_in_batch) {
        for (uint32_t c = 0; c < a.ncols; ++c) {
            const FusedCol &col = a.cols[c];
            if (col.width == 4) {
                uint8_t *page = col.out + (a.off_p32[pg] + batch) * (uint64_t)kPage;
                write_page<4>(col, win[c], pb, pp, 0, rows_in_batch, page, lds4);
            } else {
                uint64_t base = a.off_p64[pg] + 2ull * batch;
                uint32_t h0 = rows_in_batch < kFusedHalf ? rows_in_batch : kFusedHalf;
                write_page<8>(col, win[c], pb, pp, 0, h0, col.out + base * kPage, lds4);
                if (rows_in_batch > kFusedHalf)
                    write_page<8>(col, win[c], pb, pp, kFusedHalf, rows_in_batch - kFusedHalf,
                                  col.out + (base + 1) * kPage, lds4);
            }
        }
        ++batch;
        __syncthreads();
    };

    for (uint32_t t0 = 0; t0 < n; t0 += kWG) {
        const uint32_t r = t0 + threadIdx.x;
        const bool inr = r < n;
        bool ok = inr && (dense || page_bit(bm, r));
        uint64_t mm;
        uint32_t tt;
        uint32_t vi = dense ? r : running + tile_rank(ok, &mm, &tt, lds4);
        ok = ok && vi < nv;
        const int32_t key = ok ? ld32(kp + 4 + 4 * (size_t)vi) : 0;
        const uint32_t prid = (uint32_t)(r0 + r);
        uint64_t rs = 0, re = 0;
        uint32_t m = ok ? lookup<NL>(key, a.ht, nl, &rs, &re) : 0;
        uint64_t tile_total;
        const uint64_t pre = block_excl_scan<uint64_t>((uint64_t)m, &tile_total, red);
        // never stage more rows than the counting pass sized the pages for
        if (tile_total > out_rows - staged) tile_total = out_rows - staged;
        staged += tile_total;
        // stage this tile's pairs, flushing full batches
        uint64_t done = 0;
        while (done < tile_total) {
            const uint32_t fill = s_fill;
            const uint64_t space = kFusedBatch - fill;
            const uint64_t take = tile_total - done < space ? tile_total - done : space;
            // my ordinals [pre, pre + m) intersect [done, done + take)
            if (m) {
                uint64_t lo = pre > done ? pre : done;
                uint64_t hi = pre + m < done + take ? pre + m : done + take;
                if (lo < hi)
                    emit_range<NL>(key, prid, a.ht, nl, rs, re, (uint32_t)(lo - pre), (uint32_t)(hi - pre),
                                   (uint32_t)(fill + (lo - done)), pb, pp);
            }
            done += take;
            __syncthreads();
            if (threadIdx.x == 0) s_fill = fill + (uint32_t)take;
            __syncthreads();
            if (s_fill == kFusedBatch) {
                flush(kFusedBatch);
                if (threadIdx.x == 0) s_fill = 0;
                __syncthreads();
            }
        }
        if (!dense) running += tt;
        if (staged == out_rows) break;  // every match of this page is staged
    }
    if (s_fill) flush(s_fill);
}

void fused_write(const FusedArgs &a, uint32_t units, bool nl, hipStream_t s) {
    if (!units) return;
    if (nl) k_fused_write<true><<<units, kWG, 0, s>>>(a);
    else k_fused_write<false><<<units, kWG, 0, s>>>(a);
    CJ_HIP(hipGetLastError());
}

}  // namespace cj
