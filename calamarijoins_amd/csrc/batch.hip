// batch.hip -- persistent work-list kernels of the batched radix join
// (batch.hpp).
//
// The per-plan algorithm is the reference's (build side = fewer rows,
// partition the build keys by hash bits, build each partition's slice of an
// unchained directory, probe, materialize: hashtable.h:197-246,405-505,
// hash_join.h:156-244, materialize.h:85-186).  What changes: the probe side
// is partitioned too, output values travel with the keys, a build partition
// is probed from LDS, and one launch serves every plan of the batch.  A
// workgroup takes global units blockIdx.x, blockIdx.x + gridDim.x, ... and
// finds each unit's plan by a binary search over the plans' unit prefix.
#include "batch.hpp"
#include "probe8.hpp"

namespace cj {

namespace {

constexpr uint32_t kMaxSlots = 1u << kBatchMaxLogPart;  // LDS arrays hold the largest partition

// The next join unit: from the phase's work queue (one atomic on one
// counter per unit claimed, every thread gets the result) or strided by
// workgroup.  Only the join phases may take the queue: their units are
// large (thousands of probe tuples) and of uneven cost; page-sized units
// would serialize on the one counter (about 6 ns a claim on MI355X).
__device__ __forceinline__ uint32_t next_unit(bool queue, uint32_t *q, uint32_t *slot, uint32_t prev, bool first) {
    if (!queue) return first ? blockIdx.x : prev + gridDim.x;
    if (threadIdx.x == 0) *slot = atomicAdd(q, 1u);
    __syncthreads();
    const uint32_t g = *slot;
    __syncthreads();
    return g;
}

// Directory slot of a key in a plan of 2^log_slots slots (log_slots <= 30):
// the top bits of a 32-bit multiplicative (Fibonacci) hash, one multiply.
// Keys of a dense domain spread evenly over the slots.
__device__ __forceinline__ uint32_t bslot_of(int32_t key, uint32_t log_slots) {
    return ((uint32_t)key * 0x9E3779B1u) >> (32 - log_slots);
}

// Largest k < n with start[k] <= g (entries with no units share a start).
template <class T>
__device__ __forceinline__ uint32_t find_owner(const T *start, uint32_t n, T g) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (start[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// find_owner that keeps its last answer's unit range: the units a wave
// takes one after another mostly lie in the range of the one before, so most
// calls skip the search (a chain of dependent scalar loads).
struct OwnerCache {
    uint32_t k = 0, lo = 1, hi = 0;
};
__device__ __forceinline__ uint32_t owner_of(OwnerCache &c, const uint32_t *start, uint32_t n, uint32_t g) {
    if (g < c.lo || g >= c.hi) {
        c.k = find_owner(start, n, g);
        c.lo = start[c.k];
        c.hi = start[c.k + 1];
    }
    return c.k;
}

}  // namespace

// ------------------------------------------------------ partition histogram
// Partition of a key = top bits of its directory slot (a plan's partitions
// own contiguous slot ranges).  NULL keys take no part in a join
// (hashtable.h:439-478, hash_join.h:156-244).
//
// A workgroup takes a contiguous range of key pages (so consecutive pages of
// one plan) and keeps one LDS histogram while the plan stays the same; it
// adds the histogram to the global counts when the plan changes and at the
// end -- one global atomic per partition per range and plan instead of per
// page.  Counting needs no row positions: a page's non-NULL keys are its
// value array, read directly (the bitmap only places them in rows).
template <bool PROBE>
__global__ __launch_bounds__(kWG) void k_bx_part_count(BatchArgs a, uint32_t units, const BPlan *__restrict__ plans,
                                                       const uint32_t *__restrict__ pstart, const uint32_t *__restrict__ part_start) {
    extern __shared__ uint32_t hist[];  // max_parts
    uint64_t *count = PROBE ? a.pcount : a.bcount;
    const uint32_t lp = a.log_part;
    const uint32_t per = (units + gridDim.x - 1) / gridDim.x;
    const uint32_t g0 = blockIdx.x * per, g1 = g0 + per < units ? g0 + per : units;
    uint32_t k = ~0u, P = 0, ls = 0;
    auto flush = [&]() {
        __syncthreads();
        const uint32_t pb = part_start[k];
        for (uint32_t i = threadIdx.x; i < P; i += kWG)
            if (hist[i]) atomicAdd((unsigned long long *)&count[pb + i], (unsigned long long)hist[i]);
    };
    for (uint32_t g = g0; g < g1; ++g) {
        const uint32_t kg = find_owner(pstart, a.nplans, g);
        if (kg != k) {
            if (k != ~0u) flush();
            __syncthreads();  // the histogram is read before it is cleared
            k = kg;
            ls = plans[k].log_slots;
            P = 1u << (ls - lp);
            for (uint32_t i = threadIdx.x; i < P; i += kWG) hist[i] = 0;
            __syncthreads();
        }
        const BSide &sd = PROBE ? plans[k].p : plans[k].b;
        const CJ_G uint8_t *p = glb(sd.k_pages) + (size_t)(g - pstart[k]) * kPage;
        const uint32_t nv = ld16(p + 2);
        const CJ_G int32_t *vals = reinterpret_cast<const CJ_G int32_t *>(p + 4);
        for (uint32_t v0 = 0; v0 < nv; v0 += kTile) {
            int32_t key[kRPT];
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) {
                const uint32_t v = v0 + kWG * j + threadIdx.x;
                key[j] = vals[v < nv ? v : nv - 1];
            }
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j)
                if (v0 + kWG * j + threadIdx.x < nv) atomicAdd(&hist[bslot_of(key[j], ls) >> lp], 1u);
        }
    }
    if (k != ~0u) flush();
}

// ------------------------------------------------------ partition scatter
// Keys go to their partition's range of the side's key array; carried
// values to the same plan-local position of the carried arrays.  Per key
// page: each non-NULL key takes a rank in its partition (LDS atomic) and is
// recorded at its value index; one global reservation per non-empty
// partition; an LDS scan orders the page's tuples by partition; then the
// keys and each carried column are written out in that order, so
// consecutive lanes store consecutive slots of a partition's run (a page
// of a plan with few partitions writes long coalesced runs instead of one
// scattered store per tuple).

// The key pages of a scatter launch dealt to the eight XCDs in contiguous
// eighths: workgroup b runs on XCD b % 8 (workgroups are dispatched to the
// XCDs in turn), so the pages one XCD takes at a time are consecutive pages
// of a few plans, and the partial cache lines that its scattered stores
// leave in those plans' partition runs fill up in that XCD's L2 before they
// are written back (instead of eight L2s each writing back a piece of every
// line).  A grid that is not a multiple of 8 strides plainly.
struct XcdUnits {
    uint32_t first, end, step;
};
__device__ __forceinline__ XcdUnits xcd_units(uint32_t units) {
    if (gridDim.x < 8 || gridDim.x % 8) return {blockIdx.x, units, gridDim.x};
    const uint32_t x = blockIdx.x & 7;
    const uint32_t lo = (uint32_t)((uint64_t)units * x / 8), hi = (uint32_t)((uint64_t)units * (x + 1) / 8);
    return {lo + (blockIdx.x >> 3), hi, gridDim.x >> 3};
}

// Carried column c of side sd at table row `row`.
__device__ __forceinline__ int32_t carried4(const BSide &sd, uint32_t c, uint64_t row) {
    const uint32_t R = sd.src_rows[c];
    if (R) {  // no NULLs, R rows on every page but the last: read in place
        const uint32_t r = (uint32_t)row, q = r / R;
        return ld32(glb(sd.src[c]) + (size_t)q * kPage + 4 + 4 * (size_t)(r - q * R));
    }
    return glb(reinterpret_cast<const int32_t *>(sd.src[c]))[row];  // decoded: NULL folded into INT32_MIN
}

constexpr uint32_t kRankBits = 11;  // ranks within a page: < 2048 non-NULL keys (a key page holds <= 2047)

template <bool PROBE>
__global__ __launch_bounds__(kWG) void k_bx_part_scatter(BatchArgs a, uint32_t units, const BPlan *__restrict__ plans,
                                                         const uint32_t *__restrict__ pstart, const uint32_t *__restrict__ part_start,
                                                         const uint64_t *__restrict__ off, uint32_t lds_parts) {
    // global run start of each partition (u64) then the counts / local
    // starts (u32): declared as u64 so the u64 half stays 8-byte aligned
    // whatever static LDS precedes the array; lds_parts >= the plans' P
    extern __shared__ uint64_t sm64[];
    uint64_t *gb = sm64;
    uint32_t *cnt = reinterpret_cast<uint32_t *>(sm64 + lds_parts);
    // (a tuple's key is read back from the page's value array at the
    // write-out: no LDS copy, so eight workgroups fit a CU)
    __shared__ uint32_t ainfo[kTile];  // by value index: partition << kRankBits | rank
    __shared__ uint16_t arow[kTile];   // row in the page
    __shared__ uint16_t order[kTile];  // partition order -> value index
    __shared__ uint32_t lds32[32];
    __shared__ uint32_t lds4[4];
    uint64_t *gcur = PROBE ? a.pcur : a.bcur;
    int32_t *keys_out = PROBE ? a.pkeys : a.bstage_keys;
    const uint32_t lp = a.log_part;
    const XcdUnits xu = xcd_units(units);
    for (uint32_t g = xu.first; g < xu.end; g += xu.step) {
        const uint32_t k = find_owner(pstart, a.nplans, g);
        const BPlan &pl = plans[k];
        const BSide &sd = PROBE ? pl.p : pl.b;
        const uint32_t ls = pl.log_slots;
        const uint32_t P = 1u << (ls - lp);
        const uint32_t page = g - pstart[k];
        for (uint32_t i = threadIdx.x; i < P; i += kWG) cnt[i] = 0;
        __syncthreads();
        const CJ_G uint8_t *p = glb(sd.k_pages) + (size_t)page * kPage;
        const uint32_t n = ld16(p), nv = ld16(p + 2);
        uint32_t running = 0;
        for (uint32_t t0 = 0; t0 < n; t0 += kTile) {
            int32_t key[kRPT];
            uint32_t vi[kRPT];
            const uint32_t valid = load_keys8(p, n, nv, n == nv, t0, &running, key, lds32, vi);
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j)
                if ((valid >> j) & 1u) {
                    const uint32_t q = (uint32_t)(bslot_of(key[j], ls) >> lp);
                    const uint32_t rank = atomicAdd(&cnt[q], 1u);
                    ainfo[vi[j]] = q << kRankBits | rank;
                    arow[vi[j]] = (uint16_t)(t0 + kWG * j + threadIdx.x);
                }
            __syncthreads();  // lds32 of the next tile's ranking
        }
        // one reservation per non-empty partition, then the counts become
        // the partitions' starts in the page's partition order
        const uint32_t pb = part_start[k];
        for (uint32_t i = threadIdx.x; i < P; i += kWG)
            gb[i] = cnt[i] ? atomicAdd((unsigned long long *)&gcur[pb + i], (unsigned long long)cnt[i]) : 0;
        const uint32_t per = (P + kWG - 1) / kWG;
        const uint32_t c0 = threadIdx.x * per;
        uint32_t local = 0;
        for (uint32_t i = 0; i < per && c0 + i < P; ++i) local += cnt[c0 + i];
        uint32_t nval;
        uint32_t pre = block_excl_scan<uint32_t>(local, &nval, lds4);  // barrier: the reservations read cnt first
        for (uint32_t i = 0; i < per && c0 + i < P; ++i) {
            const uint32_t c = cnt[c0 + i];
            cnt[c0 + i] = pre;
            pre += c;
        }
        __syncthreads();
        for (uint32_t v = threadIdx.x; v < nval; v += kWG) {
            const uint32_t in = ainfo[v];
            order[cnt[in >> kRankBits] + (in & ((1u << kRankBits) - 1))] = (uint16_t)v;
        }
        __syncthreads();
        // write-out in partition order: keys, then each carried column
        const uint64_t plan_base = off[pb];
        const uint64_t r0 = glb(sd.k_rs)[page];
        const CJ_G int32_t *vals = reinterpret_cast<const CJ_G int32_t *>(p + 4);
        const uint32_t ncar = sd.ncar;
        // keys and a first 4-byte carried column in one pass (their loads
        // in flight together), other columns after
        const bool first4 = ncar > 0 && sd.car_w[0] == 4;
        CJ_G int32_t *dst0 = first4 ? glb(reinterpret_cast<int32_t *>(sd.car[0])) : nullptr;
        for (uint32_t s = threadIdx.x; s < nval; s += kWG) {
            const uint32_t v = order[s], in = ainfo[v];
            const uint64_t d = gb[in >> kRankBits] + (in & ((1u << kRankBits) - 1));
            const int32_t key = vals[v];
            if (first4) dst0[d - plan_base] = carried4(sd, 0, r0 + arow[v]);
            keys_out[d] = key;
        }
        for (uint32_t c = first4 ? 1 : 0; c < ncar; ++c) {
            if (sd.car_w[c] == 4) {
                CJ_G int32_t *dst = glb(reinterpret_cast<int32_t *>(sd.car[c]));
                for (uint32_t s = threadIdx.x; s < nval; s += kWG) {
                    const uint32_t v = order[s], in = ainfo[v];
                    dst[gb[in >> kRankBits] + (in & ((1u << kRankBits) - 1)) - plan_base] = carried4(sd, c, r0 + arow[v]);
                }
            } else {
                const CJ_G uint64_t *src = glb(reinterpret_cast<const uint64_t *>(sd.src[c]));
                const CJ_G uint32_t *sv = glb(sd.src_valid[c]);
                CJ_G uint64_t *dst = glb(reinterpret_cast<uint64_t *>(sd.car[c]));
                CJ_G uint8_t *dv = glb(sd.car_vld[c]);
                for (uint32_t s = threadIdx.x; s < nval; s += kWG) {
                    const uint32_t v = order[s], in = ainfo[v];
                    const uint64_t row = r0 + arow[v];
                    const uint64_t d = gb[in >> kRankBits] + (in & ((1u << kRankBits) - 1)) - plan_base;
                    dst[d] = src[row];
                    dv[d] = (uint8_t)valid_bit(sv, row);
                }
            }
        }
        __syncthreads();
    }
}

// The same scatter for plans with many partitions per page (a page's runs
// would be a tuple or two: staging buys nothing there): every tuple takes
// its position from an LDS cursor and is stored directly.  The global
// reservations are issued kResPer at a time per thread before any result is
// awaited.
constexpr uint32_t kResPer = 16;

template <bool PROBE>
__global__ __launch_bounds__(kWG) void k_bx_part_scatter_direct(BatchArgs a, uint32_t units,
                                                                const BPlan *__restrict__ plans,
                                                                const uint32_t *__restrict__ pstart,
                                                                const uint32_t *__restrict__ part_start,
                                                                const uint64_t *__restrict__ off, uint32_t lds_parts) {
    extern __shared__ uint64_t sm64[];  // cursors (u64) then hist[lds_parts] (u32)
    uint64_t *cur = sm64;
    uint32_t *hist = reinterpret_cast<uint32_t *>(sm64 + lds_parts);
    __shared__ uint32_t lds32[32];
    uint64_t *gcur = PROBE ? a.pcur : a.bcur;
    int32_t *keys_out = PROBE ? a.pkeys : a.bstage_keys;
    const uint32_t lp = a.log_part;
    const XcdUnits xu = xcd_units(units);
    for (uint32_t g = xu.first; g < xu.end; g += xu.step) {
        const uint32_t k = find_owner(pstart, a.nplans, g);
        const BPlan &pl = plans[k];
        const BSide &sd = PROBE ? pl.p : pl.b;
        const uint32_t ls = pl.log_slots;
        const uint32_t P = 1u << (ls - lp);
        const uint32_t page = g - pstart[k];
        const CJ_G uint8_t *p = glb(sd.k_pages) + (size_t)page * kPage;
        // the histogram counts the page's non-NULL keys straight from its
        // value array (at most 2047 values: one slot a thread; addresses stay
        // inside the page), loaded with the header
        {
            int32_t key[kRPT];
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) {
                const uint32_t v = kWG * j + threadIdx.x;
                key[j] = ld32(p + 4 + 4 * (v < kTile - 2 ? v : kTile - 2));
            }
            const uint32_t nv0 = ld16(p + 2);
            for (uint32_t i = threadIdx.x; i < P; i += kWG) hist[i] = 0;
            __syncthreads();
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j)
                if (kWG * j + threadIdx.x < nv0) atomicAdd(&hist[bslot_of(key[j], ls) >> lp], 1u);
            __syncthreads();
        }
        const uint32_t n = ld16(p), nv = ld16(p + 2);
        uint32_t running = 0;
        const uint32_t pb = part_start[k];
        for (uint32_t i0 = 0; i0 < P; i0 += kWG * kResPer) {
            uint64_t r[kResPer];
#pragma unroll
            for (uint32_t q = 0; q < kResPer; ++q) {
                const uint32_t i = i0 + q * kWG + threadIdx.x;
                const uint32_t h = i < P ? hist[i] : 0u;
                r[q] = h ? atomicAdd((unsigned long long *)&gcur[pb + i], (unsigned long long)h) : 0;
            }
#pragma unroll
            for (uint32_t q = 0; q < kResPer; ++q) {
                const uint32_t i = i0 + q * kWG + threadIdx.x;
                if (i < P) cur[i] = r[q];
            }
        }
        __syncthreads();
        const uint64_t plan_base = off[pb];
        const uint64_t r0 = glb(sd.k_rs)[page];
        const uint32_t ncar = sd.ncar;
        running = 0;
        for (uint32_t t0 = 0; t0 < n; t0 += kTile) {
            int32_t key[kRPT];
            const uint32_t valid = load_keys8(p, n, nv, n == nv, t0, &running, key, lds32);
            uint64_t pos[kRPT];
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) {
                pos[j] = 0;
                if ((valid >> j) & 1u)
                    pos[j] = atomicAdd((unsigned long long *)&cur[bslot_of(key[j], ls) >> lp], 1ull);
            }
            // per carried column: all loads, then all stores
            for (uint32_t c = 0; c < ncar; ++c) {
                if (sd.car_w[c] == 4) {
                    int32_t v[kRPT];
#pragma unroll
                    for (uint32_t j = 0; j < kRPT; ++j)
                        if ((valid >> j) & 1u) v[j] = carried4(sd, c, r0 + t0 + kWG * j + threadIdx.x);
                    CJ_G int32_t *dst = glb(reinterpret_cast<int32_t *>(sd.car[c]));
#pragma unroll
                    for (uint32_t j = 0; j < kRPT; ++j)
                        if ((valid >> j) & 1u) dst[pos[j] - plan_base] = v[j];
                } else {
                    const CJ_G uint64_t *src = glb(reinterpret_cast<const uint64_t *>(sd.src[c]));
                    const CJ_G uint32_t *sv = glb(sd.src_valid[c]);
                    uint64_t v[kRPT];
                    uint32_t vb = 0;
#pragma unroll
                    for (uint32_t j = 0; j < kRPT; ++j)
                        if ((valid >> j) & 1u) {
                            const uint64_t row = r0 + t0 + kWG * j + threadIdx.x;
                            v[j] = src[row];
                            vb |= (uint32_t)valid_bit(sv, row) << j;
                        }
                    CJ_G uint64_t *dst = glb(reinterpret_cast<uint64_t *>(sd.car[c]));
                    CJ_G uint8_t *dv = glb(sd.car_vld[c]);
#pragma unroll
                    for (uint32_t j = 0; j < kRPT; ++j)
                        if ((valid >> j) & 1u) {
                            dst[pos[j] - plan_base] = v[j];
                            dv[pos[j] - plan_base] = (uint8_t)((vb >> j) & 1u);
                        }
                }
            }
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j)
                if ((valid >> j) & 1u) keys_out[pos[j]] = key[j];
            __syncthreads();
        }
    }
}

// --------------------------------------------- per-partition build (slot order)
// build_partition (hashtable.h:197-246): count per slot, exclusive prefix,
// place the tuples; the directory keeps each slot's partition-local end.
// A partition of at most kBatchSortCap tuples is ordered by (slot, key), so
// the matches of a probe key are consecutive and a slot can be searched by
// bisection: keys and source tuples in LDS, each slot sorted by one thread
// (insertion by bisection; a key already in place, such as the next of a
// run of equal keys, costs one comparison), then written out in order.  A
// larger one is left in slot order.
// A long slot (many copies of a few keys) is ordered by the whole
// workgroup, one distinct key per round, ascending: the smallest key left and
// its count c (reductions), then the copies outside the first c places and
// the other keys inside them are listed (in LDS, any order) and swapped pair
// by pair.  O(slot length / kWG) a round; an insertion sort by one thread
// would move each copy of one key past the copies of another.
constexpr uint32_t kLongSlot = 32;  // longer slots are ordered by the workgroup

__device__ __forceinline__ int32_t block_min_i32(int32_t v, int32_t *red) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    if (lane_id() == 0) red[wave_id()] = v;
    __syncthreads();
    int32_t r = red[0];
#pragma unroll
    for (uint32_t w = 1; w < kWG / 64; ++w) r = min(r, red[w]);
    __syncthreads();
    return r;
}
__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, uint32_t *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane_id() == 0) red[wave_id()] = v;
    __syncthreads();
    uint32_t r = 0;
#pragma unroll
    for (uint32_t w = 0; w < kWG / 64; ++w) r += red[w];
    __syncthreads();
    return r;
}

// lists: 2 x half u16 entries (positions < 2^16); cnt2: two LDS counters.
__device__ void group_slot_wg(int32_t *skey, uint16_t *sidx, uint32_t lo, uint32_t hi, uint16_t *lists, uint32_t half,
                              uint32_t *cnt2, int32_t *redi, uint32_t *redu) {
    uint32_t w = lo;
    while (w < hi) {
        int32_t mn = INT32_MAX;
        for (uint32_t i = w + threadIdx.x; i < hi; i += kWG) mn = min(mn, skey[i]);
        mn = block_min_i32(mn, redi);
        uint32_t c = 0;
        for (uint32_t i = w + threadIdx.x; i < hi; i += kWG) c += skey[i] == mn;
        c = block_sum_u32(c, redu);
        if (w + c >= hi) break;  // the largest key is left: in place
        if (threadIdx.x == 0) cnt2[0] = cnt2[1] = 0;
        __syncthreads();
        // others in [w, w + c) and copies in [w + c, hi): as many of each
        // (at most half of the slot: c of one kind and hi - w - c of the
        // other bound them both)
        for (uint32_t i = w + threadIdx.x; i < w + c; i += kWG)
            if (skey[i] != mn) lists[atomicAdd(&cnt2[0], 1u)] = (uint16_t)i;
        for (uint32_t i = w + c + threadIdx.x; i < hi; i += kWG)
            if (skey[i] == mn) lists[half + atomicAdd(&cnt2[1], 1u)] = (uint16_t)i;
        __syncthreads();
        const uint32_t np = cnt2[0];
        for (uint32_t q = threadIdx.x; q < np; q += kWG) {
            const uint32_t x = lists[q], y = lists[half + q];
            const int32_t tk = skey[x];
            const uint16_t ti = sidx[x];
            skey[x] = skey[y];
            sidx[x] = sidx[y];
            skey[y] = tk;
            sidx[y] = ti;
        }
        __syncthreads();
        w += c;
    }
}

__device__ __forceinline__ void bx_build_part(const BatchArgs &a, const BPlan *__restrict__ plans,
                                              const uint32_t *__restrict__ part_start, const uint64_t *__restrict__ boff,
                                              const int32_t *__restrict__ bstage_keys, uint32_t g, uint32_t sort_cap,
                                              uint32_t *cnt, uint32_t *lds4, int32_t *skey, uint16_t *sidx,
                                              uint16_t *lists, uint32_t *longs, uint32_t *misc, int32_t *redi) {
    const uint32_t S = 1u << a.log_part, smask = S - 1;
    const uint32_t k = find_owner(part_start, a.nplans, g);
    const BPlan &pl = plans[k];
    const uint32_t ls = pl.log_slots;
    const uint64_t t_a = boff[g], t_e = boff[g + 1];
    const uint64_t plan_base = boff[part_start[k]];
    const bool sorted = t_e - t_a <= sort_cap;
    for (uint32_t i = threadIdx.x; i < S; i += kWG) cnt[i] = 0;
    __syncthreads();
    for (uint64_t i = t_a + threadIdx.x; i < t_e; i += kWG)
        atomicAdd(&cnt[(uint32_t)bslot_of(bstage_keys[i], ls) & smask], 1u);
    __syncthreads();
    const uint32_t per = S / kWG;
    const uint32_t s0 = threadIdx.x * per;
    uint32_t local = 0;
#pragma unroll
    for (uint32_t i = 0; i < per; ++i) local += cnt[s0 + i];
    uint32_t tot;
    uint32_t pre = block_excl_scan<uint32_t>(local, &tot, lds4);
    for (uint32_t i = 0; i < per; ++i) {
        const uint32_t c = cnt[s0 + i];
        cnt[s0 + i] = pre;
        pre += c;
    }
    __syncthreads();
    const uint32_t ncar = pl.b.ncar;
    if (sorted) {
        // slot order in LDS: cnt[s] ends as the end of slot s
        if (threadIdx.x == 0) misc[2] = 0;  // long slots found
        for (uint64_t i = t_a + threadIdx.x; i < t_e; i += kWG) {
            const int32_t key = bstage_keys[i];
            const uint32_t pos = atomicAdd(&cnt[(uint32_t)bslot_of(key, ls) & smask], 1u);
            skey[pos] = key;
            sidx[pos] = (uint16_t)(i - t_a);
        }
        __syncthreads();
        for (uint32_t sl = threadIdx.x; sl < S; sl += kWG) {
            const uint32_t lo = sl ? cnt[sl - 1] : 0u, hi = cnt[sl];
            if (hi - lo > kLongSlot) {
                longs[atomicAdd(&misc[2], 1u)] = sl;
                continue;
            }
            for (uint32_t i = lo + 1; i < hi; ++i) {
                const int32_t kx = skey[i];
                if (skey[i - 1] <= kx) continue;
                const uint16_t ix = sidx[i];
                // first of [lo, i) above kx, then [that, i) one place up
                uint32_t x = lo, y = i - 1;
                while (x < y) {
                    const uint32_t mid = (x + y) >> 1;
                    if (skey[mid] > kx) y = mid;
                    else x = mid + 1;
                }
                uint32_t j = i;
                for (; j >= x + 4; j -= 4) {
                    const int32_t k0 = skey[j - 1], k1 = skey[j - 2], k2 = skey[j - 3], k3 = skey[j - 4];
                    const uint16_t i0 = sidx[j - 1], i1 = sidx[j - 2], i2 = sidx[j - 3], i3 = sidx[j - 4];
                    skey[j] = k0, skey[j - 1] = k1, skey[j - 2] = k2, skey[j - 3] = k3;
                    sidx[j] = i0, sidx[j - 1] = i1, sidx[j - 2] = i2, sidx[j - 3] = i3;
                }
                for (; j > x; --j) {
                    skey[j] = skey[j - 1];
                    sidx[j] = sidx[j - 1];
                }
                skey[x] = kx;
                sidx[x] = ix;
            }
        }
        __syncthreads();
        const uint32_t nlong = misc[2];
        for (uint32_t q = 0; q < nlong; ++q) {
            const uint32_t sl = longs[q];
            group_slot_wg(skey, sidx, sl ? cnt[sl - 1] : 0u, cnt[sl], lists, sort_cap / 2, misc, redi, lds4);
        }
        const uint32_t nb = (uint32_t)(t_e - t_a);
        for (uint32_t pos = threadIdx.x; pos < nb; pos += kWG) {
            a.bkeys[t_a + pos] = skey[pos];
            const uint64_t src = t_a + sidx[pos] - plan_base, dst = t_a + pos - plan_base;
            for (uint32_t c = 0; c < ncar; ++c) {
                if (pl.b.car_w[c] == 4) {
                    glb(reinterpret_cast<int32_t *>(pl.bslot_car[c]))[dst] =
                        glb(reinterpret_cast<const int32_t *>(pl.b.car[c]))[src];
                } else {
                    glb(reinterpret_cast<uint64_t *>(pl.bslot_car[c]))[dst] =
                        glb(reinterpret_cast<const uint64_t *>(pl.b.car[c]))[src];
                    glb(pl.bslot_vld[c])[dst] = glb(pl.b.car_vld[c])[src];
                }
            }
        }
    } else {
        for (uint64_t i = t_a + threadIdx.x; i < t_e; i += kWG) {
            const int32_t key = bstage_keys[i];
            const uint32_t pos = atomicAdd(&cnt[(uint32_t)bslot_of(key, ls) & smask], 1u);
            a.bkeys[t_a + pos] = key;
            const uint64_t src = i - plan_base, dst = t_a + pos - plan_base;
            for (uint32_t c = 0; c < ncar; ++c) {
                if (pl.b.car_w[c] == 4) {
                    glb(reinterpret_cast<int32_t *>(pl.bslot_car[c]))[dst] =
                        glb(reinterpret_cast<const int32_t *>(pl.b.car[c]))[src];
                } else {
                    glb(reinterpret_cast<uint64_t *>(pl.bslot_car[c]))[dst] =
                        glb(reinterpret_cast<const uint64_t *>(pl.b.car[c]))[src];
                    glb(pl.bslot_vld[c])[dst] = glb(pl.b.car_vld[c])[src];
                }
            }
        }
    }
    __syncthreads();
    uint32_t *d = a.dir + pl.dir_base + ((uint64_t)(g - part_start[k]) << a.log_part);
    for (uint32_t i = threadIdx.x; i < S; i += kWG) d[i] = cnt[i];
    __syncthreads();
}

// The first launch (big = false) takes the partitions of at most sort_cap =
// lds_cap tuples (sorted) and those above kBatchSortCap (slot order only),
// strided by workgroup; the second (big = true, sort_cap = kBatchSortCap,
// more LDS) the ones in between: a workgroup checks kWG partitions at once
// and builds the ones it finds.
__global__ __launch_bounds__(kWG) void k_bx_build(BatchArgs a, const BPlan *__restrict__ plans,
                                                  const uint32_t *__restrict__ part_start, const uint64_t *__restrict__ boff,
                                                  const int32_t *__restrict__ bstage_keys, uint32_t sort_cap, bool big) {
    __shared__ uint32_t cnt[kMaxSlots];
    __shared__ uint32_t lds4[4];
    __shared__ uint32_t found[kWG];
    __shared__ uint32_t nfound;
    __shared__ uint32_t misc[4];
    __shared__ int32_t redi[kWG / 64];
    // long slots of a partition: at most its tuples / (kLongSlot + 1)
    __shared__ uint32_t longs[kBatchSortCap / (kLongSlot + 1) + 1];
    // keys, source tuples, and the swap lists of the long slots' ordering
    extern __shared__ __align__(16) uint8_t bdyn[];
    int32_t *skey = reinterpret_cast<int32_t *>(bdyn);
    uint16_t *sidx = reinterpret_cast<uint16_t *>(bdyn + 4 * (size_t)sort_cap);
    uint16_t *lists = reinterpret_cast<uint16_t *>(bdyn + 6 * (size_t)sort_cap);
    if (!big) {
        for (uint32_t g = blockIdx.x; g < a.nparts; g += gridDim.x) {
            const uint64_t nb = boff[g + 1] - boff[g];
            if (nb > a.lds_cap && nb <= kBatchSortCap) continue;  // the second launch's
            bx_build_part(a, plans, part_start, boff, bstage_keys, g, sort_cap, cnt, lds4, skey, sidx, lists, longs, misc,
                          redi);
        }
        return;
    }
    for (uint32_t g0 = blockIdx.x * kWG; g0 < a.nparts; g0 += gridDim.x * kWG) {
        if (threadIdx.x == 0) nfound = 0;
        __syncthreads();
        const uint32_t g = g0 + threadIdx.x;
        if (g < a.nparts) {
            const uint64_t nb = boff[g + 1] - boff[g];
            if (nb > a.lds_cap && nb <= kBatchSortCap) found[atomicAdd(&nfound, 1u)] = g;
        }
        __syncthreads();
        const uint32_t nf = nfound;
        for (uint32_t f = 0; f < nf; ++f)
            bx_build_part(a, plans, part_start, boff, bstage_keys, found[f], sort_cap, cnt, lds4, skey, sidx, lists,
                          longs, misc, redi);
        __syncthreads();
    }
}

// ------------------------------------------------------------- join units
// A probe partition is cut into chunks of a.chunk tuples; a partition
// whose build side is empty has no units (it cannot match).
__global__ void k_bx_units(BatchArgs a) {
    const uint32_t g = blockIdx.x * kWG + threadIdx.x;
    if (g >= a.nparts) return;
    const uint64_t np = a.poff[g + 1] - a.poff[g];
    const bool build = a.boff[g + 1] > a.boff[g];
    a.ucount[g] = build ? (np + a.chunk - 1) / a.chunk : 0;
}

// ------------------------------------------------------------------ join
namespace {

// The build partition a join unit probes: in LDS (ends as u16, keys and
// carried values) when it holds at most lds_cap tuples, else read in place.
struct BuildPart {
    const uint16_t *l_ends;
    const uint8_t *l_dyn;
    const uint32_t *g_ends;
    const int32_t *g_keys;  // bkeys + t_a
    uint64_t g_car0;        // plan-local index of the partition's first tuple
    uint32_t smask;         // slots per partition - 1
    uint32_t cap;           // LDS layout: keys[cap], values of each carried column, validity bytes
    uint32_t ncar;
    uint32_t w8;            // bit c: carried column c is 8 bytes wide
    // write pass: the tile's values of the probe side's carried column 0,
    // staged in LDS by tile row (low words, high words, validity bytes)
    uint32_t *pt_lo, *pt_hi;
    uint8_t *pt_vb;
    // byte offset of carried column c's values / validity bytes in LDS
    __device__ uint32_t coff(uint32_t c) const {
        return cap * (4 + 4 * c + 4 * __popc(w8 & ((1u << c) - 1)));
    }
    __device__ uint32_t voff(uint32_t c) const {
        return cap * (4 + 4 * ncar + 4 * __popc(w8)) + cap * __popc(w8 & ((1u << c) - 1));
    }
};

template <bool LDS>
__device__ __forceinline__ uint32_t bp_end(const BuildPart &bp, uint32_t s) {
    if constexpr (LDS) return lds(bp.l_ends)[s];
    else return bp.g_ends[s];
}
template <bool LDS>
__device__ __forceinline__ int32_t bp_key(const BuildPart &bp, uint32_t t) {
    if constexpr (LDS) return lds(reinterpret_cast<const int32_t *>(bp.l_dyn))[t];
    else return bp.g_keys[t];
}

// The match window of the write pass: (probe row in the tile) << 22 | build
// tuple of the partition, kMatchWin of them.
constexpr uint32_t kMatchWin = 2048;
// Match descriptor of a probe tuple whose build partition is staged in LDS
// (written over its key by the count pass, read by the write pass): the
// partition-local index of its first matching build tuple (the partition is
// ordered by slot, then key, so its matches are consecutive) and their
// count, 1023 standing for "1023 or more".
constexpr uint32_t kDescFirstBits = 22;
constexpr uint32_t kDescFirst = (1u << kDescFirstBits) - 1;
constexpr uint32_t kDescMany = 1023;
constexpr uint32_t kMatchTBits = 22;
constexpr uint32_t kMatchT = (1u << kMatchTBits) - 1;
static_assert(kBatchMaxBuildRows <= (1ull << kMatchTBits), "build tuple index must fit the match entry");
static_assert(kWG * 4 <= (1u << (32 - kMatchTBits)), "tile row must fit the match entry");

// A NULL at slot `slot` of page `pgi` of an output column of `rows` rows
// (R a page): clear its validity bit (page-init set them all) and count it
// on the page, whose values the page-fix pass then packs.
__device__ __forceinline__ void bx_null(uint8_t *page, uint32_t pgi, uint32_t slot, uint32_t rows, uint32_t R,
                                        uint32_t *count) {
    const uint32_t r0 = pgi * R;
    const uint32_t n = rows - r0 < R ? rows - r0 : R;
    const uint32_t byte = kPage - ((n + 7) >> 3) + (slot >> 3);
    // the word may also hold bytes before the bitmap: AND leaves them alone
    __hip_atomic_fetch_and(glb(reinterpret_cast<uint32_t *>(page + (byte & ~3u))),
                           ~(1u << ((byte & 3u) * 8 + (slot & 7u))), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(glb(count), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One output column of a plan while the write pass emits it (uniform).
struct BCol {
    uint32_t src, cj, w, rows, pg0;
    uint8_t *out;
};

// The value of output column `col` (width W, source SRC: 0 the join key, 1
// the build side, 2 the probe side, 3 the probe side's carried column 0
// staged in LDS for the tile) for match entry (key, build tuple t of the
// partition, probe tuple pt): the key, carried build value g_car0 + t
// (plan-local index of the slot-ordered carried columns), carried probe
// value pt (plan-local index of the partition-ordered carried columns) or,
// staged, the value of tile row pt.
// `ok` is false for a NULL.
template <int W, int SRC, bool LDS>
__device__ __forceinline__ uint64_t bx_get(const BPlan &pl, const BCol &col, const BuildPart &bp, int32_t key,
                                           uint32_t t, uint32_t pt, bool &ok) {
    if constexpr (W == 4) {
        int32_t v;
        if constexpr (SRC == 0) {
            v = key;  // equal keys: the build key is the probe key
        } else if constexpr (SRC == 1) {
            if constexpr (LDS) v = lds(reinterpret_cast<const int32_t *>(bp.l_dyn + bp.coff(col.cj)))[t];
            else v = glb(reinterpret_cast<const int32_t *>(pl.bslot_car[col.cj]))[bp.g_car0 + t];
        } else if constexpr (SRC == 3) {
            v = (int32_t)lds(bp.pt_lo)[pt];
        } else {
            v = glb(reinterpret_cast<const int32_t *>(pl.p.car[col.cj]))[pt];
        }
        ok = v != kI32Null;  // page_builders.h:123
        return (uint32_t)v;
    } else {
        uint64_t v;
        uint8_t b;
        if constexpr (SRC == 1) {
            if constexpr (LDS) {
                v = lds(reinterpret_cast<const uint64_t *>(bp.l_dyn + bp.coff(col.cj)))[t];
                b = lds(bp.l_dyn + bp.voff(col.cj))[t];
            } else {
                v = glb(reinterpret_cast<const uint64_t *>(pl.bslot_car[col.cj]))[bp.g_car0 + t];
                b = glb(pl.bslot_vld[col.cj])[bp.g_car0 + t];
            }
        } else if constexpr (SRC == 3) {
            v = ((uint64_t)lds(bp.pt_hi)[pt] << 32) | lds(bp.pt_lo)[pt];
            b = lds(bp.pt_vb)[pt];
        } else {
            v = glb(reinterpret_cast<const uint64_t *>(pl.p.car[col.cj]))[pt];
            b = glb(pl.p.car_vld[col.cj])[pt];
        }
        ok = b != 0;
        return v;
    }
}

// Store output row o of column `col`: slot o % R of page o / R.
template <int W>
__device__ __forceinline__ void bx_store(const BCol &col, uint32_t o, uint64_t v, bool ok, uint32_t *page_nulls) {
    constexpr uint32_t R = W == 4 ? kRowsI32 : kRowsI64;
    const uint32_t pgi = o / R, slot = o - pgi * R;
    uint8_t *page = col.out + (size_t)pgi * kPage;
    if (ok) {
        if constexpr (W == 4) glb(reinterpret_cast<uint32_t *>(page + 4))[slot] = (uint32_t)v;
        else glb(reinterpret_cast<uint64_t *>(page + 8))[slot] = v;
    } else {
        bx_null(page, pgi, slot, col.rows, R, page_nulls + col.pg0 + pgi);
    }
}

constexpr uint32_t kEmitBatch = 4;  // window entries per thread whose loads are issued together

// One column of a match window: rows ob + i for entries i of win (probe row
// in the tile << kMatchTBits | build tuple), consecutive rows on
// consecutive lanes; the column's width and source are template
// parameters, so the loop holds no per-entry branch on them.  A batch of
// entries per thread reads the window, the keys and the values with
// clamped indices (every load of the batch in flight together), then
// stores the rows that exist.
template <int W, int SRC, bool LDS>
__device__ __forceinline__ void bx_emit_col(const BPlan &pl, const BCol &col, const BuildPart &bp, const uint32_t *win,
                                            uint32_t cnt, uint32_t ob, uint32_t pt0, uint32_t *page_nulls) {
    const CJ_L uint32_t *lw = lds(win);
    for (uint32_t i0 = threadIdx.x; i0 < cnt; i0 += kEmitBatch * kWG) {
        uint32_t en[kEmitBatch];
#pragma unroll
        for (uint32_t b = 0; b < kEmitBatch; ++b) {
            const uint32_t i = i0 + b * kWG;
            en[b] = lw[i < cnt ? i : cnt - 1];
        }
        uint64_t v[kEmitBatch];
        bool ok[kEmitBatch];
#pragma unroll
        for (uint32_t b = 0; b < kEmitBatch; ++b) {
            const uint32_t t = en[b] & kMatchT;
            v[b] = bx_get<W, SRC, LDS>(pl, col, bp, SRC == 0 ? bp_key<LDS>(bp, t) : 0, t,
                                       pt0 + (en[b] >> kMatchTBits), ok[b]);
        }
#pragma unroll
        for (uint32_t b = 0; b < kEmitBatch; ++b) {
            const uint32_t i = i0 + b * kWG;
            if (i < cnt) bx_store<W>(col, ob + i, v[b], ok[b], page_nulls);
        }
    }
}

// Probe tuples [p0, p1) of the unit, kRPT per thread per tile: returns the
// thread's match count (count pass) or writes the matches (write pass).
template <bool WRITE, bool LDS, uint32_t RPT>
__device__ __forceinline__ uint64_t bx_probe(const int32_t *pkeys, const BPlan &pl, const BOut *bo,
                                             const BuildPart &bp, uint64_t p0, uint64_t p1, uint64_t pbase,
                                             uint64_t obase, uint32_t *lds32, uint32_t *win, uint32_t *page_nulls,
                                             uint32_t *desc, bool sorted) {
    uint64_t mine = 0, base = obase;
    const uint32_t ls = pl.log_slots;
    // the keys of the next tile are loaded while this one is probed
    // (clamped: past the unit's last tuple the loads read its last key; a
    // unit always holds a probe tuple, k_bx_units)
    int32_t nkey[RPT];
#pragma unroll
    for (uint32_t j = 0; j < RPT; ++j) {
        const uint64_t idx = p0 + kWG * j + threadIdx.x;
        nkey[j] = pkeys[idx < p1 ? idx : p1 - 1];
    }
    for (uint64_t i0 = p0; i0 < p1; i0 += kWG * RPT) {
        int32_t key[RPT];
        uint32_t s[RPT], e[RPT], m[RPT];
#pragma unroll
        for (uint32_t j = 0; j < RPT; ++j) {
            key[j] = nkey[j];
            const uint64_t idx = i0 + kWG * (RPT + j) + threadIdx.x;
            nkey[j] = pkeys[idx < p1 ? idx : p1 - 1];
        }
#pragma unroll
        for (uint32_t j = 0; j < RPT; ++j) {
            const uint32_t sl = (uint32_t)bslot_of(key[j], ls) & bp.smask;
            const uint32_t hi = bp_end<LDS>(bp, sl);
            const uint32_t lo = sl ? bp_end<LDS>(bp, sl - 1) : 0u;
            const bool in = i0 + kWG * j + threadIdx.x < p1;
            s[j] = in ? lo : 0u;
            e[j] = in ? hi : 0u;
        }
        // first two matching build tuples (write pass over a partition not
        // ordered by key; count pass over an ordered one: the first)
        uint32_t f0[RPT], f1[RPT];
        if (!WRITE && sorted) {
            // the slot's keys are ascending: its run of the probe key by bisection
#pragma unroll
            for (uint32_t j = 0; j < RPT; ++j) {
                uint32_t lo = s[j], hi = e[j];
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (bp_key<LDS>(bp, mid) < key[j]) lo = mid + 1;
                    else hi = mid;
                }
                uint32_t ul = lo, uh = e[j];
                while (ul < uh) {
                    const uint32_t mid = (ul + uh) >> 1;
                    if (bp_key<LDS>(bp, mid) <= key[j]) ul = mid + 1;
                    else uh = mid;
                }
                m[j] = ul - lo;
                f0[j] = lo;
                f1[j] = 0;
            }
        } else {
#pragma unroll
            for (uint32_t j = 0; j < RPT; ++j) {
                uint32_t c = 0, a0 = 0, a1 = 0;
                for (uint32_t t = s[j]; t < e[j]; ++t) {
                    const bool hit = bp_key<LDS>(bp, t) == key[j];
                    if (WRITE && hit) {
                        a1 = c == 1 ? t : a1;
                        a0 = c == 0 ? t : a0;
                    }
                    c += hit;
                }
                m[j] = c;
                f0[j] = a0;
                f1[j] = a1;
            }
        }
        if constexpr (!WRITE) {
#pragma unroll
            for (uint32_t j = 0; j < RPT; ++j) {
                mine += m[j];
                if (sorted) {
                    // the tuple's key is not read again: its descriptor
                    // takes its place
                    const uint64_t idx = i0 + kWG * j + threadIdx.x;
                    if (idx < p1) desc[idx] = (min(m[j], kDescMany) << kDescFirstBits) | (m[j] ? f0[j] : 0u);
                }
            }
        } else {
            // exclusive prefix in tuple order i0 + 256 j + t
            const uint32_t w = wave_id(), lane = lane_id();
            uint32_t inc[RPT], pre[RPT];
#pragma unroll
            for (uint32_t j = 0; j < RPT; ++j) {
                inc[j] = wave_incl_scan<uint32_t>(m[j]);
                if (lane == 63) lds32[4 * j + w] = inc[j];
            }
            __syncthreads();
            uint32_t run = 0;
#pragma unroll
            for (uint32_t j = 0; j < RPT; ++j)
#pragma unroll
                for (uint32_t ww = 0; ww < 4; ++ww) {
                    if (ww == w) pre[j] = run + inc[j] - m[j];
                    run += lds32[4 * j + ww];
                }
            __syncthreads();
            // expand the tile's matches into the LDS window (probe row in
            // the tile, build tuple), a window at a time; then the threads
            // emit them column by column, consecutive rows on consecutive
            // lanes (full lanes, whatever the per-tuple match counts)
            for (uint32_t w0 = 0; w0 < run; w0 += kMatchWin) {
                const uint32_t w1 = run - w0 < kMatchWin ? run : w0 + kMatchWin;
#pragma unroll
                for (uint32_t j = 0; j < RPT; ++j) {
                    if (!m[j] || pre[j] >= w1 || pre[j] + m[j] <= w0) continue;
                    const uint32_t row = (kWG * j + threadIdx.x) << kMatchTBits;
                    if (m[j] <= 2 && pre[j] >= w0 && pre[j] + m[j] <= w1) {
                        win[pre[j] - w0] = row | f0[j];  // the usual case: the recorded matches
                        if (m[j] == 2) win[pre[j] + 1 - w0] = row | f1[j];
                        continue;
                    }
                    uint32_t q = pre[j];
                    for (uint32_t t = s[j]; t < e[j] && q < w1; ++t) {
                        if (bp_key<LDS>(bp, t) != key[j]) continue;
                        if (q >= w0) win[q - w0] = row | t;
                        ++q;
                    }
                }
                __syncthreads();
                const uint32_t cnt = w1 - w0;
                const uint32_t ob = (uint32_t)(base + w0), pt0 = (uint32_t)(i0 - pbase);
#pragma unroll 1
                for (uint32_t c = 0; c < bo->nout; ++c) {
                    const BCol col{bo->src[c], bo->j[c], bo->w[c], bo->rows, bo->pg0[c], bo->out[c]};
                    if (col.w == 4) {
                        if (col.src == 0) bx_emit_col<4, 0, LDS>(pl, col, bp, win, cnt, ob, pt0, page_nulls);
                        else if (col.src == 1) bx_emit_col<4, 1, LDS>(pl, col, bp, win, cnt, ob, pt0, page_nulls);
                        else bx_emit_col<4, 2, LDS>(pl, col, bp, win, cnt, ob, pt0, page_nulls);
                    } else {  // 8-byte columns are never keys
                        if (col.src == 1) bx_emit_col<8, 1, LDS>(pl, col, bp, win, cnt, ob, pt0, page_nulls);
                        else bx_emit_col<8, 2, LDS>(pl, col, bp, win, cnt, ob, pt0, page_nulls);
                    }
                }
                __syncthreads();
            }
            base += run;
        }
    }
    return mine;
}

// The write pass over a partition ordered by (slot, key) (staged in LDS or
// read in place), from the count pass's match descriptors: no key lookups,
// no candidate walks.  Per tile of kWG *
// RPT probe tuples: the descriptors (the next tile's loaded while this one
// is written), a prefix of the threads' match counts (a thread's matches
// take consecutive rows; a plan's result is a row multiset, so any order of
// the tile's matches serves), then the window: entry (tile row, first + x)
// for match x of a tuple, emitted column by column.
template <uint32_t RPT, bool LDS>
__device__ __forceinline__ void bx_write_desc(const uint32_t *desc, const BPlan &pl, const BOut *bo,
                                              const BuildPart &bp, uint32_t nb, uint64_t p0, uint64_t p1,
                                              uint64_t pbase, uint64_t obase, uint32_t *lds32, uint32_t *win,
                                              uint32_t *page_nulls) {
    static_assert(RPT <= 4, "the staged probe values hold four rows a thread");
    uint64_t base = obase;
    // the probe side's carried column 0 (if any) is loaded with the
    // descriptors, a tile ahead, and staged in LDS, so emitting it reads no
    // global memory
    const bool stage = pl.p.ncar > 0;
    const bool stage8 = stage && pl.p.car_w[0] == 8;
    const CJ_G uint32_t *c32 = glb(reinterpret_cast<const uint32_t *>(pl.p.car[0]));
    const CJ_G uint8_t *cvb = glb(pl.p.car_vld[0]);
    uint32_t nd[RPT], nlo[RPT], nhi[RPT], nvb[RPT];
    auto load_tile = [&](uint64_t t0) {
#pragma unroll
        for (uint32_t j = 0; j < RPT; ++j) {
            const uint64_t idx = t0 + kWG * j + threadIdx.x;
            const uint64_t ic = idx < p1 ? idx : p1 - 1;
            nd[j] = desc[ic];
            const uint64_t pi = ic - pbase;  // plan-local index of the carried values
            nlo[j] = stage ? (stage8 ? c32[2 * pi] : c32[pi]) : 0u;
            nhi[j] = stage8 ? c32[2 * pi + 1] : 0u;
            nvb[j] = stage8 ? cvb[pi] : 0u;
        }
    };
    load_tile(p0);
    const uint32_t w = wave_id(), lane = lane_id();
    for (uint64_t i0 = p0; i0 < p1; i0 += kWG * RPT) {
        uint32_t m[RPT], f[RPT], mt = 0;
        uint32_t d_[RPT];
#pragma unroll
        for (uint32_t j = 0; j < RPT; ++j) {
            d_[j] = nd[j];
            // staged for this tile (read after the window barrier; the
            // previous tile's emission ended with a barrier)
            if (stage) {
                lds(bp.pt_lo)[kWG * j + threadIdx.x] = nlo[j];
                if (stage8) {
                    lds(bp.pt_hi)[kWG * j + threadIdx.x] = nhi[j];
                    lds(bp.pt_vb)[kWG * j + threadIdx.x] = (uint8_t)nvb[j];
                }
            }
        }
        load_tile(i0 + kWG * RPT);
#pragma unroll
        for (uint32_t j = 0; j < RPT; ++j) {
            const uint32_t d = d_[j];
            const bool in = i0 + kWG * j + threadIdx.x < p1;
            f[j] = d & kDescFirst;
            m[j] = in ? d >> kDescFirstBits : 0u;
            if (m[j] == kDescMany) {
                // a longer run: its equal keys are consecutive in the partition
                const int32_t kk = bp_key<LDS>(bp, f[j]);
                uint32_t t = f[j] + kDescMany;
                while (t < nb && bp_key<LDS>(bp, t) == kk) ++t;
                m[j] = t - f[j];
            }
            mt += m[j];
        }
        const uint32_t inc = wave_incl_scan<uint32_t>(mt);
        if (lane == 63) lds32[w] = inc;
        __syncthreads();
        uint32_t run = 0, pre = inc - mt;
#pragma unroll
        for (uint32_t ww = 0; ww < kWG / 64; ++ww) {
            const uint32_t x = lds32[ww];
            pre += ww < w ? x : 0u;
            run += x;
        }
        // (the next tile writes lds32 after this tile's window barriers, or
        // the barrier of a tile without matches: every wave has read it)
        for (uint32_t w0 = 0; w0 < run; w0 += kMatchWin) {
            const uint32_t w1 = run - w0 < kMatchWin ? run : w0 + kMatchWin;
            // the thread's entries [pre, pre + mt) that fall in [w0, w1)
            uint32_t q = pre;
#pragma unroll
            for (uint32_t j = 0; j < RPT; ++j) {
                const uint32_t row = (kWG * j + threadIdx.x) << kMatchTBits;
                const uint32_t a = q > w0 ? q : w0, b = q + m[j] < w1 ? q + m[j] : w1;
                for (uint32_t x = a; x < b; ++x) lds(win)[x - w0] = row | (f[j] + (x - q));
                q += m[j];
            }
            __syncthreads();
            const uint32_t cnt = w1 - w0;
            const uint32_t ob = (uint32_t)(base + w0), pt0 = (uint32_t)(i0 - pbase);
#pragma unroll 1
            for (uint32_t c = 0; c < bo->nout; ++c) {
                const BCol col{bo->src[c], bo->j[c], bo->w[c], bo->rows, bo->pg0[c], bo->out[c]};
                const bool staged = col.src == 2 && col.cj == 0;  // (stage holds: column 0 exists)
                if (col.w == 4) {
                    if (col.src == 0) bx_emit_col<4, 0, LDS>(pl, col, bp, win, cnt, ob, pt0, page_nulls);
                    else if (col.src == 1) bx_emit_col<4, 1, LDS>(pl, col, bp, win, cnt, ob, pt0, page_nulls);
                    else if (staged) bx_emit_col<4, 3, LDS>(pl, col, bp, win, cnt, ob, 0u, page_nulls);
                    else bx_emit_col<4, 2, LDS>(pl, col, bp, win, cnt, ob, pt0, page_nulls);
                } else {  // 8-byte columns are never keys
                    if (col.src == 1) bx_emit_col<8, 1, LDS>(pl, col, bp, win, cnt, ob, pt0, page_nulls);
                    else if (staged) bx_emit_col<8, 3, LDS>(pl, col, bp, win, cnt, ob, 0u, page_nulls);
                    else bx_emit_col<8, 2, LDS>(pl, col, bp, win, cnt, ob, pt0, page_nulls);
                }
            }
            __syncthreads();
        }
        base += run;
        // the next tile's staging overwrites the probe values: every
        // thread is past this tile's emission (the window loop ends with a
        // barrier; a tile without matches has none, so wait here)
        if (!run) __syncthreads();
    }
}

}  // namespace

// dst[0 .. n) = src[0 .. n), eight loads per thread in flight before the
// stores (clamped addresses: a load under a condition is awaited alone).
template <class T, class S>
__device__ __forceinline__ void stage_lds(T *dst, S src, uint32_t n) {
    for (uint32_t i0 = 0; i0 < n; i0 += 8 * kWG) {
        T v[8];
#pragma unroll
        for (uint32_t q = 0; q < 8; ++q) {
            const uint32_t i = i0 + q * kWG + threadIdx.x;
            v[q] = src[i < n ? i : n - 1];
        }
#pragma unroll
        for (uint32_t q = 0; q < 8; ++q) {
            const uint32_t i = i0 + q * kWG + threadIdx.x;
            if (i < n) dst[i] = v[q];
        }
    }
}

template <bool WRITE>
__global__ __launch_bounds__(kWG) void k_bx_join(BatchArgs a, const BOut *__restrict__ outs, const BPlan *__restrict__ plans,
                                                 const uint32_t *__restrict__ part_start, const uint64_t *__restrict__ uoff,
                                                 const uint64_t *__restrict__ boff, const uint64_t *__restrict__ poff,
                                                 const uint64_t *__restrict__ moff, int32_t *pkeys,
                                                 const int32_t *__restrict__ bkeys, const uint32_t *__restrict__ dir,
                                                 uint32_t *page_nulls) {
    // slot ends (u16, 2^log_part of them), then the keys and carried build values
    extern __shared__ __align__(16) uint8_t dyn[];
    uint16_t *ends = reinterpret_cast<uint16_t *>(dyn);
    __shared__ uint32_t lds32[32];
    __shared__ uint64_t red[4];
    __shared__ uint32_t unit;
    __shared__ uint32_t win[WRITE ? kMatchWin : 1];
    // write pass: the tile's values of the probe side's carried column 0
    __shared__ uint32_t ptile_lo[WRITE ? kWG * 4 : 1], ptile_hi[WRITE ? kWG * 4 : 1];
    __shared__ uint8_t ptile_vb[WRITE ? kWG * 4 : 1];
    const uint64_t total = uoff[a.nparts];  // < 2^32 (max_units)
    const uint32_t S = 1u << a.log_part;
    uint32_t *q = a.queue + (WRITE ? kQJoinWrite : kQJoinCount);
    for (uint32_t u = next_unit(a.join_queue, q, &unit, 0, true); u < total;
         u = next_unit(a.join_queue, q, &unit, u, false)) {
        const uint32_t g = find_owner<uint64_t>(uoff, a.nparts, u);
        const uint32_t k = find_owner(part_start, a.nplans, g);
        const BPlan &pl = plans[k];
        const uint64_t t_a = boff[g], t_e = boff[g + 1];
        const uint32_t nb = (uint32_t)(t_e - t_a);
        const uint64_t bbase = boff[part_start[k]], pbase = poff[part_start[k]];
        const uint64_t p0 = poff[g] + (u - uoff[g]) * a.chunk;
        const uint64_t pe = poff[g + 1];
        const uint64_t p1 = p0 + a.chunk < pe ? p0 + a.chunk : pe;
        BuildPart bp;
        bp.l_ends = ends;
        bp.l_dyn = dyn + (2u << a.log_part);
        bp.g_ends = dir + pl.dir_base + ((uint64_t)(g - part_start[k]) << a.log_part);
        bp.smask = S - 1;
        bp.g_keys = bkeys + t_a;
        bp.g_car0 = t_a - bbase;
        const uint32_t ncar = pl.b.ncar;
        bp.cap = a.lds_cap;
        bp.ncar = ncar;
        bp.w8 = 0;
        for (uint32_t c = 0; c < ncar; ++c) bp.w8 |= (pl.b.car_w[c] == 8 ? 1u : 0u) << c;
        bp.pt_lo = ptile_lo;
        bp.pt_hi = ptile_hi;
        bp.pt_vb = ptile_vb;
        uint64_t obase = 0;
        if (WRITE) obase = moff[u] - moff[uoff[part_start[k]]];
        // the plan's output columns, read through the pointer (uniform
        // addresses: scalar loads; a local copy indexed by the column loop
        // would live in scratch memory)
        const BOut *ob = WRITE ? outs + k : nullptr;
        uint64_t mine;
        if (nb <= a.lds_cap) {
            // stage the partition: each batch of loads is issued before its LDS stores
            // (the write pass reads descriptors, not slots: no slot ends)
            if (!WRITE) {
                uint32_t ev[kMaxSlots / kWG];
#pragma unroll
                for (uint32_t q = 0; q < kMaxSlots / kWG; ++q)
                    ev[q] = bp.g_ends[q * kWG + threadIdx.x < S ? q * kWG + threadIdx.x : 0];
#pragma unroll
                for (uint32_t q = 0; q < kMaxSlots / kWG; ++q)
                    if (q * kWG + threadIdx.x < S) ends[q * kWG + threadIdx.x] = (uint16_t)ev[q];
            }
            uint8_t *tup = dyn + (2u << a.log_part);  // == bp.l_dyn
            stage_lds<int32_t>(reinterpret_cast<int32_t *>(tup), bp.g_keys, nb);
            if (WRITE)
                for (uint32_t c = 0; c < ncar; ++c) {
                    if (pl.b.car_w[c] == 4) {
                        stage_lds<int32_t>(reinterpret_cast<int32_t *>(tup + bp.coff(c)),
                                           glb(reinterpret_cast<const int32_t *>(pl.bslot_car[c])) + bp.g_car0, nb);
                    } else {
                        stage_lds<uint64_t>(reinterpret_cast<uint64_t *>(tup + bp.coff(c)),
                                            glb(reinterpret_cast<const uint64_t *>(pl.bslot_car[c])) + bp.g_car0, nb);
                        stage_lds<uint8_t>(tup + bp.voff(c), glb(pl.bslot_vld[c]) + bp.g_car0, nb);
                    }
                }
            __syncthreads();
            if constexpr (WRITE) {
                bx_write_desc<4u, true>(reinterpret_cast<const uint32_t *>(pkeys), pl, ob, bp, nb, p0, p1, pbase,
                                        obase, lds32, win, page_nulls);
                mine = 0;
            } else {
                mine = bx_probe<false, true, kRPT>(pkeys, pl, ob, bp, p0, p1, pbase, obase, lds32, win, page_nulls,
                                                   reinterpret_cast<uint32_t *>(pkeys), true);
            }
        } else {
            // a partition too large for LDS is read in place; ordered by
            // key when it holds at most kBatchSortCap tuples (k_bx_build)
            const bool sorted = nb <= kBatchSortCap;
            if constexpr (WRITE) {
                if (sorted)
                    bx_write_desc<4u, false>(reinterpret_cast<const uint32_t *>(pkeys), pl, ob, bp, nb, p0, p1, pbase,
                                             obase, lds32, win, page_nulls);
                else
                    bx_probe<true, false, 4u>(pkeys, pl, ob, bp, p0, p1, pbase, obase, lds32, win, page_nulls, nullptr,
                                              false);
                mine = 0;
            } else {
                mine = bx_probe<false, false, kRPT>(pkeys, pl, ob, bp, p0, p1, pbase, obase, lds32, win, page_nulls,
                                                    reinterpret_cast<uint32_t *>(pkeys), sorted);
            }
        }
        if (!WRITE) {
            uint64_t tot;
            block_excl_scan<uint64_t>(mine, &tot, red);
            if (threadIdx.x == 0) a.mcount[u] = tot;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------ decode (dense copy)
// INT32 NULLs are written as INT32_MIN (the value the output writer reads
// as NULL); 8-byte columns also set a validity bitmap.  A thread takes 8
// consecutive rows (4 for 8-byte values): one bitmap byte (nibble), one
// block scan for the value ranks, all page loads before the dense stores.
template <uint32_t W>
__device__ __forceinline__ void bx_decode_page(const BDecode &d, const CJ_G uint8_t *p, uint64_t base,
                                               uint32_t *lds4) {
    constexpr uint32_t RPT = W == 4 ? 8 : 4;
    const uint32_t n = ld16(p), nv = ld16(p + 2);
    const CJ_G uint8_t *bm = p + kPage - ((n + 7) >> 3);
    const bool dense = n == nv;
    const uint32_t blast = n ? (n - 1) >> 3 : 0, vlast = nv ? nv - 1 : 0;
    uint32_t running = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += kWG * RPT) {
        const uint32_t i0 = c0 + RPT * threadIdx.x;
        const uint32_t in = i0 >= n ? 0u : (n - i0 >= RPT ? (1u << RPT) - 1 : (1u << (n - i0)) - 1);
        uint32_t bits = in;
        if (!dense) {
            const uint32_t byte = bm[(i0 >> 3) < blast ? (i0 >> 3) : blast];
            bits &= W == 4 ? byte : (byte >> (i0 & 4)) & 15u;
        }
        uint32_t tot;
        const uint32_t pre = running + block_excl_scan<uint32_t>(__popc(bits), &tot, lds4);
        if constexpr (W == 4) {
            int32_t v[RPT];
#pragma unroll
            for (uint32_t q = 0; q < RPT; ++q) {
                const uint32_t vi = pre + __popc(bits & ((1u << q) - 1));
                const int32_t x = ld32(p + 4 + 4 * (size_t)(vi < nv ? vi : vlast));
                v[q] = ((bits >> q) & 1u) && vi < nv ? x : kI32Null;
            }
            CJ_G int32_t *out = glb(static_cast<int32_t *>(d.values)) + base + i0;
#pragma unroll
            for (uint32_t q = 0; q < RPT; ++q)
                if ((in >> q) & 1u) out[q] = v[q];
        } else {
            uint64_t v[RPT];
            uint32_t ok = 0;
#pragma unroll
            for (uint32_t q = 0; q < RPT; ++q) {
                const uint32_t vi = pre + __popc(bits & ((1u << q) - 1));
                v[q] = ld64(p + 8 + 8 * (size_t)(vi < nv ? vi : vlast));
                ok |= (((bits >> q) & 1u) && vi < nv ? 1u : 0u) << q;
            }
            CJ_G uint64_t *out = glb(static_cast<uint64_t *>(d.values)) + base + i0;
#pragma unroll
            for (uint32_t q = 0; q < RPT; ++q)
                if ((in >> q) & 1u) out[q] = (ok >> q) & 1u ? v[q] : 0;
            if (ok) {
                const uint64_t r = base + i0;
                const uint32_t sh = (uint32_t)(r & 31);
                CJ_G uint32_t *vw = glb(d.valid) + (r >> 5);
                __hip_atomic_fetch_or(vw, ok << sh, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (sh + RPT > 32 && (ok >> (32 - sh)))
                    __hip_atomic_fetch_or(vw + 1, ok >> (32 - sh), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        running += tot;
    }
}

__global__ __launch_bounds__(kWG) void k_bx_decode(const BDecode *__restrict__ jobs, const uint32_t *__restrict__ page_start, uint32_t njobs,
                                                   uint32_t units) {
    __shared__ uint32_t lds4[4];
    for (uint32_t g = blockIdx.x; g < units; g += gridDim.x) {
        const uint32_t j = find_owner(page_start, njobs, g);
        const BDecode d = jobs[j];
        const uint32_t pg = g - page_start[j];
        const CJ_G uint8_t *p = glb(d.pages) + (size_t)pg * kPage;
        const uint64_t base = glb(d.rs)[pg];
        if (d.width == 4) bx_decode_page<4>(d, p, base, lds4);
        else bx_decode_page<8>(d, p, base, lds4);
        __syncthreads();
    }
}

// ---------------------------------------------------------- output pages
// One wave per output page: header (rows, rows: every value present) and an
// all-ones validity bitmap at the page end (page_builders.h:92-151 layout),
// NULL counter zero.  The write pass then stores each value at its slot.
__global__ __launch_bounds__(kWG) void k_bx_page_init(const BPageCol *__restrict__ cols,
                                                      const uint32_t *__restrict__ page_start, uint32_t ncols,
                                                      uint32_t pages, uint32_t *page_nulls) {
    const uint32_t lane = lane_id();
    OwnerCache oc;
    uint32_t mc = ~0u;  // the column whose descriptor m holds
    BPageCol m{};
    for (uint32_t g = blockIdx.x * (kWG / 64) + wave_id(); g < pages; g += gridDim.x * (kWG / 64)) {
        const uint32_t c = owner_of(oc, page_start, ncols, g);
        if (c != mc) {
            m = cols[c];
            mc = c;
        }
        const uint32_t R = m.width == 4 ? kRowsI32 : kRowsI64;
        const uint64_t r0 = (uint64_t)(g - oc.lo) * R;  // (oc.lo = page_start[c])
        const uint32_t n = (uint32_t)(m.rows - r0 < R ? m.rows - r0 : R);
        CJ_G uint8_t *p = glb(m.out) + (uint64_t)(g - oc.lo) * kPage;
        // the bitmap ends at the page end: its last byte holds the last rows;
        // whole 4-byte words from the first aligned one, bytes before it
        const uint32_t nbytes = (n + 7) >> 3, b0 = kPage - nbytes, a0 = (b0 + 3) & ~3u;
        const uint32_t tail = n & 7;  // rows of the last byte (0: a full byte)
        for (uint32_t w = a0 / 4 + lane; w < kPage / 4; w += 64) {
            uint32_t v = ~0u;
            if (w == kPage / 4 - 1 && tail) v = 0x00ffffffu | (((1u << tail) - 1u) << 24);
            reinterpret_cast<CJ_G uint32_t *>(p)[w] = v;
        }
        if (lane < a0 - b0)  // (<= 3 bytes; the page's last byte only when no whole word is left)
            p[b0 + lane] = (uint8_t)(b0 + lane == kPage - 1 && tail ? (1u << tail) - 1u : 0xffu);
        if (lane == 0) {
            *reinterpret_cast<CJ_G uint32_t *>(p) = n | (n << 16);
            page_nulls[g] = 0;
        }
    }
}

// A page whose rows include NULLs: the write pass left every value at its
// row's slot and cleared the NULL rows' bits; pack the present values in row
// order (in place: all loads before a barrier, then the stores) and set the
// header's value count.  One workgroup per page, 8 (4) rows per thread.
__global__ __launch_bounds__(kWG) void k_bx_page_fix(const BPageCol *__restrict__ cols,
                                                     const uint32_t *__restrict__ page_start, uint32_t ncols,
                                                     uint32_t pages, const uint32_t *__restrict__ page_nulls) {
    __shared__ uint32_t lds4[4];
    __shared__ uint32_t flagged[kWG], nflag;
    // the NULL counts of kWG pages are read at once (one per thread) and the
    // pages that have any are listed; the workgroup then packs them one by one
    for (uint32_t base = blockIdx.x * kWG; base < pages; base += gridDim.x * kWG) {
        if (threadIdx.x == 0) nflag = 0;
        __syncthreads();
        const uint32_t gt = base + threadIdx.x;
        if (gt < pages && page_nulls[gt]) flagged[atomicAdd(&nflag, 1u)] = gt;
        __syncthreads();
        const uint32_t nf = nflag;
        for (uint32_t fi = 0; fi < nf; ++fi) {
            const uint32_t g = flagged[fi];
            const uint32_t c = find_owner(page_start, ncols, g);
            const BPageCol m = cols[c];
            const uint32_t R = m.width == 4 ? kRowsI32 : kRowsI64;
            const uint64_t r0 = (uint64_t)(g - page_start[c]) * R;
            const uint32_t n = (uint32_t)(m.rows - r0 < R ? m.rows - r0 : R);
            CJ_G uint8_t *p = glb(m.out) + (uint64_t)(g - page_start[c]) * kPage;
            const CJ_G uint8_t *bm = p + kPage - ((n + 7) >> 3);
            uint32_t tot;
            if (m.width == 4) {
                const uint32_t i0 = 8 * threadIdx.x;
                uint32_t bits = i0 < n ? bm[i0 >> 3] : 0u;  // 8 rows a thread: one bitmap byte
                CJ_G int32_t *vals = reinterpret_cast<CJ_G int32_t *>(p + 4);
                int32_t v[8];
#pragma unroll
                for (uint32_t q = 0; q < 8; ++q) v[q] = vals[i0 + q < n ? i0 + q : 0];
                uint32_t o = block_excl_scan<uint32_t>(__popc(bits), &tot, lds4);  // its barrier orders loads before stores
#pragma unroll
                for (uint32_t q = 0; q < 8; ++q)
                    if ((bits >> q) & 1u) vals[o++] = v[q];
            } else {
                const uint32_t i0 = 4 * threadIdx.x;
                uint32_t bits = i0 < n ? (bm[i0 >> 3] >> (i0 & 4)) & 15u : 0u;
                CJ_G uint64_t *vals = reinterpret_cast<CJ_G uint64_t *>(p + 8);
                uint64_t v[4];
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q) v[q] = vals[i0 + q < n ? i0 + q : 0];
                uint32_t o = block_excl_scan<uint32_t>(__popc(bits), &tot, lds4);
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q)
                    if ((bits >> q) & 1u) vals[o++] = v[q];
            }
            if (threadIdx.x == 0) *reinterpret_cast<CJ_G uint16_t *>(p + 2) = (uint16_t)tot;
            __syncthreads();
        }
        __syncthreads();  // flagged and nflag are rewritten for the next pages
    }
}

__global__ void k_bx_totals(BatchArgs a, uint64_t *totals) {
    const uint32_t k = blockIdx.x * kWG + threadIdx.x;
    if (k < a.nplans) totals[k] = a.moff[a.uoff[a.part_start[k + 1]]] - a.moff[a.uoff[a.part_start[k]]];
}

__global__ void k_copy_u64(const uint64_t *src, uint64_t *dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)kWG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kWG) dst[i] = src[i];
}

// -------------------------------------------------------------- launchers
static unsigned pgrid(uint64_t units, uint32_t grid) { return (unsigned)(units < grid ? units : grid); }
// a scatter grid: a multiple of 8 where it has 8 workgroups or more (xcd_units)
static unsigned sgrid(uint64_t units, uint32_t grid) {
    const unsigned g = pgrid(units, grid);
    return g >= 8 ? g & ~7u : g;
}

// dynamic LDS of each kernel
static size_t part_count_lds(const BatchArgs &a) { return a.max_parts * sizeof(uint32_t); }
static size_t scatter_lds(uint32_t parts) { return parts * (sizeof(uint64_t) + sizeof(uint32_t)); }
// keys, source tuples and swap lists of the partitions a build launch orders
static size_t build_lds(const BatchArgs &a) { return (size_t)a.lds_cap * (sizeof(int32_t) + 2 * sizeof(uint16_t)); }
static size_t build_big_lds() { return (size_t)kBatchSortCap * (sizeof(int32_t) + 2 * sizeof(uint16_t)); }
static size_t join_count_lds(const BatchArgs &a) { return (2u << a.log_part) + (size_t)a.lds_cap * 4; }  // slot ends + keys
static size_t join_write_lds(const BatchArgs &a) {
    return (2u << a.log_part) + (size_t)a.lds_cap * a.tuple_bytes;  // slot ends + tuples
}

template <class K>
static uint32_t resident(K kernel, size_t dyn, int cus) {
    int per = 0;
    CJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kWG, dyn));
    return (uint32_t)(per > 0 ? per : 1) * (uint32_t)cus;
}

// A kernel whose static + dynamic LDS must fit a workgroup's share of the
// CU; dynamic LDS above 64 KiB needs the kernel attribute raised.
template <class K>
static void allow_lds(K kernel, size_t dyn, int device, const char *what) {
    hipFuncAttributes fa;
    CJ_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(kernel)));
    int max_lds = 0;
    CJ_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    if (fa.sharedSizeBytes + dyn > (size_t)max_lds)
        throw UnsupportedError(std::string("batch configuration: ") + what +
                               " needs more LDS than a workgroup has (too many partitions for 2^log_part slots)");
    if (dyn > 64 * 1024)
        CJ_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
}

BatchGrids batch_resident_grids(const BatchArgs &a, int device) {
    allow_lds(k_bx_part_count<true>, part_count_lds(a), device, "the partition histogram");
    allow_lds(k_bx_part_count<false>, part_count_lds(a), device, "the partition histogram");
    allow_lds(k_bx_part_scatter<true>, scatter_lds(a.staged_parts), device, "the partition scatter");
    allow_lds(k_bx_part_scatter<false>, scatter_lds(a.staged_parts), device, "the partition scatter");
    allow_lds(k_bx_part_scatter_direct<true>, scatter_lds(a.direct_parts), device, "the partition scatter");
    allow_lds(k_bx_part_scatter_direct<false>, scatter_lds(a.direct_parts), device, "the partition scatter");
    allow_lds(k_bx_build, std::max(build_lds(a), build_big_lds()), device, "the partition build");
    allow_lds(k_bx_join<true>, join_write_lds(a), device, "the join");
    allow_lds(k_bx_join<false>, join_count_lds(a), device, "the join");
    int cus = 0;
    CJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    BatchGrids g;
    g.decode = resident(k_bx_decode, 0, cus);
    g.part_count = resident(k_bx_part_count<true>, part_count_lds(a), cus);
    g.part_scatter = resident(k_bx_part_scatter<true>, scatter_lds(a.staged_parts), cus);
    g.part_scatter_direct = resident(k_bx_part_scatter_direct<true>, scatter_lds(a.direct_parts), cus);
    g.build = resident(k_bx_build, build_lds(a), cus);
    g.build_big = resident(k_bx_build, build_big_lds(), cus);
    g.join_count = resident(k_bx_join<false>, join_count_lds(a), cus);
    g.join_write = resident(k_bx_join<true>, join_write_lds(a), cus);
    g.page_init = resident(k_bx_page_init, 0, cus);
    g.page_fix = resident(k_bx_page_fix, 0, cus);
    return g;
}

void batch_part_count(const BatchArgs &a, bool probe, uint32_t units, uint32_t grid, hipStream_t s) {
    if (!units) return;
    const size_t sh = part_count_lds(a);
    if (probe) k_bx_part_count<true><<<pgrid(units, grid), kWG, sh, s>>>(a, units, a.plans, a.ppage_start, a.part_start);
    else k_bx_part_count<false><<<pgrid(units, grid), kWG, sh, s>>>(a, units, a.plans, a.bpage_start, a.part_start);
    CJ_HIP(hipGetLastError());
}

void batch_part_scatter(const BatchArgs &a, bool probe, const uint32_t *pstart, uint32_t units, bool staged,
                        uint32_t grid, hipStream_t s) {
    if (!units) return;
    const uint32_t parts = staged ? a.staged_parts : a.direct_parts;
    const size_t sh = scatter_lds(parts);
    const uint64_t *off = probe ? a.poff : a.boff;
    if (staged) {
        if (probe) k_bx_part_scatter<true><<<sgrid(units, grid), kWG, sh, s>>>(a, units, a.plans, pstart, a.part_start, off, parts);
        else k_bx_part_scatter<false><<<sgrid(units, grid), kWG, sh, s>>>(a, units, a.plans, pstart, a.part_start, off, parts);
    } else {
        if (probe)
            k_bx_part_scatter_direct<true><<<sgrid(units, grid), kWG, sh, s>>>(a, units, a.plans, pstart, a.part_start, off, parts);
        else
            k_bx_part_scatter_direct<false><<<sgrid(units, grid), kWG, sh, s>>>(a, units, a.plans, pstart, a.part_start, off,
                                                                                 parts);
    }
    CJ_HIP(hipGetLastError());
}

void batch_build(const BatchArgs &a, uint32_t grid, uint32_t grid_big, hipStream_t s) {
    if (!a.nparts) return;
    k_bx_build<<<pgrid(a.nparts, grid), kWG, build_lds(a), s>>>(a, a.plans, a.part_start, a.boff, a.bstage_keys,
                                                                 a.lds_cap, false);
    k_bx_build<<<pgrid(a.nparts, grid_big), kWG, build_big_lds(), s>>>(a, a.plans, a.part_start, a.boff,
                                                                        a.bstage_keys, kBatchSortCap, true);
    CJ_HIP(hipGetLastError());
}

void batch_units(const BatchArgs &a, hipStream_t s) {
    if (!a.nparts) return;
    k_bx_units<<<(a.nparts + kWG - 1) / kWG, kWG, 0, s>>>(a);
    CJ_HIP(hipGetLastError());
}

void batch_join_count(const BatchArgs &a, uint32_t grid, hipStream_t s) {
    if (!a.max_units) return;
    const size_t sh = join_count_lds(a);
    k_bx_join<false><<<pgrid(a.max_units, grid), kWG, sh, s>>>(a, nullptr, a.plans, a.part_start, a.uoff, a.boff,
                                                                        a.poff, a.moff, a.pkeys, a.bkeys, a.dir, nullptr);
    CJ_HIP(hipGetLastError());
}

void batch_join_write(const BatchArgs &a, const BOut *outs, uint32_t *page_nulls, uint32_t grid, hipStream_t s) {
    if (!a.max_units) return;
    const size_t sh = join_write_lds(a);
    k_bx_join<true><<<pgrid(a.max_units, grid), kWG, sh, s>>>(a, outs, a.plans, a.part_start, a.uoff, a.boff,
                                                                       a.poff, a.moff, a.pkeys, a.bkeys, a.dir,
                                                                       page_nulls);
    CJ_HIP(hipGetLastError());
}

void batch_decode(const BDecode *jobs, const uint32_t *page_start, uint32_t njobs, uint32_t units, uint32_t grid,
                  hipStream_t s) {
    if (!units || !njobs) return;
    k_bx_decode<<<pgrid(units, grid), kWG, 0, s>>>(jobs, page_start, njobs, units);
    CJ_HIP(hipGetLastError());
}

void batch_page_init(const BPageCol *cols, const uint32_t *page_start, uint32_t ncols, uint32_t pages,
                     uint32_t *page_nulls, uint32_t grid, hipStream_t s) {
    if (!pages || !ncols) return;
    k_bx_page_init<<<pgrid((pages + kWG / 64 - 1) / (kWG / 64), grid), kWG, 0, s>>>(cols, page_start, ncols, pages,
                                                                                  page_nulls);
    CJ_HIP(hipGetLastError());
}

void batch_page_fix(const BPageCol *cols, const uint32_t *page_start, uint32_t ncols, uint32_t pages,
                    const uint32_t *page_nulls, uint32_t grid, hipStream_t s) {
    if (!pages || !ncols) return;
    k_bx_page_fix<<<pgrid((pages + kWG - 1) / kWG, grid), kWG, 0, s>>>(cols, page_start, ncols, pages, page_nulls);
    CJ_HIP(hipGetLastError());
}

void batch_totals(const BatchArgs &a, uint64_t *totals, hipStream_t s) {
    if (!a.nplans) return;
    k_bx_totals<<<(a.nplans + kWG - 1) / kWG, kWG, 0, s>>>(a, totals);
    CJ_HIP(hipGetLastError());
}

void batch_copy_u64(const uint64_t *src, uint64_t *dst, uint64_t n, hipStream_t s) {
    if (!n) return;
    const uint64_t g = (n + kWG - 1) / kWG;
    k_copy_u64<<<(unsigned)(g < 4096 ? g : 4096), kWG, 0, s>>>(src, dst, n);
    CJ_HIP(hipGetLastError());
}

}  // namespace cj
