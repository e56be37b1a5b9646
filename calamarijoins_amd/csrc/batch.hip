// batch.hip -- persistent work-list kernels of the batched join (batch.hpp).
//
// The per-plan algorithm is the reference's (select build side, radix
// partition the build keys, build each partition's slice of the unchained
// directory with Bloom tags, probe, materialize: hashtable.h:197-246,405-505,
// hash_join.h:156-244, materialize.h:85-186); what changes is that one
// launch serves every plan of the batch.  A workgroup takes global units
// blockIdx.x, blockIdx.x + gridDim.x, ... and finds each unit's plan by a
// binary search over the plans' unit prefix (a few hundred entries, cached).
#include "batch.hpp"
#include "probe8.hpp"

namespace cj {

__device__ __forceinline__ uint64_t bslot_of(uint64_t h, uint32_t log_slots) { return h >> (64 - log_slots); }

// Largest k < n with start[k] <= g (plans with no units share a start).
__device__ __forceinline__ uint32_t find_unit_owner(const uint32_t *start, uint32_t n, uint32_t g) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (start[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------ radix partition of builds
// Partition of a key = top bits of its directory slot (a plan's partitions
// own contiguous slot ranges).  NULL build keys are not inserted
// (build_columnar, hashtable.h:439-478).
__global__ __launch_bounds__(kWG) void k_batch_part_count(BatchArgs a, uint32_t units) {
    extern __shared__ uint32_t hist[];  // max_parts
    __shared__ uint32_t lds32[32];
    for (uint32_t g = blockIdx.x; g < units; g += gridDim.x) {
        const uint32_t k = find_unit_owner(a.bpage_start, a.nplans, g);
        const BPlan pl = a.plans[k];
        const uint32_t P = 1u << (pl.log_slots - pl.log_part);
        for (uint32_t i = threadIdx.x; i < P; i += kWG) hist[i] = 0;
        __syncthreads();
        const uint8_t *p = pl.bk_pages + (size_t)(g - a.bpage_start[k]) * kPage;
        const uint32_t n = ld16(p), nv = ld16(p + 2);
        uint32_t running = 0;
        for (uint32_t t0 = 0; t0 < n; t0 += kTile) {
            int32_t key[kRPT];
            const uint32_t valid = load_keys8(p, n, nv, n == nv, t0, &running, key, lds32);
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j)
                if ((valid >> j) & 1u) atomicAdd(&hist[bslot_of(hash_key(key[j]), pl.log_slots) >> pl.log_part], 1u);
        }
        __syncthreads();
        const uint32_t pb = a.part_start[k];
        for (uint32_t i = threadIdx.x; i < P; i += kWG)
            if (hist[i]) atomicAdd((unsigned long long *)&a.part_count[pb + i], (unsigned long long)hist[i]);
        __syncthreads();
    }
}

__global__ __launch_bounds__(kWG) void k_batch_part_scatter(BatchArgs a, uint32_t units) {
    extern __shared__ uint32_t sm[];  // hist[max_parts] then cursors (u64)
    uint32_t *hist = sm;
    uint64_t *cur = reinterpret_cast<uint64_t *>(sm + ((a.max_parts + 1) & ~1u));
    __shared__ uint32_t lds32[32];
    for (uint32_t g = blockIdx.x; g < units; g += gridDim.x) {
        const uint32_t k = find_unit_owner(a.bpage_start, a.nplans, g);
        const BPlan pl = a.plans[k];
        const uint32_t P = 1u << (pl.log_slots - pl.log_part);
        const uint32_t page = g - a.bpage_start[k];
        for (uint32_t i = threadIdx.x; i < P; i += kWG) hist[i] = 0;
        __syncthreads();
        const uint8_t *p = pl.bk_pages + (size_t)page * kPage;
        const uint32_t n = ld16(p), nv = ld16(p + 2);
        uint32_t running = 0;
        for (uint32_t t0 = 0; t0 < n; t0 += kTile) {
            int32_t key[kRPT];
            const uint32_t valid = load_keys8(p, n, nv, n == nv, t0, &running, key, lds32);
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j)
                if ((valid >> j) & 1u) atomicAdd(&hist[bslot_of(hash_key(key[j]), pl.log_slots) >> pl.log_part], 1u);
        }
        __syncthreads();
        const uint32_t pb = a.part_start[k];
        for (uint32_t i = threadIdx.x; i < P; i += kWG)
            cur[i] = hist[i] ? atomicAdd((unsigned long long *)&a.part_cur[pb + i], (unsigned long long)hist[i]) : 0;
        __syncthreads();
        const uint64_t r0 = pl.bk_rs[page];
        running = 0;
        for (uint32_t t0 = 0; t0 < n; t0 += kTile) {
            int32_t key[kRPT];
            const uint32_t valid = load_keys8(p, n, nv, n == nv, t0, &running, key, lds32);
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j)
                if ((valid >> j) & 1u) {
                    const uint32_t part = (uint32_t)(bslot_of(hash_key(key[j]), pl.log_slots) >> pl.log_part);
                    const uint64_t pos = atomicAdd((unsigned long long *)&cur[part], 1ull);
                    a.stage_keys[pos] = key[j];
                    a.stage_rids[pos] = (uint32_t)(r0 + t0 + kWG * j + threadIdx.x);
                }
        }
        __syncthreads();
    }
}

// --------------------------------------------- per-partition build in LDS
// build_partition (hashtable.h:197-246): count per slot, exclusive prefix,
// place keys/row ids, OR Bloom tags, write (end << 16) | tag with the end as
// a global tuple index.
__global__ __launch_bounds__(kWG) void k_batch_build(BatchArgs a, uint32_t nparts) {
    extern __shared__ uint32_t sm[];  // cnt[4096], tag[4096]
    __shared__ uint32_t lds4[4];
    for (uint32_t g = blockIdx.x; g < nparts; g += gridDim.x) {
        const uint32_t k = find_unit_owner(a.part_start, a.nplans, g);
        const BPlan pl = a.plans[k];
        const uint32_t S = 1u << pl.log_part, smask = S - 1;
        uint32_t *cnt = sm, *tag = sm + S;
        const uint64_t t_a = a.part_off[g], t_e = a.part_off[g + 1];
        for (uint32_t i = threadIdx.x; i < S; i += kWG) {
            cnt[i] = 0;
            tag[i] = 0;
        }
        __syncthreads();
        for (uint64_t i = t_a + threadIdx.x; i < t_e; i += kWG) {
            const uint64_t h = hash_key(a.stage_keys[i]);
            const uint32_t ls = (uint32_t)bslot_of(h, pl.log_slots) & smask;
            atomicAdd(&cnt[ls], 1u);
            atomicOr(&tag[ls], (uint32_t)c_bloom.t[h & 0x7FF]);
        }
        __syncthreads();
        const uint32_t per = (S + kWG - 1) / kWG;
        const uint32_t s0 = threadIdx.x * per;
        uint32_t local = 0;
        for (uint32_t i = 0; i < per; ++i)
            if (s0 + i < S) local += cnt[s0 + i];
        uint32_t tot;
        uint32_t pre = block_excl_scan<uint32_t>(local, &tot, lds4);
        for (uint32_t i = 0; i < per; ++i)
            if (s0 + i < S) {
                const uint32_t c = cnt[s0 + i];
                cnt[s0 + i] = pre;
                pre += c;
            }
        __syncthreads();
        for (uint64_t i = t_a + threadIdx.x; i < t_e; i += kWG) {
            const int32_t key = a.stage_keys[i];
            const uint32_t ls = (uint32_t)bslot_of(hash_key(key), pl.log_slots) & smask;
            const uint64_t pos = t_a + atomicAdd(&cnt[ls], 1u);
            a.keys[pos] = key;
            a.rids[pos] = a.stage_rids[i];
        }
        __syncthreads();
        // directory ends are plan-local tuple indices (the plan's tuples start
        // at tb), so slot 0 of every plan starts at its own first tuple
        const uint64_t tb = a.part_off[a.part_start[k]];
        uint64_t *d = a.dir + pl.dir_base + (uint64_t)(g - a.part_start[k]) * S;
        for (uint32_t i = threadIdx.x; i < S; i += kWG)
            d[i] = ((t_a - tb + cnt[i]) << 16) | (uint64_t)(tag[i] & 0xFFFF);
        __syncthreads();
    }
}

// ------------------------------------------------------------------ probe
__global__ __launch_bounds__(kWG) void k_batch_probe_count(BatchArgs a, uint32_t units) {
    __shared__ uint32_t lds32[32];
    __shared__ uint64_t red[4];
    const NLKeys nl{};
    const DirectTable dt{};
    for (uint32_t g = blockIdx.x; g < units; g += gridDim.x) {
        const uint32_t k = find_unit_owner(a.ppage_start, a.nplans, g);
        const BPlan pl = a.plans[k];
        const uint64_t tb = a.part_off[a.part_start[k]];
        const HashTableDev ht{a.dir + pl.dir_base, a.keys + tb, a.rids + tb, (int)pl.log_slots, 0};
        const uint8_t *p = pl.pk_pages + (size_t)(g - a.ppage_start[k]) * kPage;
        const uint32_t n = ld16(p), nv = ld16(p + 2);
        uint32_t running = 0;
        uint64_t mine = 0;
        for (uint32_t t0 = 0; t0 < n; t0 += kTile) {
            int32_t key[kRPT];
            const uint32_t valid = load_keys8(p, n, nv, n == nv, t0, &running, key, lds32);
            uint32_t m[kRPT], s[kRPT], e[kRPT];
            lookup8<kModeHash>(valid, key, ht, nl, dt, m, s, e);
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) mine += m[j];
        }
        uint64_t tot;
        block_excl_scan<uint64_t>(mine, &tot, red);
        if (threadIdx.x == 0) a.pcount[g] = tot;
        __syncthreads();
    }
}

// Exclusive prefix (row order t0 + 256 j + t) of 8 per-row counts.
__device__ __forceinline__ uint32_t batch_scan8(const uint32_t cnt[kRPT], uint32_t pre[kRPT], uint32_t *lds32) {
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

__global__ __launch_bounds__(kWG) void k_batch_probe_write(BatchArgs a, uint32_t units) {
    __shared__ uint32_t lds32[32];
    const NLKeys nl{};
    const DirectTable dt{};
    for (uint32_t g = blockIdx.x; g < units; g += gridDim.x) {
        uint64_t base = a.poff[g];
        const uint64_t end = a.poff[g + 1];
        if (base == end) continue;
        const uint32_t k = find_unit_owner(a.ppage_start, a.nplans, g);
        const BPlan pl = a.plans[k];
        const uint64_t tb = a.part_off[a.part_start[k]];
        const HashTableDev ht{a.dir + pl.dir_base, a.keys + tb, a.rids + tb, (int)pl.log_slots, 0};
        const uint32_t page = g - a.ppage_start[k];
        const uint8_t *p = pl.pk_pages + (size_t)page * kPage;
        const uint32_t n = ld16(p), nv = ld16(p + 2);
        const uint64_t r0 = pl.pk_rs[page];
        uint32_t running = 0;
        for (uint32_t t0 = 0; t0 < n && base < end; t0 += kTile) {
            int32_t key[kRPT];
            const uint32_t valid = load_keys8(p, n, nv, n == nv, t0, &running, key, lds32);
            uint32_t m[kRPT], s[kRPT], e[kRPT], pre[kRPT];
            lookup8<kModeHash>(valid, key, ht, nl, dt, m, s, e);
            const uint32_t tot = batch_scan8(m, pre, lds32);
#pragma unroll
            for (uint32_t j = 0; j < kRPT; ++j) {
                if (!m[j]) continue;
                const uint32_t prid = (uint32_t)(r0 + t0 + kWG * j + threadIdx.x);
                uint64_t o = base + pre[j];
                for (uint32_t i = s[j]; i < e[j]; ++i)
                    if (ht.keys[i] == key[j]) {
                        if (o < end) {
                            a.out_b[o] = ht.rids[i];
                            a.out_p[o] = prid;
                        }
                        ++o;
                    }
            }
            base += tot;
        }
    }
}

// ------------------------------------------------------ decode (dense copy)
__global__ __launch_bounds__(kWG) void k_batch_decode(const BDecode *jobs, const uint32_t *page_start, uint32_t njobs,
                                                      uint32_t units) {
    __shared__ uint32_t lds4[4];
    for (uint32_t g = blockIdx.x; g < units; g += gridDim.x) {
        const uint32_t j = find_unit_owner(page_start, njobs, g);
        const BDecode d = jobs[j];
        const uint32_t pg = g - page_start[j];
        const uint8_t *p = d.pages + (size_t)pg * kPage;
        const uint32_t n = ld16(p), nv = ld16(p + 2);
        const uint64_t base = d.rs[pg];
        const uint8_t *bm = p + kPage - ((n + 7) >> 3);
        const bool dense = n == nv;
        const uint32_t nlast = n ? n - 1 : 0, vlast = nv ? nv - 1 : 0;
        uint32_t running = 0;
        for (uint32_t t0 = 0; t0 < n; t0 += kWG) {
            const uint32_t r = t0 + threadIdx.x;
            const bool inr = r < n;
            const uint32_t rc = inr ? r : nlast;
            const bool ok = inr && (dense || page_bit(bm, rc));
            uint64_t m;
            uint32_t tot;
            const uint32_t vi = running + tile_rank(ok, &m, &tot, lds4);
            const bool rd = ok && vi < nv;
            const uint32_t vc = vi < nv ? vi : vlast;
            if (d.width == 4) {
                const int32_t v = ld32(p + 4 + 4 * (size_t)vc);
                if (inr) static_cast<int32_t *>(d.values)[base + r] = rd ? v : 0;
            } else {
                const uint64_t v = ld64(p + 8 + 8 * (size_t)vc);
                if (inr) static_cast<uint64_t *>(d.values)[base + r] = rd ? v : 0;
            }
            if (lane_id() == 0) or_mask_bits(d.valid, base + t0 + 64 * wave_id(), m);
            running += tot;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ materialize
// Output page k of a column holds result rows [k R, (k + 1) R), R = 1984
// (INT32) or 1007 (8-byte): gather by row id, rank the non-NULL rows.
__global__ __launch_bounds__(kWG) void k_batch_materialize(const BMat *cols, const uint32_t *page_start,
                                                           uint32_t ncols, uint32_t units) {
    __shared__ uint32_t lds4[4];
    for (uint32_t g = blockIdx.x; g < units; g += gridDim.x) {
        const uint32_t c = find_unit_owner(page_start, ncols, g);
        const BMat m = cols[c];
        const uint32_t R = m.width == 4 ? kRowsI32 : kRowsI64;
        const uint64_t pg = g - page_start[c];
        const uint64_t r0 = pg * R;
        const uint32_t cnt = (uint32_t)(m.rows - r0 < R ? m.rows - r0 : R);
        uint8_t *p = m.out + pg * kPage;
        const uint32_t nbytes = (cnt + 7) >> 3;
        uint8_t *bm = p + kPage - nbytes;
        uint32_t running = 0;
        for (uint32_t t0 = 0; t0 < cnt; t0 += kWG) {
            const uint32_t i = t0 + threadIdx.x;
            const uint64_t ri = r0 + (i < cnt ? i : cnt - 1);
            const uint32_t rid = m.ids[ri];
            bool ok = i < cnt && valid_bit(m.valid, rid);
            uint64_t v;
            if (m.width == 4) {
                const int32_t x = static_cast<const int32_t *>(m.values)[rid];
                if (m.is_int32 && x == kI32Null) ok = false;  // page_builders.h:123
                v = (uint32_t)x;
            } else {
                v = static_cast<const uint64_t *>(m.values)[rid];
            }
            uint64_t mk;
            uint32_t tot;
            const uint32_t vi = running + tile_rank(ok, &mk, &tot, lds4);
            if (ok) {
                if (m.width == 4) *reinterpret_cast<uint32_t *>(p + 4 + 4 * (size_t)vi) = (uint32_t)v;
                else *reinterpret_cast<uint64_t *>(p + 8 + 8 * (size_t)vi) = v;
            }
            const uint32_t lane = lane_id();
            if (lane < 8) {
                const uint32_t b = ((t0 + 64 * wave_id()) >> 3) + lane;
                if (b < nbytes) bm[b] = (uint8_t)(mk >> (8 * lane));
            }
            running += tot;
        }
        if (threadIdx.x == 0) {
            *reinterpret_cast<uint16_t *>(p) = (uint16_t)cnt;
            *reinterpret_cast<uint16_t *>(p + 2) = (uint16_t)running;
        }
        __syncthreads();
    }
}

// -------------------------------------------------------------- launchers
static unsigned pgrid(uint32_t units, uint32_t grid) { return units < grid ? units : grid; }

void batch_part_count(const BatchArgs &a, uint32_t units, uint32_t grid, hipStream_t s) {
    if (!units) return;
    k_batch_part_count<<<pgrid(units, grid), kWG, a.max_parts * sizeof(uint32_t), s>>>(a, units);
    CJ_HIP(hipGetLastError());
}

void batch_part_scatter(const BatchArgs &a, uint32_t units, uint32_t grid, hipStream_t s) {
    if (!units) return;
    const size_t sh = ((a.max_parts + 1) & ~1u) * sizeof(uint32_t) + a.max_parts * sizeof(uint64_t);
    k_batch_part_scatter<<<pgrid(units, grid), kWG, sh, s>>>(a, units);
    CJ_HIP(hipGetLastError());
}

void batch_build(const BatchArgs &a, uint32_t nparts, uint32_t grid, hipStream_t s) {
    if (!nparts) return;
    const size_t sh = 2 * sizeof(uint32_t) << kBatchMaxLogPart;
    k_batch_build<<<pgrid(nparts, grid), kWG, sh, s>>>(a, nparts);
    CJ_HIP(hipGetLastError());
}

void batch_probe_count(const BatchArgs &a, uint32_t units, uint32_t grid, hipStream_t s) {
    if (!units) return;
    k_batch_probe_count<<<pgrid(units, grid), kWG, 0, s>>>(a, units);
    CJ_HIP(hipGetLastError());
}

void batch_probe_write(const BatchArgs &a, uint32_t units, uint32_t grid, hipStream_t s) {
    if (!units) return;
    k_batch_probe_write<<<pgrid(units, grid), kWG, 0, s>>>(a, units);
    CJ_HIP(hipGetLastError());
}

void batch_decode(const BDecode *jobs, const uint32_t *page_start, uint32_t njobs, uint32_t units, uint32_t grid,
                  hipStream_t s) {
    if (!units || !njobs) return;
    k_batch_decode<<<pgrid(units, grid), kWG, 0, s>>>(jobs, page_start, njobs, units);
    CJ_HIP(hipGetLastError());
}

void batch_materialize(const BMat *cols, const uint32_t *page_start, uint32_t ncols, uint32_t units, uint32_t grid,
                       hipStream_t s) {
    if (!units || !ncols) return;
    k_batch_materialize<<<pgrid(units, grid), kWG, 0, s>>>(cols, page_start, ncols, units);
    CJ_HIP(hipGetLastError());
}

__global__ void k_batch_totals(const uint64_t *poff, const uint32_t *ppage_start, uint32_t n, uint64_t *totals) {
    const uint32_t k = blockIdx.x * kWG + threadIdx.x;
    if (k < n) totals[k] = poff[ppage_start[k + 1]] - poff[ppage_start[k]];
}

void batch_totals(const uint64_t *poff, const uint32_t *ppage_start, uint32_t n, uint64_t *totals, hipStream_t s) {
    if (!n) return;
    k_batch_totals<<<(n + kWG - 1) / kWG, kWG, 0, s>>>(poff, ppage_start, n, totals);
    CJ_HIP(hipGetLastError());
}

__global__ void k_copy_u64(const uint64_t *src, uint64_t *dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)kWG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kWG) dst[i] = src[i];
}

void batch_copy_u64(const uint64_t *src, uint64_t *dst, uint64_t n, hipStream_t s) {
    if (!n) return;
    const uint64_t g = (n + kWG - 1) / kWG;
    k_copy_u64<<<(unsigned)(g < 4096 ? g : 4096), kWG, 0, s>>>(src, dst, n);
    CJ_HIP(hipGetLastError());
}

}  // namespace cj
