// join.hip -- radix partitioning, unchained hash-table build, probe and the
// tiny-build nested loop.
//
// Algorithm follows the reference's UnchainedHashtable
// (join_execution/hashtable.h:63-505): a power-of-two directory of u64
// entries (end_offset << 16 | 16-bit Bloom tag), keys and row ids stored
// contiguously in slot order, built partition by partition after a radix
// partition pass; probe = Bloom test, then a scan of [prev_end, end)
// (hashtable.h:298-313, hash_join.h:156-244).  Tiny build sides (< 8 rows)
// use the nested loop of join_execution/nested_loop.h:107-244 with the
// build keys broadcast to every lane.
//
// gfx950 mapping: one workgroup per key page (or 2048-row chunk of a row-id
// list); partition histograms and slot counts live in LDS; the partition
// pass reserves output ranges with one global atomic per (workgroup,
// partition); a build workgroup owns one partition (<= 4096 slots) so its
// counts, prefix and Bloom tags stay in LDS and the directory is written
// coalesced.
#include "cj_internal.hpp"
#include "bloom.hpp"
#include "device_util.hpp"

namespace cj {



// Iterate the keys of one work unit in row order, 256 rows per tile.  For
// each tile, `fn(in_range, ok, key, rid)` is called by every thread (so it
// may use workgroup barriers).  `want_nulls` (hash build over an
// intermediate) passes NULL keys through as INT32_MIN, the reference's
// value_t encoding (hashtable.h:357 inserts them unfiltered).
template <class F>
__device__ __forceinline__ void for_each_key(const KeySource &k, uint32_t unit, bool want_nulls, uint32_t *lds4,
                                             F &&fn) {
    if (k.mode == KEY_COLUMNAR) {
        const uint8_t *p = k.pages + (size_t)unit * kPage;
        const uint32_t n = ld16(p), nv = ld16(p + 2);
        const uint64_t base = k.row_start[unit];
        const uint8_t *bm = p + kPage - ((n + 7) >> 3);
        const bool dense = n == nv;
        uint32_t running = 0;
        for (uint32_t t0 = 0; t0 < n; t0 += kWG) {
            uint32_t r = t0 + threadIdx.x;
            bool inr = r < n;
            bool ok = inr && (dense || page_bit(bm, r));
            uint64_t m;
            uint32_t tot;
            uint32_t vi = dense ? r : running + tile_rank(ok, &m, &tot, lds4);
            ok = ok && vi < nv;  // malformed bitmap: stay inside the value area
            int32_t key = ok ? ld32(p + 4 + 4 * (size_t)vi) : 0;
            fn(inr, ok, key, (uint32_t)(base + r));
            if (!dense) running += tot;
        }
    } else {
        const uint64_t a = (uint64_t)unit * kInterChunk;
        const uint64_t e = a + kInterChunk < k.rows ? a + kInterChunk : k.rows;
        for (uint64_t t0 = a; t0 < e; t0 += kWG) {
            uint64_t j = t0 + threadIdx.x;
            bool inr = j < e;
            int32_t key = kI32Null;
            if (inr) {
                uint32_t rid = k.rids ? k.rids[j] : (uint32_t)j;
                if (!k.values) key = dense_page_value(k.dense_pages, k.dense_rows, rid);
                else if (valid_bit(k.valid_bits, rid)) key = k.values[rid];
            }
            bool ok = inr && (want_nulls || key != kI32Null);
            fn(inr, ok, key, (uint32_t)j);
        }
    }
}

__device__ __forceinline__ uint64_t slot_of(uint64_t h, int log_slots) { return h >> (64 - log_slots); }

// ---------------------------------------------------- radix partition pass
// Partition of a key = the top bits of its directory slot, so each build
// partition owns a contiguous slot range (hashtable.h:439-478 partitions by
// hash bits too, into per-thread chunk lists).
__device__ __forceinline__ uint32_t part_of(int32_t key, const HashGeom &g) {
    return (uint32_t)(slot_of(hash_key(key), g.log_slots) >> g.log_part);
}

// Wave-aggregated add on an LDS counter array: the active lanes that carry
// the same bin find each other with one ballot per bin bit, their lowest
// lane adds the group's size once, and every lane gets the old value plus
// its rank in the group (a unique slot).  A skewed key (thousands of
// duplicates) costs one LDS atomic per wave instead of 64.
__device__ __forceinline__ uint32_t agg_add(uint32_t *ctr, uint32_t bin, bool active, int bin_bits) {
    const uint64_t lt = lanemask_lt();
    uint64_t peers = __ballot(active);
    for (int b = 0; b < bin_bits; ++b) {
        const bool bit = (bin >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
    }
    const uint32_t leader = peers ? (uint32_t)__ffsll((unsigned long long)peers) - 1u : 0u;
    uint32_t base = 0;
    if (active && lane_id() == leader) base = atomicAdd(&ctr[bin], (uint32_t)__popcll(peers));
    base = (uint32_t)__shfl((int)base, (int)leader, 64);
    return base + (uint32_t)__popcll(peers & lt);
}

// Histogram of one unit's keys per partition (LDS, wave-aggregated), then
// one global add per non-empty partition.
__global__ __launch_bounds__(kWG) void k_part_count(KeySource k, HashGeom g, uint64_t *part_count) {
    extern __shared__ uint32_t hist[];  // num_partitions
    __shared__ uint32_t lds4[4];
    const int bits = g.log_slots - g.log_part;
    for (uint32_t i = threadIdx.x; i < g.num_partitions; i += kWG) hist[i] = 0;
    __syncthreads();
    const bool want_nulls = k.mode == KEY_INTER;
    for_each_key(k, blockIdx.x, want_nulls, lds4, [&](bool, bool ok, int32_t key, uint32_t) {
        agg_add(hist, ok ? part_of(key, g) : 0u, ok, bits);
    });
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < g.num_partitions; i += kWG)
        if (hist[i]) atomicAdd((unsigned long long *)&part_count[i], (unsigned long long)hist[i]);
}

// LDS-staged scatter of one unit (at most kUnitTuples tuples): histogram,
// block scan into per-partition starts in the staging buffer, one global
// reservation per non-empty partition, the tuples staged grouped by
// partition, then written out in staged order so consecutive lanes store
// consecutive addresses of one partition's run.
constexpr uint32_t kUnitTuples = 2048;  // a key page holds <= 2047 INT32 values; a row-id unit 2048 rows

__global__ __launch_bounds__(kWG) void k_part_scatter(KeySource k, HashGeom g, uint64_t *part_cursor,
                                                      int32_t *tup_keys, uint32_t *tup_rids) {
    extern __shared__ uint32_t sm[];  // start[P] (then the running cursor), delta[P]
    const uint32_t P = g.num_partitions;
    uint32_t *cur = sm, *delta = sm + P;
    __shared__ int32_t skey[kUnitTuples];
    __shared__ uint32_t srid[kUnitTuples];
    __shared__ uint32_t lds4[4];
    __shared__ uint32_t wsum[kWG / 64];
    const int bits = g.log_slots - g.log_part;
    for (uint32_t i = threadIdx.x; i < P; i += kWG) cur[i] = 0;
    __syncthreads();
    const bool want_nulls = k.mode == KEY_INTER;
    for_each_key(k, blockIdx.x, want_nulls, lds4, [&](bool, bool ok, int32_t key, uint32_t) {
        agg_add(cur, ok ? part_of(key, g) : 0u, ok, bits);
    });
    __syncthreads();
    // exclusive scan of the counts (each thread owns a run of bins); each
    // non-empty bin reserves its range of the partition: delta = global
    // position - staged position
    const uint32_t per = (P + kWG - 1) / kWG, b0 = threadIdx.x * per;
    uint32_t local = 0;
    for (uint32_t i = 0; i < per; ++i)
        if (b0 + i < P) local += cur[b0 + i];
    const uint32_t inc = wave_incl_scan<uint32_t>(local);
    if (lane_id() == 63) wsum[wave_id()] = inc;
    __syncthreads();
    uint32_t pre = inc - local;
    for (uint32_t w = 0; w < wave_id(); ++w) pre += wsum[w];
    const uint32_t staged = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    for (uint32_t i = 0; i < per; ++i)
        if (b0 + i < P) {
            const uint32_t c = cur[b0 + i];
            const uint64_t gpos = c ? atomicAdd((unsigned long long *)&part_cursor[b0 + i], (unsigned long long)c) : 0;
            cur[b0 + i] = pre;
            delta[b0 + i] = (uint32_t)gpos - pre;  // modulo 2^32: tuple positions are below 2^32
            pre += c;
        }
    __syncthreads();
    for_each_key(k, blockIdx.x, want_nulls, lds4, [&](bool, bool ok, int32_t key, uint32_t rid) {
        const uint32_t at = agg_add(cur, ok ? part_of(key, g) : 0u, ok, bits);
        if (ok) {
            skey[at] = key;
            srid[at] = rid;
        }
    });
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < staged; i += kWG) {
        const int32_t key = skey[i];
        const uint32_t pos = delta[part_of(key, g)] + i;
        tup_keys[pos] = key;
        tup_rids[pos] = srid[i];
    }
}

void build_partition_hist(const KeySource &k, const HashGeom &g, uint64_t *part_count, hipStream_t s) {
    uint32_t units = key_source_units(k);
    if (!units) return;
    size_t sh = g.num_partitions * sizeof(uint32_t);
    k_part_count<<<units, kWG, sh, s>>>(k, g, part_count);
    CJ_HIP(hipGetLastError());
}

void build_partition_scatter(const KeySource &k, const HashGeom &g, uint64_t *part_cursor, int32_t *tup_keys,
                             uint32_t *tup_rids, hipStream_t s) {
    uint32_t units = key_source_units(k);
    if (!units) return;
    // part_cursor starts as a copy of the partition offsets and is advanced
    // by one atomic per (workgroup, non-empty partition).
    size_t sh = 2 * (size_t)g.num_partitions * sizeof(uint32_t);
    k_part_scatter<<<units, kWG, sh, s>>>(k, g, part_cursor, tup_keys, tup_rids);
    CJ_HIP(hipGetLastError());
}

// -------------------------------------------- per-partition build in LDS
// build_partition (hashtable.h:197-246): count per slot, exclusive prefix,
// scatter keys/row ids, OR Bloom tags, write (end << 16) | tag.  1024
// threads per partition: a hot key of a skewed build side fills one
// partition, whose workgroup sets the kernel's length.  A partition holds
// 2^log_part slots; the workgroup counts 2^log_sub of them at a time in LDS
// (a count and a tag word per slot: up to 16384 slots, 128 KiB), reading
// the partition's tuples twice per sub-range.
constexpr uint32_t kBuildWG = 1024;

__global__ __launch_bounds__(kBuildWG) void k_build_partitions(HashGeom g, const uint64_t *part_start,
                                                               const uint64_t *part_end, const int32_t *tup_keys,
                                                               const uint32_t *tup_rids, uint64_t *dir,
                                                               int32_t *keys, uint32_t *rids) {
    extern __shared__ uint32_t sm[];
    const uint32_t S = 1u << g.log_sub;
    uint32_t *cnt = sm, *tag = sm + S;
    __shared__ uint32_t wsum[kBuildWG / 64];
    const uint32_t p = blockIdx.x;
    const uint64_t a = part_start[p], e = part_end[p];
    const uint32_t pmask = (1u << g.log_part) - 1u;
    const uint32_t nsub = 1u << (g.log_part - g.log_sub);
    uint64_t base = a;  // first tuple position of the sub-range
    for (uint32_t sub = 0; sub < nsub; ++sub) {
        for (uint32_t i = threadIdx.x; i < S; i += kBuildWG) {
            cnt[i] = 0;
            tag[i] = 0;
        }
        __syncthreads();
        for (uint64_t i = a + threadIdx.x; i < e; i += kBuildWG) {
            const uint64_t h = hash_key(tup_keys[i]);
            const uint32_t ps = (uint32_t)slot_of(h, g.log_slots) & pmask;
            if ((ps >> g.log_sub) != sub) continue;
            const uint32_t ls = ps & (S - 1u);
            atomicAdd(&cnt[ls], 1u);
            atomicOr(&tag[ls], (uint32_t)c_bloom.t[h & 0x7FF]);
        }
        __syncthreads();
        // exclusive prefix over the S slot counts: each thread owns S/1024 slots
        const uint32_t per = (S + kBuildWG - 1) / kBuildWG;
        const uint32_t s0 = threadIdx.x * per;
        uint32_t local = 0;
        for (uint32_t i = 0; i < per; ++i)
            if (s0 + i < S) local += cnt[s0 + i];
        const uint32_t inc = wave_incl_scan<uint32_t>(local);
        if (lane_id() == 63) wsum[wave_id()] = inc;
        __syncthreads();
        uint32_t pre = inc - local, subtotal = 0;
        for (uint32_t w = 0; w < kBuildWG / 64; ++w) {
            pre += w < wave_id() ? wsum[w] : 0u;
            subtotal += wsum[w];
        }
        for (uint32_t i = 0; i < per; ++i)
            if (s0 + i < S) {
                const uint32_t c = cnt[s0 + i];
                cnt[s0 + i] = pre;
                pre += c;
            }
        __syncthreads();
        for (uint64_t i = a + threadIdx.x; i < e; i += kBuildWG) {
            const int32_t key = tup_keys[i];
            const uint32_t ps = (uint32_t)slot_of(hash_key(key), g.log_slots) & pmask;
            if ((ps >> g.log_sub) != sub) continue;
            const uint64_t pos = base + atomicAdd(&cnt[ps & (S - 1u)], 1u);
            keys[pos] = key;
            rids[pos] = tup_rids[i];
        }
        __syncthreads();
        uint64_t *d = dir + ((uint64_t)p << g.log_part) + ((uint64_t)sub << g.log_sub);
        for (uint32_t i = threadIdx.x; i < S; i += kBuildWG)
            d[i] = ((base + cnt[i]) << 16) | (uint64_t)(tag[i] & 0xFFFF);
        base += subtotal;
        __syncthreads();
    }
}

void build_partitions(const HashGeom &g, const uint64_t *part_off, const int32_t *tup_keys, const uint32_t *tup_rids,
                      uint64_t *dir, int32_t *keys, uint32_t *rids, hipStream_t s) {
    const size_t sh = 2 * (size_t)(1u << g.log_sub) * sizeof(uint32_t);
    static bool lds_set = false;  // room for 128 KiB of dynamic LDS (set once per process)
    if (!lds_set) {
        CJ_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_build_partitions),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 2 * 16384 * 4));
        lds_set = true;
    }
    k_build_partitions<<<g.num_partitions, kBuildWG, sh, s>>>(g, part_off, part_off + 1, tup_keys, tup_rids, dir,
                                                              keys, rids);
    CJ_HIP(hipGetLastError());
}

// ------------------------------------------------------------ nested loop
// visit_rows (nested_loop.h:50-95): the first (up to) 8 non-NULL build keys
// in row order.  The build side has fewer than 8 rows.
__global__ void k_nl_collect(KeySource k, NLKeys *out) {
    if (threadIdx.x != 0) return;
    NLKeys r{};
    r.count = 0;
    uint32_t units = k.mode == KEY_COLUMNAR ? k.num_pages : (uint32_t)((k.rows + kInterChunk - 1) / kInterChunk);
    for (uint32_t u = 0; u < units && r.count < 8; ++u) {
        if (k.mode == KEY_COLUMNAR) {
            const uint8_t *p = k.pages + (size_t)u * kPage;
            uint32_t n = ld16(p), nv = ld16(p + 2), vi = 0;
            const uint8_t *bm = p + kPage - ((n + 7) >> 3);
            for (uint32_t rr = 0; rr < n && r.count < 8; ++rr) {
                bool ok = n == nv || page_bit(bm, rr);
                if (!ok || vi >= nv) continue;
                r.keys[r.count] = ld32(p + 4 + 4 * (size_t)(n == nv ? rr : vi));
                r.rids[r.count] = (uint32_t)(k.row_start[u] + rr);
                r.count++;
                vi++;
            }
        } else {
            uint64_t a = (uint64_t)u * kInterChunk, e = a + kInterChunk < k.rows ? a + kInterChunk : k.rows;
            for (uint64_t j = a; j < e && r.count < 8; ++j) {
                uint32_t rid = k.rids ? k.rids[j] : (uint32_t)j;
                int32_t key = !k.values ? dense_page_value(k.dense_pages, k.dense_rows, rid)
                                        : (valid_bit(k.valid_bits, rid) ? k.values[rid] : kI32Null);
                if (key == kI32Null) continue;
                r.keys[r.count] = key;
                r.rids[r.count] = (uint32_t)j;
                r.count++;
            }
        }
    }
    *out = r;
}
void nl_collect(const KeySource &k, NLKeys *out, hipStream_t s) {
    k_nl_collect<<<1, 64, 0, s>>>(k, out);
    CJ_HIP(hipGetLastError());
}

}  // namespace cj
