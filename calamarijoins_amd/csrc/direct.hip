// direct.hip -- direct-mapped build side for unique, dense INT32 keys.
//
// When every non-NULL build key is distinct and the keys span a range of at
// most ~2x the build rows (primary keys are the common case), the unchained
// hash table (hashtable.h:63-505) is replaced by a table indexed by
// key - min: a presence bitmap (one bit per key of the range, small enough
// to stay in L2) answers the counting pass, and a record array holds, per
// key, the build side's projected payload values with their NULL flags, so
// the probe fetches one 16-byte record per match instead of a directory
// entry, a key, a row id and one value per build column.  The join result
// is identical: a probe key matches exactly the build row carrying that key.
//
// Building takes no atomics on the table: records are zeroed, each build
// row stores its record (one 16-byte store per chunk) with a present bit,
// and a streaming pass turns the present bits into the bitmap, counting
// them.  Two rows with the same key store to the same slot, so fewer present
// slots than keyed rows means duplicates, and the executor falls back to the
// hash table.
#include "cj_internal.hpp"
#include "device_util.hpp"
#include "direct.hpp"

namespace cj {

// Each thread reduces 16 consecutive keys (four 16-byte loads, half a
// validity word) of a decoded column, or 16 rows of a column read in place
// from its pages (rows kWG apart, so a wave's loads stay contiguous); a workgroup reduces through shuffles and LDS and writes
// its partial to its own slot; one workgroup combines the partials.  (Many
// workgroups updating one address with device atomics serialize.)
__global__ __launch_bounds__(kWG) void k_key_minmax(const int32_t *keys, const uint32_t *valid,
                                                    const uint8_t *key_pages, uint32_t key_rows, uint64_t n,
                                                    int32_t *part) {
    __shared__ int32_t red[2][kWG / 64];
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    const uint64_t groups = (n + 15) / 16;
    // in place (every row keyed): a workgroup takes 16 * kWG rows, lane-contiguous
    for (uint64_t c = blockIdx.x; !keys && c * 16 * kWG < n; c += gridDim.x) {
        int32_t k[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint64_t r = c * 16 * kWG + (uint64_t)i * kWG + threadIdx.x;
            k[i] = dense_page_value(key_pages, key_rows, (uint32_t)(r < n ? r : n - 1));
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            lo = min(lo, k[i]);
            hi = max(hi, k[i]);
        }
    }
    for (uint64_t g = blockIdx.x * (uint64_t)kWG + threadIdx.x; keys && g < groups; g += (uint64_t)gridDim.x * kWG) {
        const uint64_t r0 = g * 16;
        const uint32_t vw = valid[r0 >> 5] >> (r0 & 16);
        if (r0 + 16 <= n) {
            const int4 *k4 = reinterpret_cast<const int4 *>(keys + r0);
            int4 q[4] = {k4[0], k4[1], k4[2], k4[3]};
            const int32_t *k = reinterpret_cast<const int32_t *>(q);
#pragma unroll
            for (int i = 0; i < 16; ++i)
                if ((vw >> i) & 1u) {
                    lo = min(lo, k[i]);
                    hi = max(hi, k[i]);
                }
        } else {
            for (uint64_t r = r0; r < n; ++r)
                if ((vw >> (r - r0)) & 1u) {
                    lo = min(lo, keys[r]);
                    hi = max(hi, keys[r]);
                }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o, 64));
        hi = max(hi, __shfl_xor(hi, o, 64));
    }
    if (lane_id() == 0) {
        red[0][wave_id()] = lo;
        red[1][wave_id()] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < kWG / 64; ++w) {
            lo = min(lo, red[0][w]);
            hi = max(hi, red[1][w]);
        }
        part[2 * blockIdx.x] = lo;
        part[2 * blockIdx.x + 1] = hi;
    }
}

__global__ __launch_bounds__(kWG) void k_minmax_final(const int32_t *part, uint32_t nparts, int32_t *mm) {
    __shared__ int32_t red[2][kWG / 64];
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (uint32_t i = threadIdx.x; i < nparts; i += kWG) {
        lo = min(lo, part[2 * i]);
        hi = max(hi, part[2 * i + 1]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o, 64));
        hi = max(hi, __shfl_xor(hi, o, 64));
    }
    if (lane_id() == 0) {
        red[0][wave_id()] = lo;
        red[1][wave_id()] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < kWG / 64; ++w) {
            lo = min(lo, red[0][w]);
            hi = max(hi, red[1][w]);
        }
        mm[0] = lo;
        mm[1] = hi;
    }
}

void key_minmax(const int32_t *keys, const uint32_t *valid, const uint8_t *key_pages, uint32_t key_rows, uint64_t n,
                int32_t *mm, int32_t *part,
                hipStream_t s) {
    const uint64_t groups = (n + 15) / 16;
    uint64_t g = (groups + kWG - 1) / kWG;
    g = g < kMinmaxParts ? g : kMinmaxParts;
    if (g) k_key_minmax<<<(unsigned)g, kWG, 0, s>>>(keys, valid, key_pages, key_rows, n, part);
    k_minmax_final<<<1, kWG, 0, s>>>(part, (uint32_t)g, mm);
    CJ_HIP(hipGetLastError());
}

__device__ __forceinline__ int32_t direct_lo(const DirectBuildArgs &a) { return a.known ? a.lo : a.minmax[0]; }

// Key range of the build side, or 0 when it has no key or exceeds max_range.
__device__ __forceinline__ uint64_t direct_range(const DirectBuildArgs &a) {
    const int32_t lo = direct_lo(a), hi = a.known ? a.hi : a.minmax[1];
    if (lo > hi) return 0;
    const uint64_t r = (uint64_t)((int64_t)hi - (int64_t)lo) + 1;
    return r <= a.max_range ? r : 0;
}

__global__ __launch_bounds__(kWG) void k_direct_zero(DirectBuildArgs a) {
    const uint64_t range = direct_range(a);
    const uint64_t chunks = range * (a.stride / 16);
    uint4 *r = reinterpret_cast<uint4 *>(a.records);
    for (uint64_t i = blockIdx.x * (uint64_t)kWG + threadIdx.x; i < chunks; i += (uint64_t)gridDim.x * kWG)
        r[i] = make_uint4(0, 0, 0, 0);
}

// Field f's value, placed at byte offset `off` of a record built in the
// registers of 16-byte chunks q[] (offsets are uniform).
template <int NQ>
__device__ __forceinline__ void rec_put(uint4 (&q)[NQ], uint32_t off, uint32_t v) {
    const uint32_t d = off >> 2;
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
        if (d == 4u * c + 0) q[c].x = v;
        if (d == 4u * c + 1) q[c].y = v;
        if (d == 4u * c + 2) q[c].z = v;
        if (d == 4u * c + 3) q[c].w = v;
    }
}

template <int NQ>
__global__ __launch_bounds__(kWG) void k_direct_scatter(DirectBuildArgs a) {
    __shared__ uint32_t red[kWG / 64];
    const uint64_t range = direct_range(a);
    uint32_t cnt = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)kWG + threadIdx.x; range && i < a.rows; i += (uint64_t)gridDim.x * kWG) {
        if (a.keys && !valid_bit(a.key_valid, i)) continue;
        ++cnt;
        const int32_t key = a.keys ? a.keys[i] : dense_page_value(a.key_pages, a.key_rows, (uint32_t)i);
        const uint64_t slot = (uint64_t)((int64_t)key - (int64_t)direct_lo(a));
        uint4 q[NQ];
#pragma unroll
        for (int c = 0; c < NQ; ++c) q[c] = make_uint4(0, 0, 0, 0);
        uint32_t flags = 1u << kDirectPresentBit;
        for (uint32_t f = 0; f < a.nfields; ++f) {
            const DirectField &fd = a.fields[f];
            if (!fd.values) {  // in place: no NULLs
                flags |= 1u << f;
                rec_put<NQ>(q, fd.offset, (uint32_t)dense_page_value(fd.pages, fd.dense_rows, (uint32_t)i));
                continue;
            }
            const bool ok = valid_bit(fd.valid_bits, i);
            flags |= (uint32_t)ok << f;
            if (fd.width == 4) {
                rec_put<NQ>(q, fd.offset, ok ? static_cast<const uint32_t *>(fd.values)[i] : 0u);
            } else {
                const uint64_t v = ok ? static_cast<const uint64_t *>(fd.values)[i] : 0ull;
                rec_put<NQ>(q, fd.offset, (uint32_t)v);
                rec_put<NQ>(q, fd.offset + 4, (uint32_t)(v >> 32));
            }
        }
        if (a.rid_offset >= 0) rec_put<NQ>(q, (uint32_t)a.rid_offset, (uint32_t)i);
        rec_put<NQ>(q, a.flags_offset, flags);
        uint4 *dst = reinterpret_cast<uint4 *>(a.records + slot * a.stride);
#pragma unroll
        for (int c = 0; c < NQ; ++c) dst[c] = q[c];
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if (lane_id() == 0) red[wave_id()] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kWG / 64; ++w) t += red[w];
        a.partials[blockIdx.x] = t;
    }
}

constexpr uint32_t kScatterRows = 4;  // rows per thread whose loads are issued together

// The validity word of a field or key read in place (no NULLs).
__device__ const uint32_t kAllValid = ~0u;

// k_direct_scatter for NF (<= 3) fields: a thread takes kScatterRows rows
// (kWG apart) per step and issues every load of them (key, validity words,
// value words) before it builds and stores their records.  The field loop
// is unrolled and the loads are unconditional (addresses are selected, not
// branched on: a field read in place or decoded, an 8-byte value's high
// word at +4 of the low one, a 4-byte value reading its own word twice), so
// no load waits for another.  KIN: the keys are read in place.
template <int NF, int NQ, bool KIN>
__global__ __launch_bounds__(kWG) void k_direct_scatter_nf(DirectBuildArgs a) {
    __shared__ uint32_t red[kWG / 64];
    const uint64_t range = direct_range(a);
    const int64_t lo = direct_lo(a);
    const uint64_t n = a.rows;
    uint32_t cnt = 0;
    const uint64_t step = (uint64_t)gridDim.x * kWG * kScatterRows;
    for (uint64_t i0 = blockIdx.x * (uint64_t)kWG * kScatterRows + threadIdx.x; range && i0 < n; i0 += step) {
        uint64_t ic[kScatterRows];
        int32_t key[kScatterRows];
        uint32_t kw[kScatterRows];
#pragma unroll
        for (uint32_t r = 0; r < kScatterRows; ++r) {
            const uint64_t i = i0 + r * kWG;
            ic[r] = i < n ? i : n - 1;
            if constexpr (KIN) {
                key[r] = dense_page_value(a.key_pages, a.key_rows, (uint32_t)ic[r]);
                kw[r] = ~0u;
            } else {
                key[r] = a.keys[ic[r]];
                kw[r] = a.key_valid[ic[r] >> 5];
            }
        }
        uint32_t vw[NF][kScatterRows], lo32[NF][kScatterRows], hi32[NF][kScatterRows];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const DirectField &fd = a.fields[f];
            const uint32_t hoff = fd.width == 8 ? 4u : 0u;
            const uint32_t dr = fd.dense_rows ? fd.dense_rows : 1u;  // (in place only)
#pragma unroll
            for (uint32_t r = 0; r < kScatterRows; ++r) {
                const uint32_t pg = (uint32_t)ic[r] / dr;
                const uint8_t *inplace = fd.pages + (size_t)pg * kPage + 4 + 4 * (size_t)((uint32_t)ic[r] - pg * dr);
                const uint8_t *dense = static_cast<const uint8_t *>(fd.values) + ic[r] * (uint64_t)fd.width;
                const uint8_t *p = fd.values ? dense : inplace;
                const uint32_t *vp = fd.valid_bits ? fd.valid_bits + (ic[r] >> 5) : &kAllValid;
                lo32[f][r] = *reinterpret_cast<const uint32_t *>(p);
                hi32[f][r] = *reinterpret_cast<const uint32_t *>(p + hoff);
                vw[f][r] = *vp;
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < kScatterRows; ++r) {
            const uint64_t i = i0 + r * kWG;
            if (i >= n || !((kw[r] >> (ic[r] & 31)) & 1u)) continue;
            ++cnt;
            const uint64_t slot = (uint64_t)((int64_t)key[r] - lo);
            uint4 q[NQ];
#pragma unroll
            for (int c = 0; c < NQ; ++c) q[c] = make_uint4(0, 0, 0, 0);
            uint32_t flags = 1u << kDirectPresentBit;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const DirectField &fd = a.fields[f];
                const bool ok = (vw[f][r] >> (ic[r] & 31)) & 1u;
                flags |= (uint32_t)ok << f;
                rec_put<NQ>(q, fd.offset, ok ? lo32[f][r] : 0u);
                if (fd.width == 8) rec_put<NQ>(q, fd.offset + 4, ok ? hi32[f][r] : 0u);
            }
            if (a.rid_offset >= 0) rec_put<NQ>(q, (uint32_t)a.rid_offset, (uint32_t)i);
            rec_put<NQ>(q, a.flags_offset, flags);
            uint4 *dst = reinterpret_cast<uint4 *>(a.records + slot * a.stride);
#pragma unroll
            for (int c = 0; c < NQ; ++c) dst[c] = q[c];
        }
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if (lane_id() == 0) red[wave_id()] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kWG / 64; ++w) t += red[w];
        a.partials[blockIdx.x] = t;
    }
}

template <int NQ, bool KIN>
static void launch_scatter_nf(const DirectBuildArgs &a, unsigned g, hipStream_t s) {
    if (a.nfields == 1) k_direct_scatter_nf<1, NQ, KIN><<<g, kWG, 0, s>>>(a);
    else if (a.nfields == 2) k_direct_scatter_nf<2, NQ, KIN><<<g, kWG, 0, s>>>(a);
    else k_direct_scatter_nf<3, NQ, KIN><<<g, kWG, 0, s>>>(a);
}

// 64 slots per wave step: ballot of their present bits -> two presence words.
__global__ __launch_bounds__(kWG) void k_direct_presence(DirectBuildArgs a) {
    __shared__ uint32_t red[kWG / 64];
    const uint64_t range = direct_range(a);
    const uint32_t lane = lane_id();
    const uint64_t steps = (range + 63) / 64;
    const uint64_t wave = blockIdx.x * (uint64_t)(kWG / 64) + wave_id();
    const uint64_t waves = (uint64_t)gridDim.x * (kWG / 64);
    uint32_t cnt = 0;
    for (uint64_t st = wave; st < steps; st += waves) {
        const uint64_t slot = st * 64 + lane;
        const uint64_t sc = slot < range ? slot : range - 1;
        const uint32_t fl = *reinterpret_cast<const uint32_t *>(a.records + sc * a.stride + a.flags_offset);
        const uint64_t b = __ballot(slot < range && ((fl >> kDirectPresentBit) & 1u));
        if (lane < 2) a.presence[2 * st + lane] = (uint32_t)(b >> (32 * lane));
        cnt += (uint32_t)__popcll(b);
    }
    if (lane == 0) red[wave_id()] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kWG / 64; ++w) t += red[w];
        a.partials[kDirectParts + blockIdx.x] = t;
    }
}

// counters[0] = keyed rows, counters[1] = present slots.
__global__ __launch_bounds__(kWG) void k_direct_finish(DirectBuildArgs a, uint32_t nscatter, uint32_t npresence) {
    __shared__ uint32_t red[2][kWG / 64];
    uint32_t k = 0, p = 0;
    for (uint32_t i = threadIdx.x; i < nscatter; i += kWG) k += a.partials[i];
    for (uint32_t i = threadIdx.x; i < npresence; i += kWG) p += a.partials[kDirectParts + i];
    for (int o = 32; o > 0; o >>= 1) {
        k += __shfl_xor(k, o, 64);
        p += __shfl_xor(p, o, 64);
    }
    if (lane_id() == 0) {
        red[0][wave_id()] = k;
        red[1][wave_id()] = p;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        k = p = 0;
        for (uint32_t w = 0; w < kWG / 64; ++w) {
            k += red[0][w];
            p += red[1][w];
        }
        a.counters[0] = k;
        a.counters[1] = p;
    }
}

// Row-id maps (no payload fields): the record of a slot is the build row id
// alone (4 bytes; ~0u marks an empty slot, a row id never reaches it), so
// clearing, scattering and the presence pass move a quarter of the bytes of
// the smallest field record.
__global__ __launch_bounds__(kWG) void k_rid_zero(DirectBuildArgs a) {
    const uint64_t chunks = (direct_range(a) + 3) / 4;
    uint4 *r = reinterpret_cast<uint4 *>(a.records);
    for (uint64_t i = blockIdx.x * (uint64_t)kWG + threadIdx.x; i < chunks; i += (uint64_t)gridDim.x * kWG)
        r[i] = make_uint4(~0u, ~0u, ~0u, ~0u);
}

__global__ __launch_bounds__(kWG) void k_rid_scatter(DirectBuildArgs a) {
    __shared__ uint32_t red[kWG / 64];
    const uint64_t range = direct_range(a);
    uint32_t *rids = reinterpret_cast<uint32_t *>(a.records);
    uint32_t cnt = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)kWG + threadIdx.x; range && i < a.rows; i += (uint64_t)gridDim.x * kWG) {
        if (a.keys && !valid_bit(a.key_valid, i)) continue;
        ++cnt;
        const int32_t key = a.keys ? a.keys[i] : dense_page_value(a.key_pages, a.key_rows, (uint32_t)i);
        rids[(uint64_t)((int64_t)key - (int64_t)direct_lo(a))] = (uint32_t)i;
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if (lane_id() == 0) red[wave_id()] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kWG / 64; ++w) t += red[w];
        a.partials[blockIdx.x] = t;
    }
}

// 64 slots per wave step: ballot of the filled slots -> two presence words.
__global__ __launch_bounds__(kWG) void k_rid_present(DirectBuildArgs a) {
    __shared__ uint32_t red[kWG / 64];
    const uint64_t range = direct_range(a);
    const uint32_t *rids = reinterpret_cast<const uint32_t *>(a.records);
    const uint32_t lane = lane_id();
    const uint64_t steps = (range + 63) / 64;
    const uint64_t wave = blockIdx.x * (uint64_t)(kWG / 64) + wave_id();
    const uint64_t waves = (uint64_t)gridDim.x * (kWG / 64);
    uint32_t cnt = 0;
    for (uint64_t st = wave; st < steps; st += waves) {
        const uint64_t slot = st * 64 + lane;
        const uint32_t r = rids[slot < range ? slot : range - 1];
        const uint64_t b = __ballot(slot < range && r != ~0u);
        if (lane < 2) a.presence[2 * st + lane] = (uint32_t)(b >> (32 * lane));
        cnt += (uint32_t)__popcll(b);
    }
    if (lane == 0) red[wave_id()] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kWG / 64; ++w) t += red[w];
        a.partials[kDirectParts + blockIdx.x] = t;
    }
}

static void rid_build(const DirectBuildArgs &a, hipStream_t s) {
    if (a.stride != 4 || a.rid_offset != 0) throw InvalidError("bad row-id map layout");
    const uint64_t zg = ((a.max_range + 3) / 4 + kWG - 1) / kWG;
    if (!(a.known && a.full)) k_rid_zero<<<(unsigned)(zg < 8192 ? zg : 8192), kWG, 0, s>>>(a);
    uint64_t sg = (a.rows + kWG - 1) / kWG;
    sg = sg < kDirectParts ? sg : kDirectParts;
    k_rid_scatter<<<(unsigned)sg, kWG, 0, s>>>(a);
    if (a.known) {
        CJ_HIP(hipGetLastError());
        if (!a.full) {
            const uint64_t steps = (a.max_range + 63) / 64;
            uint64_t pg = (steps + kWG / 64 - 1) / (kWG / 64);
            k_rid_present<<<(unsigned)(pg < kDirectParts ? pg : kDirectParts), kWG, 0, s>>>(a);
            CJ_HIP(hipGetLastError());
        }
        return;
    }
    const uint64_t steps = (a.max_range + 63) / 64;
    uint64_t pg = (steps + kWG / 64 - 1) / (kWG / 64);
    pg = pg < kDirectParts ? pg : kDirectParts;
    k_rid_present<<<(unsigned)pg, kWG, 0, s>>>(a);
    k_direct_finish<<<1, kWG, 0, s>>>(a, (uint32_t)sg, (uint32_t)pg);
    CJ_HIP(hipGetLastError());
}

void direct_build(const DirectBuildArgs &a, hipStream_t s) {
    if (!a.rows || !a.max_range) return;
    if (a.nfields == 0 && a.rid_offset >= 0) return rid_build(a, s);
    if (a.stride % 16 || a.stride == 0 || a.stride > 80) throw InvalidError("bad direct-map record stride");
    const uint64_t chunks = a.max_range * (a.stride / 16);
    const uint64_t zg = (chunks + kWG - 1) / kWG;
    if (!(a.known && a.full)) k_direct_zero<<<(unsigned)(zg < 8192 ? zg : 8192), kWG, 0, s>>>(a);
    uint64_t sg = (a.rows + kWG - 1) / kWG;
    sg = sg < kDirectParts ? sg : kDirectParts;
    const uint32_t nq = a.stride / 16;
    if (a.nfields >= 1 && a.nfields <= 3 && nq <= 2) {
        // batched loads; a thread takes kScatterRows rows a step
        uint64_t bg = (a.rows + (uint64_t)kWG * kScatterRows - 1) / ((uint64_t)kWG * kScatterRows);
        bg = bg < kDirectParts ? bg : kDirectParts;
        sg = bg;
        const unsigned g = (unsigned)bg;
        const bool kin = a.keys == nullptr;
        if (nq == 1) kin ? launch_scatter_nf<1, true>(a, g, s) : launch_scatter_nf<1, false>(a, g, s);
        else kin ? launch_scatter_nf<2, true>(a, g, s) : launch_scatter_nf<2, false>(a, g, s);
    } else {
        switch (nq) {
            case 1: k_direct_scatter<1><<<(unsigned)sg, kWG, 0, s>>>(a); break;
            case 2: k_direct_scatter<2><<<(unsigned)sg, kWG, 0, s>>>(a); break;
            case 3: k_direct_scatter<3><<<(unsigned)sg, kWG, 0, s>>>(a); break;
            case 4: k_direct_scatter<4><<<(unsigned)sg, kWG, 0, s>>>(a); break;
            default: k_direct_scatter<5><<<(unsigned)sg, kWG, 0, s>>>(a); break;
        }
    }
    const uint64_t steps = (a.max_range + 63) / 64;
    uint64_t pg = (steps + kWG / 64 - 1) / (kWG / 64);
    pg = pg < kDirectParts ? pg : kDirectParts;
    if (!(a.known && a.full)) k_direct_presence<<<(unsigned)pg, kWG, 0, s>>>(a);
    if (!a.known) k_direct_finish<<<1, kWG, 0, s>>>(a, (uint32_t)sg, (uint32_t)pg);
    CJ_HIP(hipGetLastError());
}

}  // namespace cj
