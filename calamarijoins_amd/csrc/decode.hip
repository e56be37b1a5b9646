// decode.hip -- page decode and device-wide scans.
//
// Page decode restates the page format the reference reads in
// data_access/columnar_reader.h:120-201 (PageIndex::build: cumulative rows,
// 0xffff counts one row, 0xfffe none) and :452-474 (bitmap + popcount to
// find a value's slot), turned into one wave (fixed width) or workgroup
// (VARCHAR) per page that writes a dense, row-indexed column plus a
// validity bitmap.
#include <type_traits>

#include "cj_internal.hpp"
#include "device_util.hpp"

namespace cj {

// ------------------------------------------------------------------ fills
__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u64(uint64_t *p, uint64_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
static unsigned grid_for(uint64_t n) {
    uint64_t g = (n + kWG - 1) / kWG;
    return (unsigned)(g < 8192 ? (g ? g : 1) : 8192);
}
void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s) {
    if (!n) return;
    k_fill_u32<<<grid_for(n), kWG, 0, s>>>(p, v, n);
    CJ_HIP(hipGetLastError());
}
void fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s) {
    if (!n) return;
    k_fill_u64<<<grid_for(n), kWG, 0, s>>>(p, v, n);
    CJ_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ scans
// Exclusive scan of u64: block sums, then each block's scan offset by the
// sum of the block sums before it (read by the block itself up to
// kScanDirect blocks, else from a recursive scan of the sums).  out has
// n + 1 entries; out[n] is the total.  In-place allowed.
constexpr int kScanItems = 8;
constexpr uint64_t kScanBlock = (uint64_t)kWG * kScanItems;

__global__ void k_scan_block_sums(const uint64_t *in, uint64_t n, uint64_t *sums) {
    __shared__ uint64_t lds[4];
    uint64_t base = blockIdx.x * kScanBlock + (uint64_t)threadIdx.x * kScanItems;
    uint64_t v[kScanItems], s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) v[i] = in[base + i < n ? base + i : n - 1];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) s += base + i < n ? v[i] : 0;
    uint64_t tot;
    block_excl_scan<uint64_t>(s, &tot, lds);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

constexpr uint64_t kScanDirect = 8 * kWG;

// RAW: block_off holds the unscanned block sums.
template <bool RAW>
__global__ void k_scan_block(const uint64_t *in, uint64_t *out, uint64_t n, const uint64_t *block_off) {
    __shared__ uint64_t lds[4];
    uint64_t boff = 0;
    if (RAW) {
        uint64_t p = 0;
        for (uint32_t i = threadIdx.x; i < blockIdx.x; i += kWG) p += block_off[i];
        block_excl_scan<uint64_t>(p, &boff, lds);
    } else if (block_off) {
        boff = block_off[blockIdx.x];
    }
    uint64_t base = blockIdx.x * kScanBlock + (uint64_t)threadIdx.x * kScanItems;
    uint64_t v[kScanItems];
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) v[i] = in[base + i < n ? base + i : n - 1];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = base + i < n ? v[i] : 0;
        s += v[i];
    }
    uint64_t tot;
    uint64_t pre = block_excl_scan<uint64_t>(s, &tot, lds);
    uint64_t off = boff + pre;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) out[base + i] = off;
        off += v[i];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kWG - 1) out[n] = off;
}

// Two launches for all arrays together: block sums of f(in), then each
// block adds the sums of the blocks before it (read directly: an array of
// at most kMultiScanMax elements has at most 512 blocks) and scans its own.
// Counts are below 2^32 (rows of one unit), so f divides in 32 bits.
constexpr uint32_t kMsItems = 8;
constexpr uint64_t kMsBlock = (uint64_t)kWG * kMsItems;  // 2048 elements a workgroup
struct MultiScanArgs {
    ScanJob job[kMultiScanJobs];
    uint32_t first_block[kMultiScanJobs + 1];  // blocks of the jobs before each job
    uint32_t njobs;
    uint64_t *sums;  // one per block
};

__device__ __forceinline__ uint32_t ms_job(const MultiScanArgs &a, uint32_t b) {
    uint32_t k = 0;
    while (k + 1 < a.njobs && a.first_block[k + 1] <= b) ++k;
    return k;
}

__device__ __forceinline__ uint64_t ms_load(const ScanJob &j, uint64_t i) {
    const uint64_t x = j.in[i < j.n ? i : j.n - 1];
    if (i >= j.n) return 0;
    return j.div ? ((uint32_t)x + (uint32_t)j.div - 1u) / (uint32_t)j.div : x;
}

__global__ __launch_bounds__(kWG) void k_multi_scan_sums(MultiScanArgs a) {
    __shared__ uint64_t lds[4];
    const uint32_t k = ms_job(a, blockIdx.x);
    const ScanJob &j = a.job[k];
    const uint64_t base = (uint64_t)(blockIdx.x - a.first_block[k]) * kMsBlock + (uint64_t)threadIdx.x * kMsItems;
    uint64_t s = 0;
#pragma unroll
    for (uint32_t i = 0; i < kMsItems; ++i) s += ms_load(j, base + i);
    uint64_t tot;
    block_excl_scan<uint64_t>(s, &tot, lds);
    if (threadIdx.x == 0) a.sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kWG) void k_multi_scan_apply(MultiScanArgs a) {
    __shared__ uint64_t lds[4];
    const uint32_t k = ms_job(a, blockIdx.x);
    const ScanJob &j = a.job[k];
    const uint32_t b = blockIdx.x - a.first_block[k];
    uint64_t p = 0;
    for (uint32_t i = threadIdx.x; i < b; i += kWG) p += a.sums[a.first_block[k] + i];
    uint64_t boff;
    block_excl_scan<uint64_t>(p, &boff, lds);
    const uint64_t base = (uint64_t)b * kMsBlock + (uint64_t)threadIdx.x * kMsItems;
    uint64_t v[kMsItems], s = 0;
#pragma unroll
    for (uint32_t i = 0; i < kMsItems; ++i) {
        v[i] = ms_load(j, base + i);
        s += v[i];
    }
    uint64_t tot;
    uint64_t off = boff + block_excl_scan<uint64_t>(s, &tot, lds);
#pragma unroll
    for (uint32_t i = 0; i < kMsItems; ++i) {
        if (base + i < j.n) j.out[base + i] = off;
        off += v[i];
    }
    const uint32_t nb = a.first_block[k + 1] - a.first_block[k];
    if (b == nb - 1 && threadIdx.x == kWG - 1) j.out[j.n] = off;
}

__global__ void k_multi_scan_empty(MultiScanArgs a) {
    if (threadIdx.x < a.njobs && a.job[threadIdx.x].n == 0) a.job[threadIdx.x].out[0] = 0;
}

void multi_scan_u64(const ScanJob *jobs, uint32_t njobs, uint64_t *sums, hipStream_t s) {
    if (!njobs) return;
    if (njobs > kMultiScanJobs) throw InvalidError("too many scan jobs in one launch");
    MultiScanArgs a{};
    a.njobs = njobs;
    a.sums = sums;
    bool empty = false;
    for (uint32_t i = 0; i < njobs; ++i) {
        if (jobs[i].n > kMultiScanMax) throw InvalidError("array too long for the multi-array scan");
        a.job[i] = jobs[i];
        if (jobs[i].div > 0xffffffffull) throw InvalidError("scan divisor above 32 bits");
        a.first_block[i + 1] = a.first_block[i] + (uint32_t)((jobs[i].n + kMsBlock - 1) / kMsBlock);
        empty |= jobs[i].n == 0;
    }
    if (empty) k_multi_scan_empty<<<1, 64, 0, s>>>(a);
    const uint32_t nb = a.first_block[njobs];
    if (nb) {
        k_multi_scan_sums<<<nb, kWG, 0, s>>>(a);
        k_multi_scan_apply<<<nb, kWG, 0, s>>>(a);
    }
    CJ_HIP(hipGetLastError());
}

uint32_t multi_scan_blocks(const ScanJob *jobs, uint32_t njobs) {
    uint64_t nb = 0;
    for (uint32_t i = 0; i < njobs; ++i) nb += (jobs[i].n + kMsBlock - 1) / kMsBlock;
    return (uint32_t)nb;
}

size_t scan_tmp_elems(uint64_t n) {
    size_t t = 0;
    while (n > kScanBlock) {
        uint64_t nb = (n + kScanBlock - 1) / kScanBlock;
        t += 2 * (nb + 1);
        n = nb;
    }
    return t + 2;
}

void scan_exclusive_u64(const uint64_t *in, uint64_t *out, uint64_t n, hipStream_t s, uint64_t *tmp,
                        size_t tmp_elems) {
    if (n == 0) {
        fill_u64(out, 0, 1, s);
        return;
    }
    uint64_t nb = (n + kScanBlock - 1) / kScanBlock;
    if (nb == 1) {
        k_scan_block<false><<<1, kWG, 0, s>>>(in, out, n, nullptr);
        CJ_HIP(hipGetLastError());
        return;
    }
    if (tmp_elems < 2 * (nb + 1)) throw DeviceError("scan scratch too small");
    uint64_t *sums = tmp, *sums_off = tmp + nb + 1;
    k_scan_block_sums<<<(unsigned)nb, kWG, 0, s>>>(in, n, sums);
    CJ_HIP(hipGetLastError());
    if (nb <= kScanDirect) {
        k_scan_block<true><<<(unsigned)nb, kWG, 0, s>>>(in, out, n, sums);
        CJ_HIP(hipGetLastError());
        return;
    }
    scan_exclusive_u64(sums, sums_off, nb, s, tmp + 2 * (nb + 1), tmp_elems - 2 * (nb + 1));
    k_scan_block<false><<<(unsigned)nb, kWG, 0, s>>>(in, out, n, sums_off);
    CJ_HIP(hipGetLastError());
}

// ------------------------------------------------------------- page index
// Rows contributed by each page (PageIndex::build, columnar_reader.h:120-160).
__global__ void k_page_row_counts(const uint8_t *pages, uint32_t num_pages, uint64_t *counts) {
    uint32_t pg = blockIdx.x * blockDim.x + threadIdx.x;
    if (pg >= num_pages) return;
    uint32_t n = ld16(pages + (size_t)pg * kPage);
    counts[pg] = n == 0xffffu ? 1u : (n == 0xfffeu ? 0u : n);
}
void page_row_counts(const uint8_t *pages, uint32_t num_pages, uint64_t *counts, hipStream_t s) {
    if (!num_pages) return;
    k_page_row_counts<<<(num_pages + kWG - 1) / kWG, kWG, 0, s>>>(pages, num_pages, counts);
    CJ_HIP(hipGetLastError());
}

// ------------------------------------------------------------ dense decode
// Segments of 512 rows (8 chunks of 64), no barriers: a page's segments go
// to as many waves as a dense page has segments (4 for INT32, 2 for 8-byte
// values), each wave taking every kDecodeWaves-th segment, so a column of
// full pages runs one segment per wave (the loads of a segment are two
// dependent round trips; one wave walking a whole page would chain eight).
// A wave starts a segment of a sparse page from the set bits of the bitmap
// bytes before it.  The bitmap bytes of the segment, then its values, are
// loaded with clamped in-page addresses, all in flight together; value
// indices come from ballots of the row bits.  Validity words wholly inside a
// page whose first row is 32-aligned are stored; the others (shared with a
// neighbouring page) are OR-ed with atomics.
constexpr uint32_t kDecodeBatch = 8;

template <int W>
__global__ __launch_bounds__(kWG) void k_decode_fixed(const uint8_t *pages, uint32_t num_pages,
                                                      const uint64_t *row_start, void *values, uint32_t *valid_bits) {
    using T = typename std::conditional<W == 4, uint32_t, uint64_t>::type;
    constexpr uint32_t kSeg = 64 * kDecodeBatch;
    constexpr uint32_t kDecodeWaves = W == 4 ? 4 : 2;  // segments of a dense page
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * (kWG / 64) + wave_id();
    const uint32_t pg = gw / kDecodeWaves;
    if (pg >= num_pages) return;
    const uint8_t *p = pages + (size_t)pg * kPage;
    const uint32_t n = ld16(p), nv = ld16(p + 2);
    const uint64_t base = row_start[pg];
    const bool dense = n == nv;
    const uint8_t *bm = p + kPage - ((n + 7) >> 3);
    const T *vals = reinterpret_cast<const T *>(p + (W == 4 ? 4 : 8));
    T *out = static_cast<T *>(values) + base;
    const uint64_t lt = lanemask_lt();
    const uint32_t nlast = n ? n - 1 : 0, vlast = nv ? nv - 1 : 0;
    const bool aligned = (base & 31) == 0;
    for (uint32_t c0 = (gw % kDecodeWaves) * kSeg; c0 < n; c0 += kDecodeWaves * kSeg) {
        uint32_t vbase = 0;  // values before the segment (sparse pages)
        if (!dense) {
            uint32_t pc = 0;
            for (uint32_t b = lane; b < c0 / 8; b += 64) pc += (uint32_t)__popc(bm[b]);
            for (int o = 32; o > 0; o >>= 1) pc += __shfl_xor(pc, o, 64);
            vbase = pc;
        }
        uint32_t okbits = 0;
        uint32_t vi[kDecodeBatch];
        if (dense) {
#pragma unroll
            for (uint32_t c = 0; c < kDecodeBatch; ++c) {
                const uint32_t r = c0 + 64 * c + lane;
                vi[c] = r;
                okbits |= (uint32_t)(r < n) << c;
            }
        } else {
            uint32_t byte[kDecodeBatch];
#pragma unroll
            for (uint32_t c = 0; c < kDecodeBatch; ++c) {
                const uint32_t r = c0 + 64 * c + lane;
                byte[c] = bm[(r < n ? r : nlast) >> 3];
            }
#pragma unroll
            for (uint32_t c = 0; c < kDecodeBatch; ++c) {
                const uint32_t r = c0 + 64 * c + lane;
                const bool ok = (r < n) & (((byte[c] >> (r & 7)) & 1u) != 0);
                const uint64_t vm = __ballot(ok);
                vi[c] = vbase + (uint32_t)__popcll(vm & lt);
                vbase += (uint32_t)__popcll(vm);
                okbits |= (uint32_t)ok << c;
            }
        }
        T v[kDecodeBatch];
#pragma unroll
        for (uint32_t c = 0; c < kDecodeBatch; ++c) v[c] = vals[vi[c] < nv ? vi[c] : vlast];
#pragma unroll
        for (uint32_t c = 0; c < kDecodeBatch; ++c) {
            const uint32_t r0 = c0 + 64 * c;
            if (r0 >= n) break;
            const uint32_t r = r0 + lane;
            // a bitmap with more set bits than n_vals is malformed: never
            // read past the page's value area
            const bool rd = ((okbits >> c) & 1u) && vi[c] < nv;
            if (r < n) out[r] = rd ? v[c] : T(0);
            const uint64_t m = __ballot(rd);
            if (aligned) {
                if (lane < 2) {
                    const uint32_t word = (uint32_t)(m >> (32 * lane));
                    uint32_t *dst = valid_bits + ((base + r0) >> 5) + lane;
                    if (r0 + 32 * lane + 32 <= n) *dst = word;
                    else if (word) atomicOr(dst, word);
                }
            } else if (lane == 0) {
                or_mask_bits(valid_bits, base + r0, m);
            }
        }
    }
}

void decode_fixed(const uint8_t *pages, uint32_t num_pages, const uint64_t *row_start, int width,
                  void *values, uint32_t *valid_bits, hipStream_t s) {
    if (!num_pages) return;
    const uint64_t waves = (uint64_t)num_pages * (width == 4 ? 4 : 2);
    const uint32_t grid = (uint32_t)((waves + kWG / 64 - 1) / (kWG / 64));
    if (width == 4) k_decode_fixed<4><<<grid, kWG, 0, s>>>(pages, num_pages, row_start, values, valid_bits);
    else k_decode_fixed<8><<<grid, kWG, 0, s>>>(pages, num_pages, row_start, values, valid_bits);
    CJ_HIP(hipGetLastError());
}

// VARCHAR: per row a packed ref (page, byte offset of the chars, length);
// a 0xffff page is one row whose ref records the length of its page run.
__global__ __launch_bounds__(kWG) void k_decode_varchar(const uint8_t *pages, uint32_t num_pages,
                                                        const uint64_t *row_start, uint64_t *refs,
                                                        uint32_t *valid_bits) {
    __shared__ uint32_t lds[4];
    const uint32_t pg = blockIdx.x;
    const uint8_t *p = pages + (size_t)pg * kPage;
    const uint32_t n = ld16(p);
    const uint64_t base = row_start[pg];
    if (n == 0xfffeu) return;
    if (n == 0xffffu) {
        if (threadIdx.x == 0) {
            uint32_t q = pg + 1;
            while (q < num_pages && ld16(pages + (size_t)q * kPage) == 0xfffeu) ++q;
            refs[base] = ((uint64_t)pg << 32) | ((uint64_t)kLongMarker << 16) | (uint64_t)(q - pg);
            atomicOr(valid_bits + (base >> 5), 1u << (base & 31));
        }
        return;
    }
    const uint32_t nv = ld16(p + 2);
    const uint8_t *bm = p + kPage - ((n + 7) >> 3);
    const uint8_t *offs = p + 4;
    const uint32_t chars = 4 + 2 * nv;
    const uint32_t max_chars = kPage - chars - ((n + 7) >> 3);  // chars end before the bitmap
    const bool dense = n == nv;
    uint32_t running = 0;
    for (uint32_t t0 = 0; t0 < n; t0 += kWG) {
        uint32_t r = t0 + threadIdx.x;
        bool inr = r < n;
        bool ok = inr && (dense || page_bit(bm, r));
        uint64_t m;
        uint32_t tot;
        uint32_t vi = running + tile_rank(ok, &m, &tot, lds);
        if (inr) {
            uint64_t ref = 0;
            if (ok && vi < nv) {
                uint32_t end = min((uint32_t)ld16(offs + 2 * vi), max_chars);
                uint32_t start = min(vi ? (uint32_t)ld16(offs + 2 * (vi - 1)) : 0u, end);
                ref = ((uint64_t)pg << 32) | ((uint64_t)(chars + start) << 16) | (uint64_t)(end - start);
            }
            refs[base + r] = ref;
        }
        if (lane_id() == 0) or_mask_bits(valid_bits, base + t0 + 64 * wave_id(), m);
        running += tot;
    }
}

void decode_varchar(const uint8_t *pages, uint32_t num_pages, const uint64_t *row_start, uint64_t *refs,
                    uint32_t *valid_bits, hipStream_t s) {
    if (!num_pages) return;
    k_decode_varchar<<<num_pages, kWG, 0, s>>>(pages, num_pages, row_start, refs, valid_bits);
    CJ_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- gathers
__global__ void k_gather_u32(const uint32_t *src, const uint32_t *idx, uint32_t *dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}
void gather_u32(const uint32_t *src, const uint32_t *idx, uint32_t *dst, uint64_t n, hipStream_t s) {
    if (!n) return;
    k_gather_u32<<<grid_for(n), kWG, 0, s>>>(src, idx, dst, n);
    CJ_HIP(hipGetLastError());
}

// Several row-id lists gathered through one index list: the index is read
// once and one launch serves every list of a join side.
__global__ void k_gather_u32_set(GatherSet g, const uint32_t *idx, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t ix = idx[i];
        uint32_t v[kGatherSet];
#pragma unroll
        for (uint32_t k = 0; k < kGatherSet; ++k) v[k] = g.src[k < g.count ? k : 0][ix];
#pragma unroll
        for (uint32_t k = 0; k < kGatherSet; ++k)
            if (k < g.count) g.dst[k][i] = v[k];
    }
}
void gather_u32_set(const GatherSet &g, const uint32_t *idx, uint64_t n, hipStream_t s) {
    if (!n || !g.count) return;
    if (g.count > kGatherSet) throw InvalidError("too many lists in one gather set");
    k_gather_u32_set<<<grid_for(n), kWG, 0, s>>>(g, idx, n);
    CJ_HIP(hipGetLastError());
}

}  // namespace cj
