// membench.hip -- memory microbenchmarks (include/cj_membench.h): a
// streaming copy and random gathers of 4/8/16-byte records, the two access
// patterns the join kernels are made of (page streams; direct-map record and
// hash-directory lookups).
#include "../../include/cj_api.h"
#include "../../include/cj_membench.h"
#include "cj_internal.hpp"

#include <algorithm>
#include <vector>
#include <string>

namespace cj {

__global__ __launch_bounds__(kWG) void k_mb_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * kWG;
    for (uint64_t i = (uint64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

constexpr uint32_t kMbBatch = 8;  // loads in flight per lane

template <int W>
__global__ __launch_bounds__(kWG) void k_mb_gather(const uint8_t *__restrict__ table, const uint32_t *__restrict__ idx,
                                                   uint64_t n, uint32_t *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kWG;
    uint32_t acc = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * kWG + threadIdx.x; base < n; base += stride * kMbBatch) {
        uint32_t ix[kMbBatch];
#pragma unroll
        for (uint32_t b = 0; b < kMbBatch; ++b) {
            const uint64_t i = base + b * stride;
            ix[b] = idx[i < n ? i : n - 1];  // clamped: every load of the batch issues together
        }
#pragma unroll
        for (uint32_t b = 0; b < kMbBatch; ++b) {
            const uint8_t *p = table + (uint64_t)ix[b] * W;
            if (W == 16) {
                const uint4 v = *reinterpret_cast<const uint4 *>(p);
                acc ^= v.x ^ v.y ^ v.z ^ v.w;
            } else if (W == 8) {
                const uint2 v = *reinterpret_cast<const uint2 *>(p);
                acc ^= v.x ^ v.y;
            } else {
                acc ^= *reinterpret_cast<const uint32_t *>(p);
            }
        }
    }
    out[(uint64_t)blockIdx.x * kWG + threadIdx.x] = acc;
}

// 16-byte records with a cache policy (mode 1: non-temporal loads, marked
// to leave the L2 first) or routed by table slice to the XCDs (mode 2: the
// lookups come grouped by slice, slice x = the x-th eighth of the table, and
// workgroup b takes slice b % 8 -- workgroups are dealt to the eight XCDs
// in turn, so each XCD's L2 sees one eighth of the table).
typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kWG) void k_mb_gather16_nt(const uint4 *__restrict__ t, const uint32_t *__restrict__ idx,
                                                        uint64_t n, uint32_t *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kWG;
    uint32_t acc = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * kWG + threadIdx.x; base < n; base += stride * kMbBatch) {
        uint32_t ix[kMbBatch];
#pragma unroll
        for (uint32_t b = 0; b < kMbBatch; ++b) {
            const uint64_t i = base + b * stride;
            ix[b] = idx[i < n ? i : n - 1];
        }
#pragma unroll
        for (uint32_t b = 0; b < kMbBatch; ++b) {
            const u32x4v v = __builtin_nontemporal_load(reinterpret_cast<const u32x4v *>(t + ix[b]));
            acc ^= v[0] ^ v[1] ^ v[2] ^ v[3];
        }
    }
    out[(uint64_t)blockIdx.x * kWG + threadIdx.x] = acc;
}

__global__ __launch_bounds__(kWG) void k_mb_gather16_sliced(const uint4 *__restrict__ t, const uint32_t *__restrict__ idx,
                                                            const uint64_t *__restrict__ slice, uint32_t *__restrict__ out) {
    const uint32_t x = blockIdx.x & 7, bi = blockIdx.x >> 3, nb = gridDim.x >> 3;
    const uint64_t lo = slice[x], n = slice[x + 1];
    const uint64_t stride = (uint64_t)nb * kWG;
    uint32_t acc = 0;
    for (uint64_t base = lo + (uint64_t)bi * kWG + threadIdx.x; base < n; base += stride * kMbBatch) {
        uint32_t ix[kMbBatch];
#pragma unroll
        for (uint32_t b = 0; b < kMbBatch; ++b) {
            const uint64_t i = base + b * stride;
            ix[b] = idx[i < n ? i : n - 1];
        }
#pragma unroll
        for (uint32_t b = 0; b < kMbBatch; ++b) {
            const uint4 v = t[ix[b]];
            acc ^= v.x ^ v.y ^ v.z ^ v.w;
        }
    }
    out[(uint64_t)blockIdx.x * kWG + threadIdx.x] = acc;
}

namespace {

struct DevBuf {
    void *p = nullptr;
    explicit DevBuf(size_t n) { CJ_HIP(hipMalloc(&p, n)); }
    ~DevBuf() { (void)hipFree(p); }
};
struct Events {
    hipEvent_t a, b;
    Events() {
        CJ_HIP(hipEventCreate(&a));
        CJ_HIP(hipEventCreate(&b));
    }
    ~Events() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

template <class F>
double best_of(uint32_t reps, F &&launch) {
    Events e;
    double best = 1e30;
    launch();  // warm
    CJ_HIP(hipDeviceSynchronize());
    for (uint32_t r = 0; r < std::max<uint32_t>(reps, 1); ++r) {
        CJ_HIP(hipEventRecord(e.a, nullptr));
        launch();
        CJ_HIP(hipEventRecord(e.b, nullptr));
        CJ_HIP(hipEventSynchronize(e.b));
        float ms = 0;
        CJ_HIP(hipEventElapsedTime(&ms, e.a, e.b));
        best = std::min(best, (double)ms);
    }
    return best;
}

template <class F>
int guarded(F &&f) {
    try {
        f();
        return CJ_OK;
    } catch (const InvalidError &e) {
        set_last_error(e.what());
        return CJ_ERR_INVALID;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return CJ_ERR_DEVICE;
    }
}
}  // namespace

}  // namespace cj

extern "C" int cj_membench_copy(int32_t device, uint64_t bytes, uint32_t reps, double *best_ms) {
    using namespace cj;
    return guarded([&] {
        if (!best_ms || bytes < 16) throw InvalidError("bad argument");
        CJ_HIP(hipSetDevice(device));
        DevBuf a(bytes), b(bytes);
        CJ_HIP(hipMemset(a.p, 1, bytes));
        const uint64_t n = bytes / 16;
        *best_ms = best_of(reps, [&] {
            k_mb_copy<<<4096, kWG>>>(static_cast<const uint4 *>(a.p), static_cast<uint4 *>(b.p), n);
            CJ_HIP(hipGetLastError());
        });
    });
}

extern "C" int cj_membench_gather(int32_t device, const uint32_t *idx, uint64_t n, uint64_t table_bytes,
                                  uint32_t record_bytes, uint32_t reps, double *best_ms) {
    using namespace cj;
    return guarded([&] {
        if (!best_ms || !idx || !n) throw InvalidError("bad argument");
        if (record_bytes != 4 && record_bytes != 8 && record_bytes != 16) throw InvalidError("record bytes 4, 8 or 16");
        const uint64_t slots = table_bytes / record_bytes;
        for (uint64_t i = 0; i < n; ++i)  // every load stays inside the table
            if (idx[i] >= slots) throw InvalidError("index beyond the table");
        CJ_HIP(hipSetDevice(device));
        DevBuf t(table_bytes + 16), d_idx(n * 4);
        const uint32_t grid = 8192;
        DevBuf out((size_t)grid * kWG * 4);
        CJ_HIP(hipMemset(t.p, 7, table_bytes));
        CJ_HIP(hipMemcpy(d_idx.p, idx, n * 4, hipMemcpyHostToDevice));
        const uint8_t *tp = static_cast<const uint8_t *>(t.p);
        const uint32_t *ip = static_cast<const uint32_t *>(d_idx.p);
        uint32_t *op = static_cast<uint32_t *>(out.p);
        *best_ms = best_of(reps, [&] {
            if (record_bytes == 16) k_mb_gather<16><<<grid, kWG>>>(tp, ip, n, op);
            else if (record_bytes == 8) k_mb_gather<8><<<grid, kWG>>>(tp, ip, n, op);
            else k_mb_gather<4><<<grid, kWG>>>(tp, ip, n, op);
            CJ_HIP(hipGetLastError());
        });
    });
}

extern "C" int cj_membench_gather16_mode(int32_t device, const uint32_t *idx, uint64_t n, uint64_t table_bytes,
                                         uint32_t mode, uint32_t reps, double *best_ms) {
    using namespace cj;
    return guarded([&] {
        if (!best_ms || !idx || !n || mode > 2) throw InvalidError("bad argument");
        const uint64_t slots = table_bytes / 16;
        if (!slots) throw InvalidError("table smaller than a record");
        for (uint64_t i = 0; i < n; ++i)  // every load stays inside the table
            if (idx[i] >= slots) throw InvalidError("index beyond the table");
        // mode 2: the lookups grouped by slice (stable), with the slice bounds
        std::vector<uint32_t> grouped;
        std::vector<uint64_t> bounds(9, 0);
        const uint32_t *src = idx;
        if (mode == 2) {
            grouped.reserve(n);
            for (uint32_t x = 0; x < 8; ++x) {
                bounds[x] = grouped.size();
                for (uint64_t i = 0; i < n; ++i)
                    if ((uint64_t)idx[i] * 8 / slots == x) grouped.push_back(idx[i]);
            }
            bounds[8] = grouped.size();
            src = grouped.data();
        }
        CJ_HIP(hipSetDevice(device));
        DevBuf t(table_bytes + 16), d_idx(n * 4), d_sl(9 * 8);
        const uint32_t grid = 8192;  // a multiple of 8: every slice gets the same workgroups
        DevBuf out((size_t)grid * kWG * 4);
        CJ_HIP(hipMemset(t.p, 7, table_bytes));
        CJ_HIP(hipMemcpy(d_idx.p, src, n * 4, hipMemcpyHostToDevice));
        CJ_HIP(hipMemcpy(d_sl.p, bounds.data(), 9 * 8, hipMemcpyHostToDevice));
        const uint4 *tp = static_cast<const uint4 *>(t.p);
        const uint32_t *ip = static_cast<const uint32_t *>(d_idx.p);
        uint32_t *op = static_cast<uint32_t *>(out.p);
        *best_ms = best_of(reps, [&] {
            if (mode == 0) k_mb_gather<16><<<grid, kWG>>>(reinterpret_cast<const uint8_t *>(tp), ip, n, op);
            else if (mode == 1) k_mb_gather16_nt<<<grid, kWG>>>(tp, ip, n, op);
            else k_mb_gather16_sliced<<<grid, kWG>>>(tp, ip, static_cast<const uint64_t *>(d_sl.p), op);
            CJ_HIP(hipGetLastError());
        });
    });
}
