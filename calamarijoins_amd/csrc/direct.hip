// direct.hip -- direct-mapped build side for unique, dense INT32 keys.
//
// When every non-NULL build key is distinct and the keys span a range of at
// most ~2x the build rows (primary keys are the common case), the unchained
// hash table (hashtable.h:63-505) is replaced by a table indexed by
// key - min: a presence bitmap (one bit per key of the range, small enough
// to stay in L2) answers the counting pass, and a record array holds, per
// key, the build side's projected payload values with their NULL flags, so
// the probe's write pass fetches one record per match instead of a
// directory entry, a key, a row id and one value per build column.  The
// join result is identical: a probe key matches exactly the build row
// carrying that key.  Duplicate keys are detected while the table is built
// and the executor then falls back to the hash table.
#include "cj_internal.hpp"
#include "device_util.hpp"
#include "direct.hpp"

namespace cj {

__global__ void k_key_minmax(const int32_t *keys, const uint32_t *valid, uint64_t n, int32_t *mm) {
    __shared__ int32_t smin[kWG], smax[kWG];
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (uint64_t i = blockIdx.x * (uint64_t)kWG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kWG)
        if (valid_bit(valid, i)) {
            int32_t k = keys[i];
            lo = k < lo ? k : lo;
            hi = k > hi ? k : hi;
        }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int s = kWG / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = min(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = max(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMin(&mm[0], smin[0]);
        atomicMax(&mm[1], smax[0]);
    }
}

void key_minmax(const int32_t *keys, const uint32_t *valid, uint64_t n, int32_t *mm, hipStream_t s) {
    if (!n) return;
    uint64_t g = (n + kWG - 1) / kWG;
    k_key_minmax<<<(unsigned)(g < 2048 ? g : 2048), kWG, 0, s>>>(keys, valid, n, mm);
    CJ_HIP(hipGetLastError());
}

// One thread per build row: set the key's presence bit (a bit already set
// means a duplicate key) and write its record.
__global__ void k_direct_build(DirectBuildArgs a) {
    const uint64_t i = blockIdx.x * (uint64_t)kWG + threadIdx.x;
    if (i >= a.rows) return;
    if (!valid_bit(a.key_valid, i)) return;
    const uint64_t slot = (uint64_t)((int64_t)a.keys[i] - (int64_t)a.min_key);
    const uint32_t bit = 1u << (slot & 31);
    const uint32_t old = atomicOr(&a.presence[slot >> 5], bit);
    if (old & bit) {
        atomicOr(a.dup_flag, 1u);
        return;
    }
    uint8_t *rec = a.records + slot * a.stride;
    uint32_t flags = 0;
    for (uint32_t f = 0; f < a.nfields; ++f) {
        const DirectField &fd = a.fields[f];
        const bool ok = valid_bit(fd.valid_bits, i);
        flags |= (uint32_t)ok << f;
        if (fd.width == 4) *reinterpret_cast<uint32_t *>(rec + fd.offset) = ok ? static_cast<const uint32_t *>(fd.values)[i] : 0u;
        else *reinterpret_cast<uint64_t *>(rec + fd.offset) = ok ? static_cast<const uint64_t *>(fd.values)[i] : 0ull;
    }
    *reinterpret_cast<uint32_t *>(rec + a.flags_offset) = flags;
}

void direct_build(const DirectBuildArgs &a, hipStream_t s) {
    if (!a.rows) return;
    k_direct_build<<<(unsigned)((a.rows + kWG - 1) / kWG), kWG, 0, s>>>(a);
    CJ_HIP(hipGetLastError());
}

}  // namespace cj
