// device_util.hpp -- wave64 / workgroup primitives shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace cj {

// Four dwords read with one 16-byte load from a 4-byte-aligned address.
typedef uint32_t u32x4u __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }
// wave index in a 1-D block, as a scalar (all lanes of a wave share it)
__device__ __forceinline__ uint32_t wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }
__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = lane_id();
    return l ? (~0ull >> (64 - l)) : 0ull;
}

// Global address space.  A pointer read from device memory (a descriptor
// array) is generic, and accesses through generic pointers compile to flat_*
// instructions, which also count against the LDS wait counter: every LDS wait
// would then wait for outstanding HBM traffic.  glb() gives such a pointer
// the global address space, so its accesses are global_* instructions.
#define CJ_G __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ CJ_G T *glb(T *p) {
    return (CJ_G T *)p;
}
// The same for a generic pointer into LDS: ds_* instructions, and the
// compiler knows its accesses cannot alias global memory (so LDS loads may
// move above global stores).
#define CJ_L __attribute__((address_space(3)))
template <class T>
__device__ __forceinline__ CJ_L T *lds(T *p) {
    return (CJ_L T *)p;
}

__device__ __forceinline__ uint16_t ld16(const uint8_t *p) { return *reinterpret_cast<const uint16_t *>(p); }
__device__ __forceinline__ int32_t ld32(const uint8_t *p) { return *reinterpret_cast<const int32_t *>(p); }
__device__ __forceinline__ uint64_t ld64(const uint8_t *p) { return *reinterpret_cast<const uint64_t *>(p); }
__device__ __forceinline__ uint16_t ld16(const CJ_G uint8_t *p) { return *reinterpret_cast<const CJ_G uint16_t *>(p); }
__device__ __forceinline__ int32_t ld32(const CJ_G uint8_t *p) { return *reinterpret_cast<const CJ_G int32_t *>(p); }
__device__ __forceinline__ uint64_t ld64(const CJ_G uint8_t *p) { return *reinterpret_cast<const CJ_G uint64_t *>(p); }

// Bit r of a page's NULL bitmap (LSB first, stored in the last bytes).
template <class P>
__device__ __forceinline__ bool page_bit(P bm, uint32_t r) {
    return (bm[r >> 3] >> (r & 7)) & 1u;
}

template <class P>
__device__ __forceinline__ bool valid_bit(P bits, uint64_t row) {
    return (bits[row >> 5] >> (row & 31)) & 1u;
}

// Inclusive prefix sum across the 64 lanes of a wave.
template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T o = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += o;
    }
    return v;
}

// Exclusive prefix sum across a 256-thread workgroup; `lds` holds 4 T.
// All threads must call it.  Returns the exclusive prefix; *total = sum.
template <class T>
__device__ __forceinline__ T block_excl_scan(T v, T *total, T *lds) {
    T inc = wave_incl_scan(v);
    uint32_t w = wave_id();
    if (lane_id() == 63) lds[w] = inc;
    __syncthreads();
    T pre = 0, tot = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        T x = lds[i];
        pre += (i < w) ? x : T(0);
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

// Ballot-based compaction index inside a 256-thread tile: returns the
// number of set predicates before this thread in the tile; *tile_total is
// the tile's count.  `lds` holds 4 uint32.
__device__ __forceinline__ uint32_t tile_rank(bool pred, uint64_t *mask_out, uint32_t *tile_total,
                                              uint32_t *lds) {
    uint64_t m = __ballot(pred);
    uint32_t w = wave_id();
    if (lane_id() == 0) lds[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        uint32_t x = lds[i];
        pre += (i < w) ? x : 0u;
        tot += x;
    }
    __syncthreads();
    *tile_total = tot;
    *mask_out = m;
    return pre + (uint32_t)__popcll(m & lanemask_lt());
}

// OR a wave's 64-row validity mask into a global bitmap starting at row g0.
__device__ __forceinline__ void or_mask_bits(uint32_t *bits, uint64_t g0, uint64_t m) {
    if (!m) return;
    uint64_t word = g0 >> 5;
    uint32_t sh = (uint32_t)(g0 & 31);
    // (m << sh) drops the top sh bits of m; they land in the third word.
    uint64_t lo = m << sh;
    uint32_t w0 = (uint32_t)lo;
    uint32_t w1 = (uint32_t)(lo >> 32);
    uint32_t w2 = sh ? (uint32_t)(m >> (64 - sh)) : 0u;
    if (w0) atomicOr(bits + word, w0);
    if (w1) atomicOr(bits + word + 1, w1);
    if (w2) atomicOr(bits + word + 2, w2);
}

}  // namespace cj
