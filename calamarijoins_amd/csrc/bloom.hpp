// bloom.hpp -- the 16-bit Bloom tag table of the hash-table directory.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace cj {

// The 2048 smallest 16-bit integers with exactly five bits set, ascending:
// the closed form of the reference's BLOOM_TAGS (bloom_tags.h:35-263).
struct BloomTable {
    uint16_t t[2048];
};
constexpr BloomTable make_bloom() {
    BloomTable b{};
    uint32_t n = 0;
    for (uint32_t v = 0; v < 65536u && n < 2048u; ++v) {
        uint32_t c = 0;
        for (uint32_t x = v; x; x &= x - 1) ++c;
        if (c == 5) b.t[n++] = (uint16_t)v;
    }
    return b;
}
// one copy per translation unit that probes or builds
static __constant__ BloomTable c_bloom = make_bloom();

}  // namespace cj
