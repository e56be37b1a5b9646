// noalloc.hpp -- the fixed device scratch of cj_execute_noalloc: one array
// per GPU, reached through the address hipGetSymbolAddress gives on that
// GPU, never through the array's own name in host code.
#pragma once

#include "../../include/cj_api.h"
#include "cj_internal.hpp"

#include <mutex>

namespace cj {

constexpr size_t kNoAllocScratchBytes = CJ_NOALLOC_SCRATCH_BYTES;

// Looks up this GPU's scratch address once (the device must be current).
void noalloc_init(int device);
// The scratch address of `device` (noalloc_init ran for it).
uint8_t *noalloc_scratch(int device);
// Serializes the calls that use one GPU's scratch.
std::mutex &noalloc_mutex(int device);

}  // namespace cj
