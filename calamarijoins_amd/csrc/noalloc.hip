// noalloc.hip -- the device scratch array behind cj_execute_noalloc.
//
// A file-scope device array of CJ_NOALLOC_SCRATCH_BYTES: every GPU the
// library runs on holds its own copy until the process ends.  Host code
// never passes the array's name to a kernel (in host code it is a stand-in,
// not a device address): noalloc_init asks the runtime for the address of
// this GPU's copy once, with the GPU current, and the executor hands that
// address to the kernels as the backing of all its temporaries.
#include "noalloc.hpp"

#include <map>
#include <memory>

namespace cj {

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_mu;
std::map<int, uint8_t *> g_ptr;                     // device -> address of its scratch
std::map<int, std::unique_ptr<std::mutex>> g_call;  // device -> call serialization
}  // namespace

void noalloc_init(int device) {
    std::lock_guard<std::mutex> g(g_mu);
    if (g_ptr.count(device)) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw DeviceError(std::string("hipGetSymbolAddress of the no-allocation scratch failed: ") +
                          hipGetErrorString(e));
    g_ptr[device] = static_cast<uint8_t *>(p);
    g_call[device] = std::make_unique<std::mutex>();
}

uint8_t *noalloc_scratch(int device) {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_ptr.find(device);
    if (it == g_ptr.end()) throw InvalidError("no-allocation scratch not set up on this device");
    return it->second;
}

std::mutex &noalloc_mutex(int device) {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_call.find(device);
    if (it == g_call.end()) throw InvalidError("no-allocation scratch not set up on this device");
    return *it->second;
}

}  // namespace cj
