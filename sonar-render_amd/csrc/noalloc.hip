// Fixed device scratch of the allocation-free entry points.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops) get entry points whose partial sums go to this
// file-scope array instead of library allocations. Every GPU the library runs
// on holds its own copy, at its own address, until the process ends; the host
// never uses the array's name as a device address, only the address
// hipGetSymbolAddress returns for the current device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "launch.h"
#include "noalloc.h"

namespace sr {

static __device__ __align__(256) float2 scratch[kNoAllocScratchBytes / sizeof(float2)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex gScrMu;
void *gScrPtr[kMaxDevices] = {};
}  // namespace

int noalloc_scratch(float2 **ptr, size_t *elems) {
    int dev = -1;
    SR_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("allocation-free scratch: device index out of range");
        return -1;
    }
    std::lock_guard<std::mutex> g(gScrMu);
    if (!gScrPtr[dev]) {  // once per device, with that device current
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("allocation-free scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return -1;
        }
        gScrPtr[dev] = p;
    }
    *ptr = static_cast<float2 *>(gScrPtr[dev]);
    *elems = kNoAllocScratchBytes / sizeof(float2);
    return 0;
}

}  // namespace sr
