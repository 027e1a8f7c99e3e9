// Fixed device scratch of the allocation-free entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace sr {

// 64 MiB per GPU: the split-K partial sums of one field-point projection
// (two halves of 32 MiB, 4 Mi complex FP32 values each).
constexpr size_t kNoAllocScratchBytes = 64ull << 20;

// The current device's copy of the scratch array (its address from
// hipGetSymbolAddress, looked up once per device) and its size in float2.
int noalloc_scratch(float2 **ptr, size_t *elems);

}  // namespace sr
