// Device helpers shared by the transfer form's kernels (transfer.hip: near
// field, B build, product; build_s.hip: the S builds).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sr {

// Two floats as one word of FP16 halves (low: a, high: b), saturated to the
// FP16 range (near-field terms can exceed the |S| bound).
__device__ inline uint32_t pack_h2(float a, float b) {
    a = fminf(fmaxf(a, -65000.f), 65000.f);
    b = fminf(fmaxf(b, -65000.f), 65000.f);
    _Float16 ha = (_Float16)a, hb = (_Float16)b;
    return (uint32_t)__builtin_bit_cast(uint16_t, ha) | ((uint32_t)__builtin_bit_cast(uint16_t, hb) << 16);
}

// Non-temporal 16-byte accesses (streamed past the caches), for an S larger
// than every cache level: written once per build and read once per bounce.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ inline void store_nt(uint4 *p, uint4 v) {
    __builtin_nontemporal_store(u32x4{v.x, v.y, v.z, v.w}, reinterpret_cast<u32x4 *>(p));
}
__device__ inline uint4 load_nt(const uint4 *p) {
    const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}

}  // namespace sr
