// Base rule of the learned near-field correction, shared by the training
// labels (learned.hip) and inference (transfer.hip) so that both see the same
// function: the pair kernel cos_s cos_i 0.9 G(r) of src/OptiX.cu:297-354 (per
// unit emitter area, no r < delta regularisation) averaged over the 2 x 2
// Gauss-Legendre points of the emitter's delta x delta cell. The rule is exact
// for bicubic variation across the cell, so what it misses of the 16 x 16
// sub-fragment quadrature is a fourth-order remainder, ~20 times smaller (RMS,
// for delta = lambda/8..lambda/4 and 3 delta <= r < 2 lambda) than what the
// one-point rule misses; the network learns that remainder.
#pragma once
#include <hip/hip_runtime.h>

#include "learned.h"

namespace sr {

// Kernel at emitter sub-point offset (receiver at (vx, vy, vz) from it), in
// the emitter's frame (z = emitter normal); cosine culls as the reference's.
// k: Re k; alpha: Im k (the attenuation, src/Model.cpp:52-53), so that
// -e^{ikr} ik / (4 pi r) = e^{-alpha r} / (4 pi r) [(alpha cos + k sin) + i (alpha sin - k cos)].
__device__ inline float2 near_kernel_at(float vx, float vy, float vz, float3 nr, float k, float alpha = 0.0f) {
    const float r = sqrtf(vx * vx + vy * vy + vz * vz);
    const float cs = vz / r, ci = -(vx * nr.x + vy * nr.y + vz * nr.z) / r;
    if (!(cs >= 1e-6f) || !(ci >= 1e-6f)) return make_float2(0.f, 0.f);
    float s, c;
    sincosf(k * r, &s, &c);
    const float w = cs * ci * 0.9f;
    if (alpha == 0.0f) {
        const float a = k * 0.0795774715459476678f / r;
        return make_float2(w * (a * s + 1.0f / (r * r)), -w * a * c);
    }
    const float a = expf(-alpha * r) * 0.0795774715459476678f / r;
    return make_float2(w * (a * (alpha * c + k * s) + 1.0f / (r * r)), w * a * (alpha * s - k * c));
}

// rv, nr: receiver point (relative to the emitter point) and receiver normal,
// emitter frame; d: cell size; k: wavenumber.
__device__ inline float2 near_gauss2(float3 rv, float3 nr, float d, float k) {
    const float o = 0.288675134594812882f * d;  // d / (2 sqrt 3)
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float ox = (q & 1) ? o : -o, oy = (q & 2) ? o : -o;
        const float2 v = near_kernel_at(rv.x - ox, rv.y - oy, rv.z, nr, k);
        acc.x += v.x;
        acc.y += v.y;
    }
    return make_float2(0.25f * acc.x, 0.25f * acc.y);
}

// Points 2 half and 2 half + 1 of the same rule, summed, unscaled: the two
// halves' sum times 1/4 is near_gauss2 (up to the order of the additions).
__device__ inline float2 near_gauss2_half(float3 rv, float3 nr, float d, float k, int half, float alpha = 0.0f) {
    const float o = 0.288675134594812882f * d;
    const float oy = half ? o : -o;
    const float2 a = near_kernel_at(rv.x + o, rv.y - oy, rv.z, nr, k, alpha);
    const float2 b = near_kernel_at(rv.x - o, rv.y - oy, rv.z, nr, k, alpha);
    return make_float2(a.x + b.x, a.y + b.y);
}

// The network's inputs (lengths in any unit, lam the wavelength; u the unit
// vector emitter -> receiver and nr the receiver normal, emitter frame).
__device__ inline float plane_offset(float r, float d, float3 u, float3 nr) {
    return fminf(fmaxf(-r * (u.x * nr.x + u.y * nr.y + u.z * nr.z) / d, -3.0f), 3.0f);
}

__device__ inline void near_features(float r, float lam, float d, float3 u, float3 nr, float *f) {
    f[0] = r / lam - 1.0f;
    f[1] = (d / lam - 0.19f) * 14.0f;
    f[2] = u.x;
    f[3] = u.y;
    f[4] = 2.0f * u.z - 1.0f;
    f[5] = nr.x;
    f[6] = nr.y;
    f[7] = nr.z;
    f[8] = 0.5f * logf(r / d);
    f[9] = plane_offset(r, d, u, nr);
}
static_assert(kMlpIn == 10, "near_features writes 10 inputs");

}  // namespace sr
