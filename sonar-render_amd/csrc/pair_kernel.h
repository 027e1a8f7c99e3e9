// Device helpers shared by the FP32 pair evaluations (gathers, exact
// recomputation of gated rows): the complex kernel of one ordered pair.
//
// Reference sums (src/OptiX.cu): fragment->point :112-190, fragment->fragment
// :192-406 (G = -e^{ikr}/(4 pi) (ik/r) + 1/r^2, ik/(4 pi) for r < delta,
// damping 0.9, both cosines). Phase in revolutions: x = r * k/(2 pi) reduced to
// [-1/2, 1/2] before v_sin_f32 / v_cos_f32.
#pragma once
#include <hip/hip_runtime.h>

#include "sums.h"

namespace sr {

constexpr float kInv4Pi = 0.0795774715459476678f;

__device__ inline float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline float hw_sin_rev(float x) { return __builtin_amdgcn_sinf(x); }
__device__ inline float hw_cos_rev(float x) { return __builtin_amdgcn_cosf(x); }
__device__ inline float hw_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

// Fractional revolutions of r * (hi + lo), error ~1e-7 rev for r * k/(2pi) < 2^22.
__device__ inline float phase_rev_split(float r, float hi, float lo) {
    float p = r * hi;
    float e = fmaf(r, hi, -p);
    float fp = p - rintf(p);
    return fp + fmaf(r, lo, e);
}
__device__ inline float phase_rev_fast(float r, float hi) {
    float p = r * hi;
    return p - rintf(p);
}

// TwoSum accumulation (Knuth): s + c exactly as hi + lo.
__device__ inline void two_sum_acc(float &s, float &lo, float c) {
    float t = s + c;
    float bp = t - s;
    lo += (s - (t - bp)) + (c - bp);
    s = t;
}

// One pair's complex kernel value K (before multiplying by the emitter term).
template <int KIND, bool ATT>
__device__ __attribute__((always_inline)) inline void pair_kernel(const Phys &ph, float4 rp, float4 rn, float4 ep, float4 en, bool bit, float &Kre,
                                   float &Kim) {
    float vx = rp.x - ep.x, vy = rp.y - ep.y, vz = rp.z - ep.z;
    float r2 = vx * vx + vy * vy + vz * vz;
    float rinv = fast_rsq(r2);
    float r = r2 * rinv;
    float de = vx * en.x + vy * en.y + vz * en.z;
    float w;
    if (KIND == KIND_F2F) {
        float dr = vx * rn.x + vy * rn.y + vz * rn.z;
        w = -(de * dr) * (rinv * rinv);
    } else {
        w = de * rinv;
    }
    float x = (KIND == KIND_F2P) ? phase_rev_split(r, ph.krevHi, ph.krevLo) : phase_rev_fast(r, ph.krevHi);
    float c = hw_cos_rev(x), s = hw_sin_rev(x);
    float a = rinv * kInv4Pi;
    if (ATT) a *= hw_exp2(-ph.alphaLog2e * r);
    float Gre, Gim;
    if (ATT) {
        Gre = a * fmaf(ph.alpha, c, ph.kr * s) + rinv * rinv;
        Gim = -a * fmaf(ph.kr, c, -ph.alpha * s);
    } else {
        Gre = (a * ph.kr) * s + rinv * rinv;
        Gim = -(a * ph.kr) * c;
    }
    if (KIND != KIND_F2P) {
        bool reg = r < ph.deltaReg;
        Gre = reg ? -ph.alpha * kInv4Pi : Gre;
        Gim = reg ? ph.kr * kInv4Pi : Gim;
        w *= ph.damp;
    }
    Kre = bit ? w * Gre : 0.0f;
    Kim = bit ? w * Gim : 0.0f;
}

template <int NB>
__device__ inline void cmac(float2 &acc, float Kre, float Kim, float2 X) {
    acc.x = fmaf(Kre, X.x, fmaf(-Kim, X.y, acc.x));
    acc.y = fmaf(Kre, X.y, fmaf(Kim, X.x, acc.y));
}

}  // namespace sr
