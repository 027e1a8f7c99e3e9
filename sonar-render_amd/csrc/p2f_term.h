// Point->fragment terms (src/OptiX.cu:35-110): one (cell, source) term
// cos_inc * A_s * e^{ikr} / r, shared by the single-scene kernels (p2f.hip)
// and the batched scenes' (sums.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "sums.h"
#include "pair_kernel.h"

namespace sr {

// Difference vector (cell - source), r = |v|, 1 / r and the receiver cosine
// cos_inc = (-v / r) . n, exactly as the reference forms them (no contraction;
// the same operations as the visibility kernels' ray set-up and cull), 1 where
// the receiver has no normal.
struct P2FGeom {
    float vx, vy, vz, r, inv, ci;
};
__device__ __attribute__((always_inline)) inline P2FGeom p2f_geom(const P2FTask &t, float4 cp, float4 cn, float4 sp) {
    P2FGeom g;
    g.vx = __fsub_rn(cp.x, sp.x);
    g.vy = __fsub_rn(cp.y, sp.y);
    g.vz = __fsub_rn(cp.z, sp.z);
    g.r = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(g.vx, g.vx), __fmul_rn(g.vy, g.vy)), __fmul_rn(g.vz, g.vz)));
    g.inv = __fdiv_rn(1.0f, g.r);
    g.ci = 1.0f;
    if (t.useNormal) {
        const float ux = -__fmul_rn(g.vx, g.inv), uy = -__fmul_rn(g.vy, g.inv), uz = -__fmul_rn(g.vz, g.inv);
        g.ci = __fadd_rn(__fadd_rn(__fmul_rn(ux, cn.x), __fmul_rn(uy, cn.y)), __fmul_rn(uz, cn.z));
    }
    return g;
}

template <bool ATT>
__device__ __attribute__((always_inline)) inline float2 p2f_value(const Phys &ph, const P2FGeom &g, float2 A) {
    const float x = phase_rev_split(g.r, ph.krevHi, ph.krevLo);
    const float cc = hw_cos_rev(x), ss = hw_sin_rev(x);
    float m = g.ci * g.inv;
    if (ATT) m *= hw_exp2(-ph.alphaLog2e * g.r);
    // m * (cc + i ss) * A, its fused multiply-adds written out so that every
    // translation unit (contracting or not) rounds it the same way
    return make_float2(m * fmaf(cc, A.x, -ss * A.y), m * fmaf(cc, A.y, ss * A.x));
}

template <bool ATT>
__device__ __attribute__((always_inline)) inline float2 p2f_term(const P2FTask &t, const Phys &ph, float4 cp, float4 cn,
                                                                 int s) {
    return p2f_value<ATT>(ph, p2f_geom(t, cp, cn, t.sp[s]), t.samp[s]);
}

// point->fragment: lanes own cells; sources are few.
template <bool ATT>
__device__ __attribute__((always_inline)) inline void p2f_elem(const P2FTask &t, const Phys &ph, int c) {
    const int lane = threadIdx.x & 63;
    const int cb = c >> 6;
    if (cb * 64 >= t.ncell) return;
    const bool valid = c < t.ncell;
    float4 cp = valid ? t.cp[c] : make_float4(0.f, 0.f, 0.f, 0.f);
    float4 cn = valid ? t.cn[c] : make_float4(0.f, 0.f, 0.f, 0.f);
    float2 sum = make_float2(0.f, 0.f);
    for (int s = 0; s < t.ns; s++) {
        uint64_t word = t.mask[(size_t)cb * t.ns + s];
        bool bit = (word >> lane) & 1ull;
        float2 contrib = make_float2(0.f, 0.f);
        if (bit) contrib = p2f_term<ATT>(t, ph, cp, cn, s);
        if (t.batched) {
            if (valid) {
                float2 &P = t.P[s * t.Pstride + c];
                P.x += contrib.x; P.y += contrib.y;
            }
        } else {
            sum.x += contrib.x; sum.y += contrib.y;
        }
    }
    if (!t.batched && valid) {
        float2 &P = t.P[c];
        P.x += sum.x; P.y += sum.y;
    }
}

}  // namespace sr
