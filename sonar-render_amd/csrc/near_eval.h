// Near-field values of fragment pairs closer than the near radius, shared by
// the transfer-matrix build (transfer.hip) and the batched scenes
// (batch_near.hip): the 16 x 16 sub-fragment quadrature of the pair kernel over
// the emitter's cell, and the learned correction (2 x 2 Gauss base rule plus a
// 10-64-64-2 tanh MLP on the BF16 matrix cores).
//
// Both evaluators read pairs and hand back values through a source object:
//   NearGeom geom(long i) const       pair i in the emitter's frame
//   bool want(long i) const           the quadrature evaluates pair i at all
//   void put(long i, float re, float im) const
// where (re, im) is cos_s cos_i 0.9 G per unit emitter area (the reference's
// A cos_s cos_i 0.9 G of src/OptiX.cu:340-354 without the factor A).
#pragma once
#include <hip/hip_runtime.h>

#include "learned.h"
#include "near_rules.h"

namespace sr {

constexpr float kInv4PiT = 0.0795774715459476678f;

// Receiver point relative to the emitter point and receiver normal, both in the
// emitter's frame (x = xAxis, y = yAxis, z = emitter normal), and the distance.
struct NearGeom {
    float3 rv, nr;
    float r;
};

// One row (sy) of the 16 x 16 midpoint quadrature of the pair kernel over the
// emitter's cell, unnormalised. The phase of each sub-point uses the hardware
// sine/cosine in revolutions (r < 2 lambda here, so r k / 2 pi < 2 keeps
// ~1e-7 of a revolution). alpha: the attenuation (Im k); the kernel then
// carries e^{-alpha r} as the reference's thrust::exp(i k r) does (OptiX.cu:350).
__device__ inline float2 quad16_row(float3 rv, float3 nr, float d, float k, float alpha, int sy) {
    const int n = 16;
    const float h = d / n;
    const float krev = k * 0.159154943091895336f;
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll 4
    for (int sx = 0; sx < n; sx++) {
        const float vx = rv.x - ((sx + 0.5f) * h - 0.5f * d), vy = rv.y - ((sy + 0.5f) * h - 0.5f * d);
        const float vz = rv.z;
        const float r = sqrtf(vx * vx + vy * vy + vz * vz);
        const float cs = vz / r, ci = -(vx * nr.x + vy * nr.y + vz * nr.z) / r;
        if (!(cs >= 1e-6f) || !(ci >= 1e-6f)) continue;
        float re, im;
        if (r < h) {  // ik / (4 pi), k = kr + i alpha
            re = -alpha * kInv4PiT;
            im = k * kInv4PiT;
        } else {
            float x = r * krev;
            x -= rintf(x);
            const float s = __builtin_amdgcn_sinf(x), c = __builtin_amdgcn_cosf(x);
            const float a = kInv4PiT / r * (alpha != 0.0f ? __expf(-alpha * r) : 1.0f);
            re = a * fmaf(alpha, c, k * s) + 1.0f / (r * r);
            im = -a * fmaf(k, c, -alpha * s);
        }
        const float w = cs * ci * 0.9f;
        acc.x += w * re;
        acc.y += w * im;
    }
    return acc;
}

// Pairs by 16 x 16 quadrature: 16 lanes per pair, one row of sub-points each,
// summed across the lanes. Block bid of 256 threads covers pairs 16 bid .. +15.
template <class Src>
__device__ inline void near_quad_block(const Src &src, long count, float d, float k, float alpha, int bid) {
    const long i = ((long)bid * blockDim.x + threadIdx.x) >> 4;
    const int sy = threadIdx.x & 15;
    const bool valid = i < count && src.want(i);  // uniform over the pair's 16 lanes
    float2 v = make_float2(0.f, 0.f);
    if (valid) {
        const NearGeom q = src.geom(i);
        v = quad16_row(q.rv, q.nr, d, k, alpha, sy);
    }
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) {
        v.x += __shfl_xor(v.x, m);
        v.y += __shfl_xor(v.y, m);
    }
    if (valid && sy == 0) src.put(i, v.x * (1.0f / 256.0f), v.y * (1.0f / 256.0f));
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int kHP = 72;  // bf16 per activation row in LDS (64 + 8: 16-byte aligned rows)

// tanh x = 1 - 2 / (e^{2x} + 1) on the hardware exp2 and reciprocal (a few ulp
// in FP32, far below the BF16 rounding of the activations it feeds); +-1 at the
// extremes (e^{2x} -> inf or 0), no cancellation issue at the BF16 scale.
__device__ inline float tanh_hw(float x) {
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(x * 2.8853900817779268f) + 1.0f);
}

// Pairs by the learned correction: value = G2 + (M / g) t e^{ikr} with G2 the
// 2 x 2 Gauss base rule (near_rules.h), t from the 10-64-64-2 tanh MLP in BF16
// (weights and activations rounded to BF16, FP32 accumulation) on the BF16
// matrix cores (v_mfma_f32_32x32x16_bf16). A wave evaluates 32 pairs:
// features -> LDS -> A operand of layer 1; each hidden layer's accumulators go
// through LDS (tanh, BF16) to become the next A operand; the 64 -> 2 output
// layer is a per-lane dot product. The layers' B operands are built once per
// workgroup in fragment order in LDS. Workgroups bid = 0 .. nblk-1 of 256
// threads stride over tiles of 128 pairs.
template <class Src>
__device__ inline void near_mlp_blocks(const Src &src, long count, const float *P, float d, float kr, float alpha,
                                       int bid, int nblk) {
    __shared__ bf16x8 sW1[2][64];
    __shared__ bf16x8 sW2[4][2][64];
    __shared__ float sB1[64], sB2[64], sW3[2][64], sB3[2];
    __shared__ __attribute__((aligned(16))) __bf16 sH[4][32][kHP];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int h = lane >> 5, j = lane & 31;
    // layer B fragments: lane l holds B[k 8 (l/32) .. +7][col l % 32] of its tile
    for (int idx = threadIdx.x; idx < 2 * 64; idx += 256) {
        const int c = idx >> 6, l = idx & 63, col = 32 * c + (l & 31), k0 = 8 * (l >> 5);
        bf16x8 v;
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = (__bf16)(k0 + q < kMlpIn ? P[col * kMlpIn + k0 + q] : 0.0f);
        sW1[c][l] = v;
    }
    for (int idx = threadIdx.x; idx < 4 * 2 * 64; idx += 256) {
        const int st = idx >> 7, c = (idx >> 6) & 1, l = idx & 63, col = 32 * c + (l & 31), k0 = 16 * st + 8 * (l >> 5);
        bf16x8 v;
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = (__bf16)P[kOffW2 + col * 64 + k0 + q];
        sW2[st][c][l] = v;
    }
    for (int idx = threadIdx.x; idx < 64; idx += 256) {
        sB1[idx] = (float)(__bf16)P[kOffB1 + idx];
        sB2[idx] = (float)(__bf16)P[kOffB2 + idx];
        sW3[0][idx] = (float)(__bf16)P[kOffW3 + idx];
        sW3[1][idx] = (float)(__bf16)P[kOffW3 + 64 + idx];
    }
    if (threadIdx.x < 2) sB3[threadIdx.x] = (float)(__bf16)P[kOffB3 + threadIdx.x];
    const float lam = 2.0f * 3.14159265358979f / kr;
    __bf16 (*H)[kHP] = sH[wave];
    // workgroup-uniform trip count: tiles of 128 pairs (32 per wave)
    for (long base = (long)bid * 128; base < count; base += (long)nblk * 128) {
        const long i = base + wave * 32 + j;
        const bool valid = i < count;
        float2 K = make_float2(0.f, 0.f);
        float M = 0.f, cs_kr = 1.f, sn_kr = 0.f;
        __syncthreads();  // the previous tile's reads of H are done
        // both halves of the wave take pair j: each evaluates two points of the
        // base rule the network corrects; half 0 also the features and e^{ikr}
        bf16x8 f0, f1;
#pragma unroll
        for (int q = 0; q < 8; q++) { f0[q] = (__bf16)0.0f; f1[q] = (__bf16)0.0f; }
        float2 Kh = make_float2(0.f, 0.f);
        if (valid) {
            const NearGeom q = src.geom(i);
            Kh = near_gauss2_half(q.rv, q.nr, d, kr, h, alpha);
            if (h == 0) {
                const float r = q.r;
                // the network's correction is the attenuation-free one; in an attenuating
                // medium it is carried by e^{-alpha r} at the pair's distance (the factor's
                // variation across the cell, alpha delta, is ~1e-3 of the correction)
                M = 0.9f * (kr * kInv4PiT / r + 1.0f / (r * r)) * (1.0f / kNfTargetGain);
                if (alpha != 0.0f) M *= expf(-alpha * r);
                sincosf(kr * r, &sn_kr, &cs_kr);
                float f[kMlpIn];
                near_features(r, lam, d, make_float3(q.rv.x / r, q.rv.y / r, q.rv.z / r), q.nr, f);
#pragma unroll
                for (int qq = 0; qq < 8; qq++) f0[qq] = (__bf16)f[qq];
#pragma unroll
                for (int qq = 8; qq < kMlpIn; qq++) f1[qq - 8] = (__bf16)f[qq];
            }
        }
        K = make_float2(0.25f * (Kh.x + __shfl_xor(Kh.x, 32)), 0.25f * (Kh.y + __shfl_xor(Kh.y, 32)));
        if (h == 0) {
            *reinterpret_cast<bf16x8 *>(&H[j][0]) = f0;
            *reinterpret_cast<bf16x8 *>(&H[j][8]) = f1;
        }
        __syncthreads();
        // layer 1: [32 x 16] x [16 x 64]
        f32x16 acc[2];
        {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&H[j][8 * h]);
#pragma unroll
            for (int c = 0; c < 2; c++) acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, sW1[c][lane], f32x16{}, 0, 0, 0);
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int row = (e >> 2) * 8 + h * 4 + (e & 3), col = 32 * c + j;
                H[row][col] = (__bf16)tanh_hw(acc[c][e] + sB1[col]);
            }
        __syncthreads();
        // layer 2: [32 x 64] x [64 x 64]
#pragma unroll
        for (int c = 0; c < 2; c++) acc[c] = f32x16{};
#pragma unroll
        for (int st = 0; st < 4; st++) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&H[j][16 * st + 8 * h]);
#pragma unroll
            for (int c = 0; c < 2; c++) acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, sW2[st][c][lane], acc[c], 0, 0, 0);
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int row = (e >> 2) * 8 + h * 4 + (e & 3), col = 32 * c + j;
                H[row][col] = (__bf16)tanh_hw(acc[c][e] + sB2[col]);
            }
        __syncthreads();
        // layer 3: lane (j, h) -> output h of pair j
        float o = sB3[h];
#pragma unroll
        for (int k8 = 0; k8 < 8; k8++) {
            const bf16x8 hv = *reinterpret_cast<const bf16x8 *>(&H[j][8 * k8]);
#pragma unroll
            for (int q = 0; q < 8; q++) o = fmaf(sW3[h][8 * k8 + q], (float)hv[q], o);
        }
        const float o1 = __shfl(o, j + 32);
        // G2 + (M / g) t e^{ikr}
        if (h == 0 && valid) src.put(i, K.x + M * (o * cs_kr - o1 * sn_kr), K.y + M * (o * sn_kr + o1 * cs_kr));
    }
}

}  // namespace sr
