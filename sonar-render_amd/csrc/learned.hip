// Learned near-field correction of fragment->fragment terms.
//
// The reference evaluates every pair with one point sample of the emitter
// fragment (src/OptiX.cu:340-354). For pairs closer than ~2 wavelengths the
// kernel varies across the fragment, so the point rule is off; the judge is a
// 16 x 16 sub-fragment quadrature of the same kernel over the delta x delta
// square centred on the emitter point (cosine culls and the r < delta/16
// regularisation applied at every sub-point).
//
// A small MLP (10 -> 64 -> 64 -> 2, tanh) learns that correction from pairs
// labelled by the quadrature, in units of the wavelength, so one network
// serves every frequency. It corrects a cheap base rule rather than the point
// value: the 2 x 2 Gauss-Legendre average of the same kernel over the cell
// (near_rules.h), whose remainder is a smooth fourth-order term. In the
// emitter's frame (x = xAxis, y = yAxis, z = normal), with u the unit vector
// emitter -> receiver:
//   features: r/lambda - 1, 14 (delta/lambda - 0.19), u.x, u.y, 2 u.z - 1,
//             n_r.x, n_r.y, n_r.z, ln(r/delta)/2, and c = -(r u . n_r)/delta
//             clamped to [-3, 3]: where the receiver's plane crosses the
//             emitter cell (the receiver cosine culls part of the cell there,
//             which is most of what the base rule misses)
//   target:   t = g (Q - G2) e^{-ikr} / M,  M = 0.9 (k/(4 pi r) + 1/r^2),
// g = kNfTargetGain,
// G2 the base rule (per unit area), for 3 delta <= r < 2 lambda.
// Pairs closer than that keep the quadrature as fallback.
// Training (forward, backward, weight gradients, Adam) and inference are the
// kernels below; inference rounds weights and activations to BF16.

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>

#include "launch.h"
#include "learned.h"
#include "near_rules.h"

namespace sr {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInv4PiL = 0.0795774715459476678f;

// Counter-based random numbers (a 32-bit integer hash of (seed, index, lane)).
__device__ inline uint32_t hash3(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u ^ (c + 0x165667B1u) * 0xC2B2AE3Du;
    h ^= h >> 16; h *= 0x7FEB352Du; h ^= h >> 15; h *= 0x846CA68Bu; h ^= h >> 16;
    return h;
}
__device__ inline float urand(uint32_t seed, uint32_t i, uint32_t k) {
    return (hash3(seed, i, k) >> 8) * (1.0f / 16777216.0f);
}

// Point-rule value without the area: cos_s cos_i 0.9 G(r), G regularised for r < reg.
__device__ inline float2 point_kernel(float r, float cs, float ci, float k, float reg) {
    float re, im;
    if (r < reg) {
        re = 0.f;
        im = k * kInv4PiL;
    } else {
        float s, c;
        sincosf(k * r, &s, &c);
        float a = k * kInv4PiL / r;
        re = a * s + 1.0f / (r * r);
        im = -a * c;
    }
    const float w = cs * ci * 0.9f;
    return make_float2(w * re, w * im);
}

// 16x16 quadrature over the emitter square (emitter frame: x, y in plane, z normal),
// per unit area. rv: receiver point relative to the emitter point; nr: receiver normal.
__device__ float2 near_quad16(float3 rv, float3 nr, bool fieldDst, float d, float k) {
    const int n = 16;
    const float h = d / n;
    float2 acc = make_float2(0.f, 0.f);
    for (int sy = 0; sy < n; sy++)
        for (int sx = 0; sx < n; sx++) {
            const float ox = (sx + 0.5f) * h - 0.5f * d, oy = (sy + 0.5f) * h - 0.5f * d;
            const float vx = rv.x - ox, vy = rv.y - oy, vz = rv.z;
            const float r = sqrtf(vx * vx + vy * vy + vz * vz);
            const float cs = vz / r;
            const float ci = fieldDst ? 1.0f : -(vx * nr.x + vy * nr.y + vz * nr.z) / r;
            if (!(cs >= 1e-6f) || !(ci >= 1e-6f)) continue;
            float2 g = point_kernel(r, cs, ci, k, h);
            acc.x += g.x;
            acc.y += g.y;
        }
    return make_float2(acc.x / (n * n), acc.y / (n * n));
}

// Training pairs in the emitter's frame, lengths in wavelengths (k = 2 pi).
__global__ void k_nf_samples(uint32_t seed, int n, float dMin, float dMax, float rMax, float *X, float *Y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float k = 2.0f * kPi, lam = 1.0f;
    const float d = dMin + (dMax - dMin) * urand(seed, i, 0);
    // distance in [3 delta, rMax], denser close in where the correction is
    // large; closer pairs are left to the quadrature
    const float u0 = urand(seed, i, 1);
    const float r0 = kNfRMinDeltas * d;
    const float r = r0 + (rMax - r0) * u0 * u0;
    float3 uv, nr;
    float cs, ci;
    // half the receivers at right angles to the emitter: right-angle junctions
    // (corner reflectors, boxes, fins) are where near pairs across facets occur
    const bool perp = urand(seed, i, 1000) < 0.5f;
    for (int tr = 0;; tr++) {
        const float zc = urand(seed, i, 2 + 4 * tr);
        const float ph = 2.0f * kPi * urand(seed, i, 3 + 4 * tr);
        const float st = sqrtf(fmaxf(0.f, 1.f - zc * zc));
        uv = make_float3(st * cosf(ph), st * sinf(ph), zc);
        const float zn = perp ? 0.f : 2.f * urand(seed, i, 4 + 4 * tr) - 1.f;
        const float pn = 2.0f * kPi * urand(seed, i, 5 + 4 * tr);
        const float sn = sqrtf(fmaxf(0.f, 1.f - zn * zn));
        nr = make_float3(sn * cosf(pn), sn * sinf(pn), zn);
        cs = uv.z;
        ci = -(uv.x * nr.x + uv.y * nr.y + uv.z * nr.z);
        if ((cs >= 1e-3f && ci >= 1e-3f) || tr > 30) break;
    }
    const float3 rv = make_float3(uv.x * r, uv.y * r, uv.z * r);
    float f[kMlpIn];
    near_features(r, lam, d, uv, nr, f);
    for (int q = 0; q < kMlpIn; q++) X[(size_t)i * kMlpIn + q] = f[q];
    const float2 Q = near_quad16(rv, nr, false, d, k);
    const float2 K = near_gauss2(rv, nr, d, k);
    const float M = 0.9f * (k * kInv4PiL / r + 1.0f / (r * r)) * (1.0f / kNfTargetGain);
    float s, c;
    sincosf(k * r, &s, &c);
    const float dre = Q.x - K.x, dim = Q.y - K.y;
    // g (dre + i dim) e^{-ikr} / M
    Y[(size_t)i * 2] = (dre * c + dim * s) / M;
    Y[(size_t)i * 2 + 1] = (dim * c - dre * s) / M;
}

// ---------------------------------------------------------------------------
// MLP: W1 [64][8], b1[64], W2 [64][64], b2[64], W3 [2][64], b3[2] packed in that order.

// A workgroup takes kMlpRows samples; thread (neuron o = t % 64, group t / 64)
// computes neuron o of 8 of them, the layer's inputs shared through LDS. Every
// sum runs in the order of the one-thread-per-sample form (the same results),
// with 64 times the threads.
constexpr int kMlpRows = 32;
constexpr int kMlpPad = 65;  // LDS row stride of a sample's 64 activations

__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ P, const float *__restrict__ X, int n,
                                                     float *__restrict__ H1, float *__restrict__ H2,
                                                     float *__restrict__ Out) {
    __shared__ float w[kMlpParams];
    __shared__ float sx[kMlpRows * kMlpIn], sh1[kMlpRows * kMlpPad], sh2[kMlpRows * kMlpPad];
    const int i0 = blockIdx.x * kMlpRows;
    for (int q = threadIdx.x; q < kMlpParams; q += blockDim.x) w[q] = P[q];
    for (int q = threadIdx.x; q < kMlpRows * kMlpIn; q += blockDim.x)
        sx[q] = i0 * kMlpIn + q < n * kMlpIn ? X[(size_t)i0 * kMlpIn + q] : 0.f;
    __syncthreads();
    const float *W1 = w, *b1 = w + kOffB1, *W2 = w + kOffW2, *b2 = w + kOffB2, *W3 = w + kOffW3, *b3 = w + kOffB3;
    const int o = threadIdx.x & 63, g = threadIdx.x >> 6;
    for (int k = 0; k < kMlpRows / 4; k++) {
        const int sl = g * (kMlpRows / 4) + k, i = i0 + sl;
        float a = b1[o];
        for (int q = 0; q < kMlpIn; q++) a = fmaf(W1[o * kMlpIn + q], sx[sl * kMlpIn + q], a);
        const float h = tanhf(a);
        sh1[sl * kMlpPad + o] = h;
        if (i < n) H1[(size_t)i * 64 + o] = h;
    }
    __syncthreads();
    for (int k = 0; k < kMlpRows / 4; k++) {
        const int sl = g * (kMlpRows / 4) + k, i = i0 + sl;
        float a = b2[o];
        for (int q = 0; q < 64; q++) a = fmaf(W2[o * 64 + q], sh1[sl * kMlpPad + q], a);
        const float h = tanhf(a);
        sh2[sl * kMlpPad + o] = h;
        if (i < n) H2[(size_t)i * 64 + o] = h;
    }
    __syncthreads();
    if (threadIdx.x < 2 * kMlpRows) {  // output c of sample sl
        const int sl = threadIdx.x >> 1, c = threadIdx.x & 1, i = i0 + sl;
        float a = b3[c];
        for (int q = 0; q < 64; q++) a = fmaf(W3[c * 64 + q], sh2[sl * kMlpPad + q], a);
        if (i < n) Out[(size_t)i * 2 + c] = a;
    }
}

// dL/dOut for L = mean over samples of |Out - Y|^2 (both components), then the
// hidden deltas D2 = (W3^T dOut) (1 - h2^2), D1 = (W2^T D2) (1 - h1^2); the
// same layout of threads over samples and neurons as the forward pass.
__global__ __launch_bounds__(256) void k_mlp_backward(const float *__restrict__ P, const float *__restrict__ Y, int n,
                                                      const float *__restrict__ H1, const float *__restrict__ H2,
                                                      const float *__restrict__ Out, float *__restrict__ G3,
                                                      float *__restrict__ D2, float *__restrict__ D1) {
    __shared__ float w[kMlpParams];
    __shared__ float sg[kMlpRows * 2], sd2[kMlpRows * kMlpPad];
    const int i0 = blockIdx.x * kMlpRows;
    for (int q = threadIdx.x; q < kMlpParams; q += blockDim.x) w[q] = P[q];
    if (threadIdx.x < 2 * kMlpRows) {
        const int sl = threadIdx.x >> 1, c = threadIdx.x & 1, i = i0 + sl;
        float gv = 0.f;
        if (i < n) {
            gv = 2.0f * (Out[(size_t)i * 2 + c] - Y[(size_t)i * 2 + c]) / (float)n;
            G3[(size_t)i * 2 + c] = gv;
        }
        sg[threadIdx.x] = gv;
    }
    __syncthreads();
    const float *W2 = w + kOffW2, *W3 = w + kOffW3;
    const int q = threadIdx.x & 63, g = threadIdx.x >> 6;
    for (int k = 0; k < kMlpRows / 4; k++) {
        const int sl = g * (kMlpRows / 4) + k, i = i0 + sl;
        const float h = i < n ? H2[(size_t)i * 64 + q] : 0.f;
        const float d2 = (W3[q] * sg[2 * sl] + W3[64 + q] * sg[2 * sl + 1]) * (1.0f - h * h);
        sd2[sl * kMlpPad + q] = d2;
        if (i < n) D2[(size_t)i * 64 + q] = d2;
    }
    __syncthreads();
    for (int k = 0; k < kMlpRows / 4; k++) {
        const int sl = g * (kMlpRows / 4) + k, i = i0 + sl;
        if (i >= n) continue;
        float a = 0.f;
        for (int o = 0; o < 64; o++) a = fmaf(W2[o * 64 + q], sd2[sl * kMlpPad + o], a);
        const float h = H1[(size_t)i * 64 + q];
        D1[(size_t)i * 64 + q] = a * (1.0f - h * h);
    }
}

// Batch sums of the parameter gradients in two fixed-order stages: thread
// (parameter p, chunk c) sums the chunk's samples in FP64, then one thread per
// parameter adds the chunks in order.
constexpr int kGradChunks = 32;

__global__ void k_mlp_wgrad(const float *__restrict__ X, const float *__restrict__ H1, const float *__restrict__ H2,
                            const float *__restrict__ G3, const float *__restrict__ D2, const float *__restrict__ D1,
                            int n, double *__restrict__ partial) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (p >= kMlpParams) return;
    const int per = (n + kGradChunks - 1) / kGradChunks;
    const int i0 = c * per, i1 = min(n, i0 + per);
    double acc = 0.0;
    if (p < kOffB1) {  // W1[o][q]
        const int o = p / kMlpIn, q = p % kMlpIn;
#pragma unroll 8
        for (int i = i0; i < i1; i++) acc += (double)D1[(size_t)i * 64 + o] * X[(size_t)i * kMlpIn + q];
    } else if (p < kOffW2) {  // b1
        const int o = p - kOffB1;
#pragma unroll 8
        for (int i = i0; i < i1; i++) acc += D1[(size_t)i * 64 + o];
    } else if (p < kOffB2) {  // W2[o][q]
        const int o = (p - kOffW2) / 64, q = (p - kOffW2) % 64;
#pragma unroll 8
        for (int i = i0; i < i1; i++) acc += (double)D2[(size_t)i * 64 + o] * H1[(size_t)i * 64 + q];
    } else if (p < kOffW3) {
        const int o = p - kOffB2;
#pragma unroll 8
        for (int i = i0; i < i1; i++) acc += D2[(size_t)i * 64 + o];
    } else if (p < kOffB3) {  // W3[o][q]
        const int o = (p - kOffW3) / 64, q = (p - kOffW3) % 64;
#pragma unroll 8
        for (int i = i0; i < i1; i++) acc += (double)G3[(size_t)i * 2 + o] * H2[(size_t)i * 64 + q];
    } else {
        const int o = p - kOffB3;
#pragma unroll 8
        for (int i = i0; i < i1; i++) acc += G3[(size_t)i * 2 + o];
    }
    partial[(size_t)c * kMlpParams + p] = acc;
}

__global__ void k_mlp_wgrad_sum(const double *__restrict__ partial, float *__restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    double acc = 0.0;
    for (int c = 0; c < kGradChunks; c++) acc += partial[(size_t)c * kMlpParams + p];
    grad[p] = (float)acc;
}

__global__ void k_adam(float *__restrict__ P, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       float lr, float b1, float b2, float eps, int t) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    m[p] = b1 * m[p] + (1.f - b1) * g[p];
    v[p] = b2 * v[p] + (1.f - b2) * g[p] * g[p];
    const float mh = m[p] / (1.f - powf(b1, (float)t)), vh = v[p] / (1.f - powf(b2, (float)t));
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void k_sq_err(const float *__restrict__ Out, const float *__restrict__ Y, int n, float *__restrict__ acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = Out[2 * i] - Y[2 * i], b = Out[2 * i + 1] - Y[2 * i + 1];
    atomicAdd(acc, a * a + b * b);
    atomicAdd(acc + 1, Y[2 * i] * Y[2 * i] + Y[2 * i + 1] * Y[2 * i + 1]);
}

// ---------------------------------------------------------------------------

int nf_samples(uint32_t seed, int n, float dMin, float dMax, float rMax, float *X, float *Y, hipStream_t st) {
    k_nf_samples<<<(n + 127) / 128, 128, 0, st>>>(seed, n, dMin, dMax, rMax, X, Y);
    SR_HIP(hipGetLastError());
    return 0;
}

int mlp_forward(const float *P, const float *X, int n, float *H1, float *H2, float *Out, hipStream_t st) {
    k_mlp_forward<<<(n + kMlpRows - 1) / kMlpRows, 256, 0, st>>>(P, X, n, H1, H2, Out);
    SR_HIP(hipGetLastError());
    return 0;
}

int mlp_backward(const float *P, const float *X, const float *Y, int n, const float *H1, const float *H2,
                 const float *Out, float *G3, float *D2, float *D1, double *partial, float *grad, hipStream_t st) {
    k_mlp_backward<<<(n + kMlpRows - 1) / kMlpRows, 256, 0, st>>>(P, Y, n, H1, H2, Out, G3, D2, D1);
    SR_HIP(hipGetLastError());
    k_mlp_wgrad<<<dim3((kMlpParams + 127) / 128, kGradChunks), 128, 0, st>>>(X, H1, H2, G3, D2, D1, n, partial);
    SR_HIP(hipGetLastError());
    k_mlp_wgrad_sum<<<(kMlpParams + 255) / 256, 256, 0, st>>>(partial, grad);
    SR_HIP(hipGetLastError());
    return 0;
}

int mlp_adam(float *P, const float *g, float *m, float *v, float lr, int t, hipStream_t st) {
    k_adam<<<(kMlpParams + 255) / 256, 256, 0, st>>>(P, g, m, v, lr, 0.9f, 0.999f, 1e-8f, t);
    SR_HIP(hipGetLastError());
    return 0;
}

int mlp_sq_err(const float *Out, const float *Y, int n, float *acc2, hipStream_t st) {
    k_sq_err<<<(n + 255) / 256, 256, 0, st>>>(Out, Y, n, acc2);
    SR_HIP(hipGetLastError());
    return 0;
}

}  // namespace sr
