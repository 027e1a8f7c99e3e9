// Learned near-field correction: training and inference launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sr {

// 10 -> 64 -> 64 -> 2 MLP, parameters packed W1[64][10] b1 W2[64][64] b2 W3[2][64] b3.
constexpr int kMlpIn = 10;
constexpr int kMlpParams = 64 * kMlpIn + 64 + 64 * 64 + 64 + 2 * 64 + 2;  // 4994
constexpr int kOffB1 = 64 * kMlpIn, kOffW2 = kOffB1 + 64, kOffB2 = kOffW2 + 4096, kOffW3 = kOffB2 + 64,
              kOffB3 = kOffW3 + 128;

// The network's target is the base rule's remainder relative to M, times this
// gain (the remainder is ~2e-4 of M, RMS; the gain brings the targets to order one).
constexpr float kNfTargetGain = 4096.0f;
// Training pairs lie at 3 or more fragment sizes: closer pairs are left to the
// quadrature (the default quadrature radius of the learned mode).
constexpr float kNfRMinDeltas = 3.0f;

// Training pairs with quadrature labels; lengths in wavelengths.
int nf_samples(uint32_t seed, int n, float dMin, float dMax, float rMax, float *X, float *Y, hipStream_t st);
int mlp_forward(const float *P, const float *X, int n, float *H1, float *H2, float *Out, hipStream_t st);
// partial: kMlpGradPartials doubles of scratch for the two-stage batch sums.
constexpr int kMlpGradPartials = 32 * kMlpParams;
int mlp_backward(const float *P, const float *X, const float *Y, int n, const float *H1, const float *H2,
                 const float *Out, float *G3, float *D2, float *D1, double *partial, float *grad, hipStream_t st);
int mlp_adam(float *P, const float *g, float *m, float *v, float lr, int t, hipStream_t st);
int mlp_sq_err(const float *Out, const float *Y, int n, float *acc2, hipStream_t st);

}  // namespace sr
