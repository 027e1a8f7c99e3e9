// Learned near-field correction: process-wide state and host entry points.
#pragma once
#include <cstdint>

namespace sr {

// Training distribution, in wavelengths: fragment edge delta and pair distance.
constexpr float kNfDeltaMin = 0.12f;
constexpr float kNfDeltaMax = 0.27f;
constexpr float kNfRadius = 2.0f;

struct NearField {
    float *P = nullptr;   // device parameters (kMlpParams FP32)
    bool trained = false;
    float trainErr = -1.f, valErr = -1.f;
    // bumped whenever the parameters change: transfer matrices holding learned
    // near terms record the generation they were built with
    unsigned gen = 0;
};

NearField &nearfield();
int nf_ensure(NearField &nf);
int nf_init(uint32_t seed);
int nf_set_params(const float *in, int n);
int nf_get_params(float *out, int n);
int nf_samples_host(uint32_t seed, int n, float *X, float *Y);
int nf_eval(const float *X, const float *Y, int n, float *out, float *grad, const float *params);
int nf_train(int nsamples, int steps, int batch, float lr, uint32_t seed, float *report);

}  // namespace sr
