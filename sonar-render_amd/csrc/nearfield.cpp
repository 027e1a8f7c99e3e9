// Host side of the learned near-field correction: parameter initialisation,
// the training loop on the GPU (the kernels of learned.hip), and the state the
// transfer-matrix build reads.
#include "nearfield.h"

#include <algorithm>
#include <cmath>
#include <vector>

#include "launch.h"
#include "learned.h"

namespace sr {

NearField &nearfield() {
    static NearField nf;
    return nf;
}

namespace {

template <class T>
int dalloc(T **p, size_t n) {
    *p = nullptr;
    SR_HIP(hipMalloc((void **)p, n * sizeof(T)));
    return 0;
}
void dfree(void *p) {
    if (p) (void)hipFree(p);
}

// Xavier-uniform weights, zero biases, from a 64-bit LCG.
std::vector<float> init_params(uint32_t seed) {
    std::vector<float> P(kMlpParams, 0.f);
    uint64_t s = 0x9E3779B97F4A7C15ull ^ seed;
    auto u = [&]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (float)((s >> 40) * (1.0 / 16777216.0));
    };
    auto fill = [&](int off, int fanIn, int fanOut, int n) {
        const float a = std::sqrt(6.0f / (fanIn + fanOut));
        for (int i = 0; i < n; i++) P[off + i] = (2.f * u() - 1.f) * a;
    };
    fill(0, kMlpIn, 64, 64 * kMlpIn);
    fill(kOffW2, 64, 64, 4096);
    fill(kOffW3, 64, 2, 128);
    return P;
}

}  // namespace

int nf_ensure(NearField &nf) {
    if (!nf.P) {
        if (dalloc(&nf.P, kMlpParams)) return -1;
        std::vector<float> P = init_params(1);
        SR_HIP(hipMemcpy(nf.P, P.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    }
    return 0;
}

int nf_init(uint32_t seed) {
    NearField &nf = nearfield();
    if (nf_ensure(nf)) return -1;
    SR_HIP(hipDeviceSynchronize());  // queued projections may still read the parameters
    std::vector<float> P = init_params(seed);
    SR_HIP(hipMemcpy(nf.P, P.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    nf.trained = false;
    nf.gen++;
    return 0;
}

int nf_set_params(const float *in, int n) {
    if (n != kMlpParams) { set_error("near-field MLP has 4994 parameters"); return -2; }
    NearField &nf = nearfield();
    if (nf_ensure(nf)) return -1;
    SR_HIP(hipDeviceSynchronize());  // queued projections may still read the parameters
    SR_HIP(hipMemcpy(nf.P, in, n * sizeof(float), hipMemcpyHostToDevice));
    nf.trained = true;
    nf.gen++;
    return 0;
}

int nf_get_params(float *out, int n) {
    if (n < kMlpParams) { set_error("buffer shorter than 4994 parameters"); return -2; }
    NearField &nf = nearfield();
    if (nf_ensure(nf)) return -1;
    SR_HIP(hipMemcpy(out, nf.P, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int nf_samples_host(uint32_t seed, int n, float *X, float *Y) {
    float *dX, *dY;
    if (dalloc(&dX, (size_t)n * kMlpIn) || dalloc(&dY, (size_t)n * 2)) return -1;
    int rc = nf_samples(seed, n, kNfDeltaMin, kNfDeltaMax, kNfRadius, dX, dY, nullptr);
    if (!rc) {
        SR_HIP(hipMemcpy(X, dX, (size_t)n * kMlpIn * sizeof(float), hipMemcpyDeviceToHost));
        SR_HIP(hipMemcpy(Y, dY, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost));
    }
    dfree(dX);
    dfree(dY);
    return rc;
}

// Forward (and optionally the full gradient) of the current network on host data.
int nf_eval(const float *X, const float *Y, int n, float *out, float *grad, const float *params) {
    NearField &nf = nearfield();
    if (nf_ensure(nf)) return -1;
    float *dP = nf.P, *dX, *dY = nullptr, *H1, *H2, *O, *G3, *D2, *D1, *dg;
    bool own = false;
    if (params) {
        if (dalloc(&dP, kMlpParams)) return -1;
        SR_HIP(hipMemcpy(dP, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
        own = true;
    }
    if (dalloc(&dX, (size_t)n * kMlpIn) || dalloc(&H1, (size_t)n * 64) || dalloc(&H2, (size_t)n * 64) ||
        dalloc(&O, (size_t)n * 2) || dalloc(&G3, (size_t)n * 2) || dalloc(&D2, (size_t)n * 64) ||
        dalloc(&D1, (size_t)n * 64) || dalloc(&dg, kMlpParams))
        return -1;
    SR_HIP(hipMemcpy(dX, X, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice));
    int rc = mlp_forward(dP, dX, n, H1, H2, O, nullptr);
    double *part = nullptr;
    if (!rc && grad) {
        if (dalloc(&dY, (size_t)n * 2) || dalloc(&part, (size_t)kMlpGradPartials)) return -1;
        SR_HIP(hipMemcpy(dY, Y, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice));
        rc = mlp_backward(dP, dX, dY, n, H1, H2, O, G3, D2, D1, part, dg, nullptr);
    }
    SR_HIP(hipDeviceSynchronize());
    if (!rc && out) SR_HIP(hipMemcpy(out, O, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost));
    if (!rc && grad) SR_HIP(hipMemcpy(grad, dg, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    for (void *p : {(void *)dX, (void *)dY, (void *)H1, (void *)H2, (void *)O, (void *)G3, (void *)D2, (void *)D1,
                    (void *)dg, (void *)part})
        dfree(p);
    if (own) dfree(dP);
    return rc;
}

// Adam on minibatches of GPU-labelled pairs; cosine learning-rate decay.
int nf_train(int nsamples, int steps, int batch, float lr, uint32_t seed, float *report) {
    NearField &nf = nearfield();
    if (nf_ensure(nf)) return -1;
    SR_HIP(hipDeviceSynchronize());  // queued projections may still read the parameters
    if (nsamples < batch || batch <= 0 || steps <= 0) { set_error("nf_train: bad sizes"); return -2; }
    // the parameters change in place from here on: not usable until training succeeds
    nf.trained = false;
    nf.gen++;
    const int nval = 16384;
    float *X, *Y, *Xv, *Yv, *H1, *H2, *O, *G3, *D2, *D1, *g, *m, *v, *acc;
    if (dalloc(&X, (size_t)nsamples * kMlpIn) || dalloc(&Y, (size_t)nsamples * 2) || dalloc(&Xv, (size_t)nval * kMlpIn) ||
        dalloc(&Yv, (size_t)nval * 2) || dalloc(&H1, (size_t)std::max(batch, nval) * 64) ||
        dalloc(&H2, (size_t)std::max(batch, nval) * 64) || dalloc(&O, (size_t)std::max(batch, nval) * 2) ||
        dalloc(&G3, (size_t)batch * 2) || dalloc(&D2, (size_t)batch * 64) || dalloc(&D1, (size_t)batch * 64) ||
        dalloc(&g, kMlpParams) || dalloc(&m, kMlpParams) || dalloc(&v, kMlpParams) || dalloc(&acc, 4))
        return -1;
    double *part = nullptr;
    if (dalloc(&part, (size_t)kMlpGradPartials)) return -1;
    int rc = nf_samples(seed, nsamples, kNfDeltaMin, kNfDeltaMax, kNfRadius, X, Y, nullptr);
    if (!rc) rc = nf_samples(seed ^ 0xA5A5A5A5u, nval, kNfDeltaMin, kNfDeltaMax, kNfRadius, Xv, Yv, nullptr);
    SR_HIP(hipMemset(m, 0, kMlpParams * sizeof(float)));
    SR_HIP(hipMemset(v, 0, kMlpParams * sizeof(float)));
    const int nb = nsamples / batch;
    for (int t = 1; !rc && t <= steps; t++) {
        const int b = (t - 1) % nb;
        const float *xb = X + (size_t)b * batch * kMlpIn, *yb = Y + (size_t)b * batch * 2;
        const float l = lr * (0.05f + 0.95f * 0.5f * (1.0f + std::cos(3.14159265f * (float)(t - 1) / steps)));
        rc = mlp_forward(nf.P, xb, batch, H1, H2, O, nullptr);
        if (!rc) rc = mlp_backward(nf.P, xb, yb, batch, H1, H2, O, G3, D2, D1, part, g, nullptr);
        if (!rc) rc = mlp_adam(nf.P, g, m, v, l, t, nullptr);
    }
    auto relerr = [&](const float *xs, const float *ys, int n, float &out) -> int {
        SR_HIP(hipMemset(acc, 0, 4 * sizeof(float)));
        if (mlp_forward(nf.P, xs, n, H1, H2, O, nullptr) || mlp_sq_err(O, ys, n, acc, nullptr)) return -1;
        float h[2];
        SR_HIP(hipMemcpy(h, acc, 2 * sizeof(float), hipMemcpyDeviceToHost));
        out = std::sqrt(h[0] / std::max(h[1], 1e-30f));
        return 0;
    };
    float tr = -1.f, va = -1.f;
    if (!rc) rc = relerr(X, Y, std::min(nsamples, nval), tr);
    if (!rc) rc = relerr(Xv, Yv, nval, va);
    SR_HIP(hipDeviceSynchronize());
    for (void *p : {(void *)X, (void *)Y, (void *)Xv, (void *)Yv, (void *)H1, (void *)H2, (void *)O, (void *)G3,
                    (void *)D2, (void *)D1, (void *)g, (void *)m, (void *)v, (void *)acc, (void *)part})
        dfree(p);
    if (rc) return rc;
    std::vector<float> chk(kMlpParams);
    SR_HIP(hipMemcpy(chk.data(), nf.P, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    for (float x : chk)
        if (!std::isfinite(x)) { set_error("near-field training diverged (non-finite parameters)"); return -1; }
    if (!std::isfinite(tr) || !std::isfinite(va)) { set_error("near-field training: non-finite error"); return -1; }
    nf.trained = true;
    nf.trainErr = tr;
    nf.valErr = va;
    if (report) { report[0] = tr; report[1] = va; }
    return 0;
}

}  // namespace sr
