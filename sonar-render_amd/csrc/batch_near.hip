// Near-field correction of the batched scenes' bounces (see batch_near.h).
//
// The gathers of a batched bounce evaluate every visible pair by the
// reference's point rule (src/OptiX.cu:340-354). For pairs closer than the near
// radius this file lists them per scene in CSR order (receiver rows, emitters
// ascending, filled by one lane per row so the order is fixed), evaluates the
// near value by the 16 x 16 quadrature or the learned correction (the
// evaluators of near_eval.h) and stores it minus the point rule, and each
// bounce adds sum_j (value - point) A_j P_in[j] per row as one more split-K
// partial: emitters on lower facets to the ungated part, the others to the
// part the zero-input gate drops (OptiX.cu:241-245).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "batch_near.h"
#include "launch.h"
#include "near_eval.h"

namespace sr {

namespace {

// r^2 of a pair in a fixed order of operations (the list build and every
// evaluation agree on which pairs are near).
__device__ inline float near_r2(float4 a, float4 b) {
    const float vx = a.x - b.x, vy = a.y - b.y, vz = a.z - b.z;
    return fmaf(vz, vz, fmaf(vy, vy, vx * vx));
}

// The reference's point rule per unit emitter area, from the pair's geometry.
__device__ inline float2 point_rule(const NearGeom &q, float kr, float alpha, float delta) {
    const float r = q.r;
    const float cs = q.rv.z / r, ci = -(q.rv.x * q.nr.x + q.rv.y * q.nr.y + q.rv.z * q.nr.z) / r;
    const float w = 0.9f * cs * ci;
    if (r < delta) return make_float2(w * (-alpha * kInv4PiT), w * (kr * kInv4PiT));  // ik / (4 pi)
    float s, c;
    sincosf(kr * r, &s, &c);
    const float a = kInv4PiT / r * (alpha != 0.0f ? expf(-alpha * r) : 1.0f);
    const float re = a * fmaf(alpha, c, kr * s) + 1.0f / (r * r);
    const float im = -a * fmaf(kr, c, -alpha * s);
    return make_float2(w * re, w * im);
}

// One scene's CSR pairs from `base` on, as a source for the evaluators.
// quadOnly: the quadrature pass of the learned mode (pairs below quadR2 only);
// mlpPass: the learned pass (pairs at or beyond quadR2 only).
struct BatchNearSrc {
    const NearBatchTask *t;
    long base;
    bool quadOnly, mlpPass;
    __device__ NearGeom geom(long i) const {
        const long k = base + i;
        const int row = t->rows[k], j = t->cols[k];
        const float4 rp = t->p[row], rn = t->nrm[row], ep = t->p[j], en = t->nrm[j];
        const float vx = rp.x - ep.x, vy = rp.y - ep.y, vz = rp.z - ep.z;
        const int fe = __float_as_int(en.w);
        const float4 xa = t->frames[2 * fe], ya = t->frames[2 * fe + 1];
        NearGeom q;
        q.rv = make_float3(vx * xa.x + vy * xa.y + vz * xa.z, vx * ya.x + vy * ya.y + vz * ya.z,
                           vx * en.x + vy * en.y + vz * en.z);
        q.nr = make_float3(rn.x * xa.x + rn.y * xa.y + rn.z * xa.z, rn.x * ya.x + rn.y * ya.y + rn.z * ya.z,
                           rn.x * en.x + rn.y * en.y + rn.z * en.z);
        q.r = sqrtf(q.rv.x * q.rv.x + q.rv.y * q.rv.y + q.rv.z * q.rv.z);
        return q;
    }
    __device__ bool close(long i) const {
        const long k = base + i;
        return near_r2(t->p[t->rows[k]], t->p[t->cols[k]]) < t->quadR2;
    }
    __device__ bool want(long i) const { return !quadOnly || close(i); }
    __device__ void put(long i, float re, float im) const {
        if (mlpPass && close(i)) return;  // the quadrature pass owns it
        const float2 pt = point_rule(geom(i), t->kr, t->alpha, t->delta);
        t->vals[base + i] = make_float2(re - pt.x, im - pt.y);
    }
};

}  // namespace

// One wave per (scene, 64-row block); lane = row. The emitters are walked in
// order, so each row's pairs land in ascending emitter order: 64 at a time,
// lane k loading emitter e0 + k's visibility word and position (one coalesced
// load each), then the emitters with a non-zero word in ascending order, their
// word and position read from lane k (a walk that loaded each emitter in turn
// waited a memory round trip per emitter).
__global__ __launch_bounds__(256) void k_near_rows(const NearBatchTask *__restrict__ tasks,
                                                   const int2 *__restrict__ items, int nitems) {
    const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (w >= nitems) return;  // uniform over the wave
    const int2 it = items[w];
    const NearBatchTask &t = tasks[it.x];
    const int lane = threadIdx.x & 63, row = it.y * 64 + lane;
    const bool rv = row < t.n;
    const float4 rp = rv ? t.p[row] : make_float4(0.f, 0.f, 0.f, 0.f);
    const bool fill = t.rowStart != nullptr;
    const int k0 = (fill && rv) ? t.rowStart[row] : 0;
    const uint64_t *mw = t.mask + (size_t)it.y * t.n;
    int cnt = 0;
    for (int e0 = 0; e0 < t.n; e0 += 64) {
        const int el = e0 + lane;
        const uint64_t wl = el < t.n ? mw[el] : 0ull;
        const float4 pl = wl ? t.p[el] : make_float4(0.f, 0.f, 0.f, 0.f);
        for (uint64_t nz = __ballot(wl != 0ull); nz; nz &= nz - 1) {
            const int k = __builtin_ctzll(nz);  // uniform
            const uint64_t word = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(wl >> 32), k) << 32) |
                                  (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)wl, k);
            const float4 ep = make_float4(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(pl.x), k)),
                                          __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pl.y), k)),
                                          __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pl.z), k)),
                                          __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pl.w), k)));
            if (rv && ((word >> lane) & 1ull) && near_r2(rp, ep) < t.radius2) {
                if (fill) {
                    t.rows[k0 + cnt] = row;
                    t.cols[k0 + cnt] = e0 + k;
                }
                cnt++;
            }
        }
    }
    if (!fill && rv) t.rowCount[row] = cnt;
}

// Learned values: one block per item of 128 pairs.
__global__ __launch_bounds__(256) void k_near_vals_mlp(const NearBatchTask *__restrict__ tasks, const float *mlp,
                                                       const int2 *__restrict__ items) {
    const int2 it = items[blockIdx.x];
    const NearBatchTask *t = tasks + it.x;
    const long total = t->rowStart[t->n];
    const long count = std::min<long>(128, total - it.y);
    near_mlp_blocks(BatchNearSrc{t, it.y, false, true}, count, mlp, t->delta, t->kr, t->alpha, 0, 1);
}

// Quadrature values: one block per item of 16 pairs.
__global__ __launch_bounds__(256) void k_near_vals_quad(const NearBatchTask *__restrict__ tasks,
                                                        const int2 *__restrict__ items) {
    const int2 it = items[blockIdx.x];
    const NearBatchTask *t = tasks + it.x;
    const long total = t->rowStart[t->n];
    const long count = std::min<long>(16, total - it.y);
    near_quad_block(BatchNearSrc{t, it.y, t->mode == 2, false}, count, t->delta, t->kr, t->alpha, 0);
}

// One thread per row: the row's correction, lower- and higher-facet emitters apart.
__global__ __launch_bounds__(256) void k_near_apply(const NearApplyTask *__restrict__ tasks,
                                                    const int2 *__restrict__ items, int nitems) {
    if ((int)blockIdx.x >= nitems) return;
    const int2 it = items[blockIdx.x];
    const NearApplyTask &t = tasks[it.x];
    const int r = it.y + threadIdx.x;
    if (r >= t.n) return;
    const int fr = __float_as_int(t.nrm[r].w);
    float2 L = make_float2(0.f, 0.f), U = make_float2(0.f, 0.f);
    for (int k = t.rowStart[r]; k < t.rowStart[r + 1]; k++) {
        const int e = t.cols[k];
        const float a = t.p[e].w;
        const float2 x = make_float2(a * t.pin[e].x, a * t.pin[e].y);
        const float2 v = t.vals[k];
        const float2 c = make_float2(v.x * x.x - v.y * x.y, v.x * x.y + v.y * x.x);
        if (__float_as_int(t.nrm[e].w) < fr) {
            L.x += c.x;
            L.y += c.y;
        } else {
            U.x += c.x;
            U.y += c.y;
        }
    }
    t.partL[r] = L;
    t.partU[r] = U;
}

int near_rows_multi_launch(const NearBatchTask *tasks, const int2 *items, int nitems, hipStream_t st) {
    if (nitems <= 0) return 0;
    k_near_rows<<<(unsigned)((nitems + 3) / 4), 256, 0, st>>>(tasks, items, nitems);
    SR_HIP(hipGetLastError());
    return 0;
}

int near_vals_multi_launch(const NearBatchTask *tasks, const float *mlp, const int2 *itemsMlp, int nMlp,
                           const int2 *itemsQuad, int nQuad, hipStream_t st) {
    if (nMlp > 0) {
        if (!mlp) { set_error("learned near field without a network"); return -2; }
        k_near_vals_mlp<<<(unsigned)nMlp, 256, 0, st>>>(tasks, mlp, itemsMlp);
        SR_HIP(hipGetLastError());
    }
    if (nQuad > 0) {
        k_near_vals_quad<<<(unsigned)nQuad, 256, 0, st>>>(tasks, itemsQuad);
        SR_HIP(hipGetLastError());
    }
    return 0;
}

int near_apply_multi_launch(const NearApplyTask *tasks, const int2 *items, int nitems, hipStream_t st) {
    if (nitems <= 0) return 0;
    k_near_apply<<<(unsigned)nitems, 256, 0, st>>>(tasks, items, nitems);
    SR_HIP(hipGetLastError());
    return 0;
}

}  // namespace sr
