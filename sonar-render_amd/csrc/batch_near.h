// Near-field correction of the batched scenes' bounces: per scene, the visible
// fragment pairs closer than the near radius in CSR order (rows by receiver,
// emitters ascending), their value by the quadrature or the learned
// correction minus the point rule the gathers already add, and a per-row
// apply that writes the correction as one more split-K partial (lower- and
// higher-facet parts apart, so the zero-input gate acts on it as on the rest).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sr {

struct NearBatchTask {
    const float4 *p, *nrm;   // active cells: point (area in w), normal (facet id bits in w)
    const float4 *frames;    // per facet: xAxis, yAxis
    const uint64_t *mask;    // SELF visibility words [ceil(n/64)][n]
    int n;
    float radius2;           // pairs with r^2 (near_r2) below this are near
    float quadR2;            // learned mode: closer pairs use the quadrature (r^2)
    float kr, alpha, delta;
    int mode;                // 1 quadrature for every near pair, 2 learned (quadrature below quadR2)
    const int *rowStart;     // n + 1 CSR offsets (null: the counting pass)
    int *rowCount;           // counting pass: near pairs of each row
    int *rows, *cols;        // CSR pairs (receiver, emitter)
    float2 *vals;            // value minus the point rule, per unit emitter area
};

struct NearApplyTask {
    const int *rowStart, *cols;
    const float2 *vals;
    const float4 *p, *nrm;   // emitter areas, facet ids
    const float2 *pin;       // this bounce's input pressures
    float2 *partL, *partU;   // the extra split's partial sums (lower-, higher-facet emitters)
    int n;
};

// items: (task, 64-row block). Counting pass when tasks' rowStart is null.
int near_rows_multi_launch(const NearBatchTask *tasks, const int2 *items, int nitems, hipStream_t st);
// items: (task, first pair), 128 pairs per item (learned) / 16 (quadrature).
int near_vals_multi_launch(const NearBatchTask *tasks, const float *mlp, const int2 *itemsMlp, int nMlp,
                           const int2 *itemsQuad, int nQuad, hipStream_t st);
// items: (task, first row), 256 rows per item.
int near_apply_multi_launch(const NearApplyTask *tasks, const int2 *items, int nitems, hipStream_t st);

}  // namespace sr
