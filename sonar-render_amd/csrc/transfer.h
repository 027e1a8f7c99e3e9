// Launch descriptors of the mixed-precision (FP16 MFMA) transfer form.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sr {

struct TransferGeom {
    const float4 *p;    // active cells: point, area
    const float4 *nrm;  // normal, facet id bits
    int n;              // active cells
    int nKs;            // K steps of 8 emitters (ceil(n/8), padded to a multiple of 8)
    int nRt;            // row tiles of 32 (padded to a multiple of 4)
};

struct TransferPhys {
    float kr, alpha, alphaLog2e, deltaReg, damp;
    float scale;        // FP16 scale of S
    double krevD;       // kr / (2 pi) in FP64
    // near field: 0 point rule (the reference), 1 16x16 quadrature, 2 learned MLP
    int nearMode;
    float nearRadius;   // metres
    const float4 *frames;  // per facet: xAxis (w unused), yAxis
    const float *mlp;      // kMlpParams FP32 weights (rounded to BF16 on load)
};

struct TransferGemm {
    const uint4 *A, *B;
    int n, nKs, nRt;
    int nplanes;        // planes of this group (<= 64)
    int nsplit, ksPerSplit;
    const int2 *wgSkip;                   // per 4-row-tile group: K steps [x, y) that hold only zeros
    float2 *partL;                        // [split][64][n]
};

int transfer_build_s(const TransferGeom &g, const TransferPhys &ph, const uint64_t *mask, uint4 *A, hipStream_t st);
int transfer_col_scale(const float4 *p, const float2 *Pin, long stride, int n, int nplanes, float *scale,
                       float *invScale, float sS, hipStream_t st);
int transfer_build_b(const float4 *p, const float2 *Pin, long stride, int n, int nKs, int nplanes, const float *scale,
                     uint4 *B, hipStream_t st);
int transfer_gemm(const TransferGemm &t, hipStream_t st);
// Rows of one object's SELF bounce whose own input pressure is exactly zero in
// some plane, recomputed with the lower-facet emitters only after a product
// that summed every emitter. Active cells ascend by facet.
struct GatedFix {
    const int *list;        // gated rows (active indices)
    int count;
    int n;
    const float4 *p, *nrm;  // active points (area in w), normals (facet id bits in w, ascending)
    const float2 *Pin;      // [plane][stride]
    long stride;
    int nplanes;
    const uint4 *A;         // FP16 transfer matrix (pair values read back), or null: evaluate in FP32
    int nKs;
    const uint64_t *mask;   // SELF visibility words (FP32 evaluation only)
    float valueScale;       // 1 / (S scale) for A, 1 otherwise
    int emPerSplit, nsplit; // emitter ranges, emPerSplit a multiple of 8
    float2 *part;           // [split][nplanes][count]
    float2 *Pout;           // [plane][stride]
    float2 *P;              // [plane][pStride] through rowCell
    long pStride;
    const int *rowCell;
};
int gated_fix_launch(const GatedFix &t, const Phys &ph, hipStream_t st);
// Lists the rows with an exactly zero input pressure in some plane (count: device int, zeroed by the caller).
int transfer_gated_rows(const float2 *Pin, int n, int nplanes, int *list, int *count, hipStream_t st);

}  // namespace sr
