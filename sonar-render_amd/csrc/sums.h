// Launch descriptors for the discrete Kirchhoff gathers (plain FP32 form).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sr {

// Physical constants of one projection, all FP32 (k_r is rounded to FP32 by
// the reference, src/Model.cpp:50-52).
struct Phys {
    float kr;         // Re k
    float krevHi;     // kr / (2 pi), split hi + lo so r * kr/(2 pi) keeps ~40 bits
    float krevLo;
    float alpha;      // Im k (the attenuation argument, used as given)
    float alphaLog2e; // alpha * log2(e)
    float deltaReg;   // fragment->fragment: r < deltaReg uses ik/(4 pi) (OptiX.cu:340-346)
    float damp;       // 0.9 for fragment->fragment (OptiX.cu:352), 1 otherwise
};

enum GatherKind {
    KIND_F2F = 0,       // fragment->fragment, both cosines (target destination)
    KIND_F2F_FIELD = 1, // fragment->fragment into a field object (cos_inc = 1)
    KIND_F2P = 2        // fragment->field point (compensated FP32 sums)
};

enum GateMode {
    GATE_NONE = 0,
    GATE_UPPER = 1, // SELF: terms from higher facets need the receiver's own P_in != 0 (OptiX.cu:241-245)
    GATE_ALL = 2    // RESP reciprocal: every term needs the receiver's own P_in != 0
};

struct GatherTask {
    // emitters (columns)
    const float4 *ep;       // position xyz, area w
    const float4 *en;       // normal xyz, facet id bits w
    const int *ecell;       // emitter -> index into eval (null: identity)
    const float2 *eval;     // emitter pressure, plane b at eval + b * evalStride
    long evalStride;
    int ne;
    // receivers (rows)
    const float4 *rp;
    const float4 *rn;
    int nr;
    const uint64_t *mask;   // [ceil(nr/64)][ne], bit = pair contributes
    // gating on the receiver's own input pressure
    const float2 *gval;
    long gvalStride;
    int gateMode;
    // batching
    int nrhs;               // planes handled by this launch
    int nsplit;             // emitter ranges (split-K)
    int emPerSplit;
    // outputs: partial sums [split][NB][nr]
    float2 *partL;          // lower / ungated terms (F2P: high part)
    float2 *partU;          // gated terms (F2P: compensation)
    // optional (profiling, matrix-core kernels): += the emitter-pair steps executed
    // (each 2 emitters x 32 receivers x the launch's planes)
    unsigned long long *workCount;
};

struct ReduceTask {
    const float2 *partL, *partU;
    int nsplit, nbStride, nrhs, nr;
    const float2 *gval;
    long gvalStride;
    int gateMode;
    float2 *out;            // optional: P_out planes (compact rows)
    long outStride;
    int outAdd;             // 1: out += value (RESP), 0: out = value
    float2 *pcell;          // optional: P += value, through rowCell
    long pcellStride;
    const int *rowCell;
    const float *colInvScale; // optional: value *= colInvScale[plane]
};

// Visibility of point->fragment pairs evaluated inside the sum (scenes of at
// most 4 occluder triangles): the sources->fragments test of the visibility
// kernels. tris null: read the visibility words instead.
struct P2FVis {
    const float4 *tris;  // the scene's occluder records (Bvh::tris), ntri x 4 float4
    int ntri;
    int rcTriBase;       // occluder index of the receiving object's facet 0 (-1: not an occluder)
    float scale, padTri, tmin;  // VisConst::scale, padTri, tminPoint
};

struct P2FTask {
    const float4 *cp;       // receiver (cell) positions
    const float4 *cn;       // cell normals
    int ncell;
    int useNormal;          // 0 for field objects (cos_inc = 1)
    const float4 *sp;       // sources
    const float2 *samp;     // source amplitudes
    int ns;
    const uint64_t *mask;   // [ceil(ncell/64)][ns]
    int batched;            // 1: source s writes plane s only
    float2 *P;
    long Pstride;
    P2FVis vis{};           // batched only: the visibility evaluated inline (mask unused)
};

int gather_launch(const GatherTask &t, const Phys &ph, int kind, hipStream_t st);
int reduce_launch(const ReduceTask &t, hipStream_t st);
// lo may be null (partials without a compensation term)
int reduce_points_launch(const float2 *hi, const float2 *lo, int nsplit, int nbStride, int nrhs, int nq,
                         double2 *fld, long fldStride, hipStream_t st);
int p2f_launch(const P2FTask &t, const Phys &ph, hipStream_t st);
int prime_launch(const float2 *P, long pStride, const int *actCell, int nact, int nrhs, float2 *Pout,
                 long outStride, hipStream_t st);
int any_zero_launch(const float2 *v, long n, int *flag, hipStream_t st);
// Fragment->fragment (kind KIND_F2F) or fragment->field point (KIND_F2P) for
// 1..64 planes with the complex products on the FP32 matrix cores. Partial
// sums [split][64][nr] (plane stride 64), FP32 within one emitter split; the
// gate is not applied (gateMode is ignored).
int f2f_mfma_launch(const GatherTask &t, const Phys &ph, int kind, hipStream_t st, bool compactSet = false);
// Gated rows (list of active indices) as a compact receiver set: positions,
// normals, first lower-facet bound and (when mask is given) visibility words
// [ceil(count/64)][ne] restricted to the lower-facet emitters.
int gated_compact_launch(const int *list, int count, const float4 *p, const float4 *nrm, const uint64_t *mask, int ne,
                         float4 *pc, float4 *nc, int *fsc, uint64_t *mc, hipStream_t st);
// Replaces the zero-input entries of the listed rows with the compact set's
// sums (partials [split][64][count], planes of one group of <= 64), each
// plane's sum times colScale[b] when colScale is given.
int gated_replace_launch(const float2 *part, int nsplit, int count, const int *list, const float2 *Pin, long stride,
                         int np, float2 *Pout, float2 *P, long pStride, const int *rowCell, const float *colScale,
                         hipStream_t st);

// ---- multi-task (many scenes in one launch) ----
struct PointsReduceTask {
    const float2 *hi, *lo;
    int nsplit, nbStride, nrhs, nq;
    double2 *fld;
    long fldStride;
};
struct PrimeTask {
    const float2 *P;
    long pStride;
    const int *actCell;
    int nact, nrhs;
    float2 *Pout;
    long outStride;
};
struct ZeroTask {
    float *p;   // float count n, zeroed with a kernel (graph capture records no memset nodes)
    long n;
};

int gather_multi_launch(const GatherTask *tasks, const Phys *phys, const int4 *items, int nitems, int *counter,
                        int kind, int nb, bool att, int grid, hipStream_t st);
// items: (task, first element) for 256-element blocks
int p2f_multi_launch(const P2FTask *tasks, const Phys *phys, const int2 *items, int nitems, bool att, hipStream_t st);
int reduce_multi_launch(const ReduceTask *tasks, const int2 *items, int nitems, hipStream_t st);
int reduce_points_multi_launch(const PointsReduceTask *tasks, const int2 *items, int nitems, hipStream_t st);
int prime_multi_launch(const PrimeTask *tasks, const int2 *items, int nitems, hipStream_t st);
int zero_multi_launch(const ZeroTask *tasks, const int2 *items, int nitems, hipStream_t st);
int counters_zero_launch(int *counters, int n, hipStream_t st);
// Zeroes count float2 ranges (start p[i], n[i] elements) with one launch per 8 ranges
// (a kernel rather than memset nodes: callers may capture it in a graph).
int zero_ranges_launch(float2 *const *p, const long *n, int count, hipStream_t st);

// Split-K choice: enough waves to occupy every SIMD a few times.
int choose_splits(int nrb, int ne, int minPerSplit);
// Plane stride of the partial-sum buffers for a launch of nrhs planes.
int gather_nb_stride(int nrhs);

}  // namespace sr
