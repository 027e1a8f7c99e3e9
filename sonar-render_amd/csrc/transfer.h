// Launch descriptors of the mixed-precision (FP16 MFMA) transfer form.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sr {

struct TransferGeom {
    const float4 *p;    // active cells: point, area
    const float4 *nrm;  // normal, facet id bits
    int n;              // active cells
    int nKs;            // K steps of 8 emitters (padded to a multiple of 8)
    int nRt;            // row tiles of 32 (padded to a multiple of 4)
    const int2 *wgSkip; // per 4-row-tile group: K steps [x, y) that hold only zeros (not written)
    // Rows of S may be padded so that every facet starts a 128-row group: rowAct
    // maps a row of S to its active cell (-1: padding), actRow the other way. The
    // emitters (K positions) are then padded the same way: K position k is the
    // emitter rowAct[k]. Both null: row = K position = active cell.
    const int *rowAct = nullptr;
    const int *actRow = nullptr;
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
    // near pairs found by the S build: by quadrature (nearQ) and learned (nearM)
    int2 *nearQ, *nearM;   // (row, emitter)
    int *nearCount;        // [2] device counters: zero before a build; the near-pair kernel clears them
    int nearCap;           // capacity of each list
    float quadRadius;      // learned mode: closer pairs use the quadrature (metres, >= delta)
    int streamS = 0;       // write S non-temporally (see kStreamSBytes)
};

// Split s of nsplit over the K steps of one 128-row group whose zero chunks are
// [skip.x, skip.y) (chunk-aligned K steps): the group's non-zero chunks of 8
// steps, numbered in K order, dealt out evenly over the splits, so that every
// (group, split) streams the same number of chunks whatever its facet. The split
// walks the chunk numbers [o0, o1); chunk number o starts at K step k(o).
struct SplitWalk {
    int below, skipEnd, o0, o1;
    __host__ __device__ int k(int o) const { return o < below ? o * 8 : (o - below) * 8 + skipEnd; }
};
__host__ __device__ inline SplitWalk split_walk(int2 skip, int nKs, int nsplit, int s) {
    const int below = skip.x / 8, total = nKs / 8 - (skip.y - skip.x) / 8;
    return SplitWalk{below, skip.y, (int)((long)s * total / nsplit), (int)((long)(s + 1) * total / nsplit)};
}

// Reduction of a gated product: per (plane, active row), the lower-facet
// partials always and the higher-facet ones where the row's own input is not
// zero (src/OptiX.cu:241-245); cls[group][split] bit 0/1 says which of the
// split's partials exist. Applies the column scale, writes P_out and adds to P.
// Words between consecutive planes' column maxima in a bounce state set: each
// plane's atomics hit a cache line of their own.
constexpr int kColStride = 32;

struct ReduceGateTask {
    const float2 *partL, *partU;
    const unsigned char *cls;
    int nsplit, nPadRows, nrhs, n;
    const int *rowAct;          // row of S -> active row (-1: padding; null: identity, nPadRows == n)
    const float2 *gval;         // P_in [plane][n]
    float2 *out;                // P_out [plane][n]
    float2 *pcell;              // P [plane][pcellStride] through rowCell
    long pcellStride;
    const int *rowCell;
    const float *colInvScale;
    // when set: per plane, max |a_r P_out[b][r]| (a_r = actP[r].w) as float bits by
    // atomicMax, zero beforehand: the next bounce's B scales without a pass over P_in
    const float4 *actP;
    unsigned *colMaxNext;
};

struct TransferGemm {
    const uint4 *A, *B;
    int n, nKs, nRt;
    int nplanes;        // planes of this group (<= 64)
    int nsplit;         // splits of each group's non-zero chunks (split_walk)
    const int2 *wgSkip;                   // per 4-row-tile group: K steps [x, y) that hold only zeros
    float2 *partL;                        // [split][64][n]
    // Zero-input gate inside the product: per 128-row group, the first K step of
    // the higher-facet emitters (whole chunks). A workgroup writes what it sums
    // below that step to partL and from it on to partU (null: partL only).
    const int *kB = nullptr;
    float2 *partU = nullptr;
    // Zero-input gate by two accumulators (any facet layout): per row of S, the
    // first emitter of its own facet; the product sends the terms below it to
    // partL and the rest to partU, both for every split (null: one accumulator).
    const int *rowFs = nullptr;
    // S larger than the GPU's memory-side cache: stream it (non-temporal loads)
    // rather than let each bounce's pass evict the cached B and partial sums
    // for data no later access in the pass reuses.
    int streamS = 0;
    // The reduction inside the product (grpCount non-null; off by default, as its
    // device-scope fences cost more than the separate pass): each workgroup counts
    // itself done on its group's counter (zero before the launch; the last one
    // resets it) and the last of a group's splits reduces the group's rows as
    // k_reduce_gate would (red: the same task), the splits summed in split order.
    ReduceGateTask red{};
    unsigned *grpCount = nullptr;
};

// S of at least this many bytes is written and read non-temporally; a smaller
// one stays in the memory-side cache (256 MiB on MI355X) between the passes.
constexpr double kStreamSBytes = 160.0 * 1024 * 1024;

// The visibility test of an object's own fragment pairs, for a scene of at most
// 4 occluder triangles, evaluated inside the symmetric S build instead of read
// from the SELF visibility words (the same bits: the visibility kernels' test on
// the same FP32 arithmetic).
struct SelfVis {
    const float4 *tris;  // the scene's occluder records (Bvh::tris), ntri x 4 float4
    int ntri;            // 1 .. 4
    int triBase;         // occluder index of this object's facet 0
    float scale, padTri, tmin;  // VisConst::scale, padTri, tminFrag
};

int transfer_build_s(const TransferGeom &g, const TransferPhys &ph, const uint64_t *mask, uint4 *A, hipStream_t st);
// The same S from the blocks on and above the diagonal (items: (row block, K chunk)
// of 64), each stored with its transpose; rows and emitters padded alike, near mode 0.
// sv: evaluate the visibility here (mask unused); null: read it from mask.
int transfer_build_s_sym(const TransferGeom &g, const TransferPhys &ph, const uint64_t *mask, uint4 *A,
                         const int2 *items, int nItems, hipStream_t st, const SelfVis *sv = nullptr);
// Writes the listed near pairs' values into S (counts read back from nearCount)
// and clears nearCount for the next build.
int transfer_near(const TransferGeom &g, const TransferPhys &ph, int nQuad, int nMlp, uint4 *A, hipStream_t st);
// Per-plane max |A_j P_in[b][j]| into colMax[b kColStride] (float bits, zeroed beforehand);
// also zeroes nextWords ints at nextState (the next bounce's state set).
int transfer_col_max(const float4 *p, const float2 *Pin, long stride, int n, int nplanes, unsigned *colMax,
                     int *nextState, int nextWords, hipStream_t st);
// B of one plane group, scaled from colMax; stores each plane's scale and the
// inverse 1 / (scale sS) the reduction applies. clearState (nullable): clearWords
// ints zeroed (the next bounce's state set, when no column-max pass runs).
int transfer_build_b(const float4 *p, const float2 *Pin, long stride, int n, int nKs, int nplanes,
                     const unsigned *colMax, float sS, float *scale, float *invScale, uint4 *B, hipStream_t st,
                     int *clearState = nullptr, int clearWords = 0, const int *kAct = nullptr);
int transfer_gemm(const TransferGemm &t, hipStream_t st);
int transfer_reduce_gate(const ReduceGateTask &t, hipStream_t st);
// Lists the rows with an exactly zero input pressure in some plane (count: device int,
// zeroed by the caller): the plain form's gated rows.
int transfer_gated_rows(const float2 *Pin, int n, int nplanes, int *list, int *count, hipStream_t st);

}  // namespace sr
