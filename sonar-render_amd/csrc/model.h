// Scene model and projection sequencing on one GPU.
//
// Mirrors the reference's Model/ModelCuda/ObjectCuda/FacetCuda state machine
// (src/Model.cpp, src/ModelCuda.cpp:242-356, src/FacetCuda.cpp:8-137) as plain
// data: per object, one SoA fragment table in HBM instead of per-facet device
// allocations, and one visibility bit-mask per (emitter set, receiver set).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "fragmenter.h"
#include "learned.h"
#include "transfer.h"
#include "sr_types.h"
#include "sums.h"

namespace sr {

struct Object {
    int type = OBJ_TARGET;
    float delta = 0.f;
    std::vector<FacetFrame> frames;
    std::vector<FacetGrid> grids;
    std::vector<long> cellOff;
    long ncell = 0;
    int nact = 0;
    std::vector<int> actCellHost;
    // device tables
    float4 *cellP = nullptr;   // point->fragment receiver point (+delta/2), area
    float4 *cellN = nullptr;   // facet normal, facet id bits
    float4 *actP = nullptr;    // active cell point (emitter form), area
    float4 *actN = nullptr;    // facet normal, facet id bits
    int *actCell = nullptr;    // active -> cell
    float4 *frameTab = nullptr;  // per facet: xAxis, yAxis
    double sameFacetPairs = 0; // sum over facets of (active cells)^2
    int triBase = -1;          // targets: index of the first facet among the occluder triangles
    std::vector<int> actFacetHost;
    // mixed-precision transfer matrix (built once per geometry and frequency)
    struct Transfer {
        uint4 *A = nullptr, *B = nullptr;
        int nKs = 0, nRt = 0;
        int2 *wgSkip = nullptr;    // per 128-row group: zero K-step range (its own single facet)
        // bounce state, two sets used in turn: [0] spare, [1 + b kColStride] max |A P_in| of plane b
        // (float bits); each bounce clears the other set for the next
        int *state = nullptr;
        int stateSet = 0;
        bool stateClear = false;   // the set stateSet names is known to be zero
        bool maxReady = false;     // ... or holds the column maxima of P_in (from the last reduction)
        int *rowFs = nullptr;      // per active row: first active index on its facet (the two-accumulator gate)
        // Zero-input gate inside the product (meshes whose facets each hold >= 128
        // active cells): rows of S padded so that every facet starts a 128-row
        // group; per group the first K step of the higher facets (kB); per
        // (group, split) which partial sums the split writes (cls: bit 0 lower, 1 higher)
        bool gateInGemm = false;
        int nPad = 0;              // rows of S (padded)
        int *rowAct = nullptr, *actRow = nullptr, *kB = nullptr;
        unsigned char *cls = nullptr;
        unsigned *grpCount = nullptr;  // per 128-row group: the product's workgroups done (zero between launches)
        int nsplit = 1;             // split-K of the bounce product (split_walk)
        int splitsReq = 0;          // the split request (Model::transferSplits) nsplit was chosen for
        // symmetric S build (padded rows): the 64 x 64 blocks (row block, K chunk) on
        // or above the diagonal that are not skipped; none when the skipped chunks
        // are not symmetric
        double zeroPairs = 0;      // ordered pairs within one run of facets with identical normals (zero in S)
        int2 *symItems = nullptr;
        int nSym = 0;
        int2 *nearQ = nullptr, *nearM = nullptr;  // near pairs: quadrature, learned
        int *nearCount = nullptr;
        bool nearClear = false;    // nearCount is known to be zero (the near-pair kernels clear it)
        int nearCap = 0;
        // near-pair counts of the last build and the settings they hold for: the
        // counts depend only on geometry, visibility and these, so a rebuild with
        // the same settings needs no read-back
        int nearKnown[2] = {0, 0};
        double nearKey[5] = {-1, 0, 0, 0, 0};
        unsigned nfGen = 0;     // learned mode: the network generation the near terms come from
        float *colScale = nullptr, *colInv = nullptr;
        float sS = 1.f;
        double aBytesRead = 0;  // bytes of S one GEMM pass reads (same-facet steps are skipped)
        double partRows = 0;    // partial-sum rows one GEMM pass writes, per plane
        bool valid = false;
    } tc;
    // pressure planes (nplanes each)
    float2 *P = nullptr;       // per cell
    float2 *Pin = nullptr;     // per active cell (targets)
    float2 *Pout = nullptr;    // per active cell (targets)
};

struct Model {
    // initial conditions (Model.cpp:44-54)
    float cp = 1480.f, freq = 0.f, atten = 0.f, density = 0.f;
    float kf = 0.f;
    std::vector<Object> targets, fields;
    // source objects (type 1): piston-like emitters, every cell at one coherent
    // amplitude (README.md:2; stored but never used by the reference, Model.cpp:33-35)
    std::vector<Object> sources;
    std::vector<float2> srcObjAmp;
    // surface colour settings of the render (SoundVisualisationInit; ModelCuda.hpp:45-48 defaults)
    float renderMaxDb = -20.f, renderMinDb = -90.f;
    bool renderPhaseTarget = false, renderPhaseField = false;
    std::vector<float> srcHost, fldHost;  // xyz triples
    bool rendered = false;
    bool batchedSources = false;
    int nplanes = 1;

    // device
    hipStream_t stream = nullptr;
    // a second stream (default priority) for the gated rows' compact set, which
    // overlaps the bounce's main product; two events order it against the main stream
    hipStream_t side = nullptr;
    hipEvent_t sideIn = nullptr, sideOut = nullptr;
    float2 *scratchC = nullptr;  // the compact set's partial sums
    size_t scratchCElems = 0;
    // counts read back while the stream runs on: pinned words, one event each
    static constexpr int kWords = 4;
    int *hostWords = nullptr;  // [0] gated-row count
    hipEvent_t wordEv[kWords] = {};
    // the visibility traversal's overflow flag: a mapped, coherent pinned word that
    // the kernels store into, read by the host after a stream wait with no copy
    int *errHost = nullptr, *errFlag = nullptr;  // host view, device view
    bool errPending = false;  // projections since the last result read may have raised it
    float4 *srcP = nullptr;
    float2 *srcAmp = nullptr;
    float4 *fldP = nullptr;
    double2 *fldVal = nullptr;  // nplanes x Q
    float4 *trisSorted = nullptr;
    float4 *nodes = nullptr;
    int ntri = 0;
    VisConst vc{};
    // lowerPending: a symmetric set whose words below the diagonal blocks are not yet
    // filled in by transposition (only the upper ones were needed so far)
    struct MaskBuf { uint64_t *p = nullptr; size_t words = 0; bool valid = false; bool lowerPending = false; };
    std::map<std::tuple<int, int, int, int, int>, MaskBuf> masks;
    int transferMode = 0;  // 0 plain FP32 gathers, 1 FP16 MFMA transfer matrix
    int transferSplits = 0;  // splits of the transfer product: 0 automatic
    // S build: 0 every block; 1 half the blocks and their transposes where possible;
    // 2 as 1, and with at most 4 occluder triangles the SELF visibility evaluated in that build
    int transferSymmetric = 2;
    // 1: with at most 4 occluder triangles the point->fragment sum (one plane per
    // source) evaluates its visibility itself; 0: always from the visibility words
    int inlineVis = 1;
    // 1: the transfer product's last workgroup of each row group reduces its split-K
    // partials (gate, P_out, P, next column maxima); 0: a pass of its own after it.
    // 0 by default: the agent-scope release and acquire the handover needs write back
    // and invalidate an XCD's L2 each (the cached B with it), which costs more than
    // the pass (measured 0.266 against 0.133 + 0.015 ms per bounce at lambda/7.5)
    int reduceInGemm = 0;
    int gatherMode = 0;    // plain fragment->fragment: 0 FP32 MFMA products from 16 planes on, 1 VALU only
    int *gatedList = nullptr, *gatedCount = nullptr;  // plain form: rows a zero input gates
    int gatedCap = 0;
    float4 *gPc = nullptr, *gNc = nullptr;  // ... as a compact receiver set
    int *gFs = nullptr;
    uint64_t *gMc = nullptr;
    int gCap = 0;
    size_t gMcCap = 0;
    int nearMode = 0;      // transfer matrix near field: 0 point rule, 1 16x16 quadrature, 2 learned
    float nearRadiusWl = 2.0f;
    float nearQuadDeltas = kNfRMinDeltas;  // learned mode: pairs closer than this many fragment sizes use the quadrature
    float2 *scratchL = nullptr, *scratchU = nullptr;
    size_t scratchElems = 0;
    // Allocation-free calls: partial sums go to a fixed buffer the caller's entry
    // point provides (two halves of fixElems each), and nothing is allocated.
    bool noAlloc = false;
    float2 *fixL = nullptr, *fixU = nullptr;
    size_t fixElems = 0;
    // the last transfer-form bounce product, for sr_time_transfer_gemm
    TransferGemm lastGemm{};
    bool haveLastGemm = false;
    // per-launch profiling (HIP events on the library's stream)
    bool profiling = false;
    // workSlot >= 0: flops = (executed steps counted in workCnt[workSlot]) x flopsPerStep
    struct Rec { int kind; hipEvent_t a, b; double pairs, flops, bytes; int workSlot; double flopsPerStep; };
    std::vector<Rec> recs;
    unsigned long long *workCnt = nullptr;  // device counters of executed matrix-core work (profiling)
    int workCap = 0, workUsed = 0;
};

// Kernel families reported by the profiler.
enum ProfKind {
    PK_P2F = 0, PK_VIS = 1, PK_F2F = 2, PK_F2F_FIELD = 3, PK_F2P = 4, PK_REDUCE = 5,
    PK_BUILD_S = 6, PK_GEMM = 7, PK_BUILD_B = 8, PK_GATED_FIX = 9, PK_NEAR = 10, PK_COUNT = 11
};

}  // namespace sr
