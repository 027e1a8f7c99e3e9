// Mixed-precision transfer form of the fragment->fragment bounce for many
// incidences sharing one geometry.
//
// The reference applies the same pair kernel K_ij = cos_s cos_i 0.9 G(r_ij)
// (src/OptiX.cu:297-354) to every incidence separately. Here K is built once
// per geometry and frequency as a complex FP16 matrix S (real and imaginary
// parts interleaved along K, scaled by one factor so |S| stays below 2^14),
// and every bounce is one real GEMM on the matrix cores:
//     C[i, 0..63]   = sum_j Re S_ij X_re[j,b] - Im S_ij X_im[j,b]
//     C[i, 64..127] = sum_j Im S_ij X_re[j,b] + Re S_ij X_im[j,b]
// with B = [X_re; X_im] (2N x 64, real and imaginary parts interleaved along K
// as in S) and two A operands formed from each S fragment in registers: the
// words (Re S, -Im S) for the real part and (Im S, Re S) for the imaginary
// part, X = A_j P_in[j,b] scaled per incidence, FP32 accumulation
// (v_mfma_f32_32x32x16_f16). The reference's zero-input gate
// (OptiX.cu:241-245: terms from higher facets need P_in[i] != 0) is kept exact
// by accumulating lower-facet and higher-facet emitters separately.
//
// Layouts (both written by this file, read in fragment order so every load is
// one coalesced 1 KiB per wave):
//   A: [row tile of 32][K step of 8 emitters][lane][8 halves]
//      lane l: row = tile*32 + l%32, emitters step*8 + (l/32)*4 + q, halves (Re, Im)
//   B: [K step][column tile of 32 (2)][lane][8 halves], same K assignment, plane
//      tile*32 + l%32.
// The phase k r is reduced modulo one revolution in FP64 before the FP32
// sin/cos (the single FP64 step of the form).

#include <hip/hip_runtime.h>
#include <cstdint>
#include <algorithm>

#include "launch.h"
#include "sums.h"
#include "pair_kernel.h"
#include "transfer.h"
#include "transfer_dev.h"
#include "learned.h"
#include "near_rules.h"
#include "near_eval.h"

namespace sr {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// Near-field values (per unit emitter area) for a pair closer than nearRadius:
// the evaluators of near_eval.h, read from the (row, emitter) lists the S build
// fills and written into the tiled FP16 matrix.

// Geometry of one near pair in the emitter's frame (x = xAxis, y = yAxis,
// z = emitter normal): receiver point relative to the emitter point, and the
// receiver normal.
__device__ inline NearGeom near_geom(const TransferGeom &g, const TransferPhys &ph, int row, int j) {
    const float4 rp = g.p[row], rn = g.nrm[row], ep = g.p[j], en = g.nrm[j];
    const float vx = rp.x - ep.x, vy = rp.y - ep.y, vz = rp.z - ep.z;
    const int fe = __float_as_int(en.w);
    const float4 xa = ph.frames[2 * fe], ya = ph.frames[2 * fe + 1];
    NearGeom q;
    q.rv = make_float3(vx * xa.x + vy * xa.y + vz * xa.z, vx * ya.x + vy * ya.y + vz * ya.z,
                       vx * en.x + vy * en.y + vz * en.z);
    q.nr = make_float3(rn.x * xa.x + rn.y * xa.y + rn.z * xa.z, rn.x * ya.x + rn.y * ya.y + rn.z * ya.z,
                       rn.x * en.x + rn.y * en.y + rn.z * en.z);
    q.r = sqrtf(q.rv.x * q.rv.x + q.rv.y * q.rv.y + q.rv.z * q.rv.z);
    return q;
}

// The FP16 word of the pair (row, K position j) in the tiled transfer matrix.
__device__ inline void store_pair(uint4 *A, int nKs, int row, int j, uint32_t w) {
    const size_t lane = (size_t)(row & 31) + 32 * ((j >> 2) & 1);
    uint32_t *q = reinterpret_cast<uint32_t *>(A + ((size_t)(row >> 5) * nKs + (j >> 3)) * 64 + lane);
    q[j & 3] = w;
}


// The near lists' pairs as a source for the evaluators: each value goes to its
// FP16 word of S (scaled as S is).
struct TransferNearSrc {
    TransferGeom g;
    TransferPhys ph;
    const int2 *list;
    uint4 *A;
    __device__ NearGeom geom(long i) const {
        const int2 pr = list[i];
        return near_geom(g, ph, pr.x, pr.y);
    }
    __device__ bool want(long) const { return true; }
    __device__ void put(long i, float re, float im) const {
        const int2 pr = list[i];
        store_pair(A, g.nKs, g.actRow ? g.actRow[pr.x] : pr.x, g.actRow ? g.actRow[pr.y] : pr.y,
                   pack_h2(re * ph.scale, im * ph.scale));
    }
};

// Both near-pair lists in one launch: blocks [0, quadBlocks) take the quadrature
// list (16 lanes per pair), the rest the learned list (128 pairs per block and
// round). Block 0 clears the near counters for the next build (nothing here reads them).
__global__ __launch_bounds__(256) void k_near(TransferGeom g, TransferPhys ph, int nQuad, int nMlp, int quadBlocks,
                                              uint4 *__restrict__ A) {
    if (blockIdx.x == 0 && threadIdx.x < 2) ph.nearCount[threadIdx.x] = 0;
    if ((int)blockIdx.x < quadBlocks) {
        near_quad_block(TransferNearSrc{g, ph, ph.nearQ, A}, nQuad, ph.deltaReg, ph.kr, ph.alpha, blockIdx.x);
    } else {
        near_mlp_blocks(TransferNearSrc{g, ph, ph.nearM, A}, nMlp, ph.mlp, ph.deltaReg, ph.kr, ph.alpha,
                        blockIdx.x - quadBlocks, gridDim.x - quadBlocks);
    }
}

// Per-incidence scale: 2^13 / max(|Re X|, |Im X|) with X = A_j P_in[j,b].
// Per-plane max |A_j P_in[b][j]| over the active cells (the FP16 scaling of
// B), atomically into colMax[b * kColStride] as float bits (non-negative floats order like
// their bit patterns). Grid (planes, row chunks). Block (0, 0) also clears the
// other bounce-state set, which the next bounce uses.
__global__ __launch_bounds__(256) void k_col_max(const float4 *__restrict__ p, const float2 *__restrict__ Pin,
                                                 long stride, int n, unsigned *__restrict__ colMax,
                                                 int *__restrict__ nextState, int nextWords) {
    __shared__ float red[4];
    const int b = blockIdx.x;
    if (b == 0 && blockIdx.y == 0)
        for (int k = threadIdx.x; k < nextWords; k += 256) nextState[k] = 0;
    const int per = (n + gridDim.y - 1) / gridDim.y;
    const int j0 = blockIdx.y * per, j1 = min(n, j0 + per);
    float m = 0.f;
#pragma unroll 4
    for (int j = j0 + threadIdx.x; j < j1; j += 256) {
        const float2 v = Pin[b * stride + j];
        const float a = p[j].w;
        m = fmaxf(m, fmaxf(fabsf(a * v.x), fabsf(a * v.y)));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(&colMax[b * kColStride], __float_as_uint(fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]))));
}

// The FP16 scale of plane b's B column: its largest entry becomes 8192.
__device__ inline float col_scale_of(unsigned bits) {
    const float mx = __uint_as_float(bits);
    return mx > 0.f ? 8192.0f / mx : 1.0f;
}

// B build for one group of up to 64 incidences.
// Scales come from colMax (k_col_max); threads 0..nplanes-1 also store each
// plane's scale and the inverse the reduction applies (1 / (scale sS)). A
// workgroup builds one quarter of a chunk of 8 K steps: one column tile (32
// planes) by one half (4 K steps, 32 K positions). It reads those planes' inputs
// along K (coalesced runs of 32 K positions) into LDS, then writes the quarter's
// B fragments (one plane per lane) from there: four workgroups per chunk, so
// that even a small B fills the GPU.
constexpr int kBStride = 33;  // LDS row stride (float2) of one plane's 32 K positions, padded
__global__ __launch_bounds__(256) void k_build_b(const float4 *__restrict__ p, const float2 *__restrict__ Pin,
                                                 long stride, int n, int nKs, int nplanes,
                                                 const unsigned *__restrict__ colMax, float sS,
                                                 float *__restrict__ scaleOut, float *__restrict__ invOut,
                                                 uint4 *__restrict__ B, int *__restrict__ clearState,
                                                 int clearWords, const int *__restrict__ kAct) {
    __shared__ float2 sX[32 * kBStride];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (clearState && t < clearWords) clearState[t] = 0;
    if (t < nplanes) {
        const float sc = col_scale_of(colMax[t * kColStride]);
        scaleOut[t] = sc;
        invOut[t] = 1.0f / (sc * sS);
    }
    const int unit = blockIdx.x;  // (chunk, column tile, half)
    const int chunk = unit >> 2, tile = (unit >> 1) & 1, half = unit & 1;
    if (chunk >= nKs / 8) return;  // uniform over the workgroup
    const int kk = threadIdx.x & 31;
    const int k = chunk * 64 + half * 32 + kk;
    const int j = kAct ? kAct[k] : (k < n ? k : -1);  // its active cell (-1: padding)
    const float area = j >= 0 ? p[j].w : 0.f;
    // all 4 of this thread's loads in flight at once
    float2 v[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int b = tile * 32 + (threadIdx.x >> 5) + 8 * i;
        v[i] = (j >= 0 && b < nplanes) ? Pin[b * stride + j] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int bl = (threadIdx.x >> 5) + 8 * i, b = tile * 32 + bl;
        float xr = 0.f, xi = 0.f;
        if (j >= 0 && b < nplanes) {
            const float a = area * col_scale_of(colMax[b * kColStride]);
            xr = a * v[i].x;
            xi = a * v[i].y;
        }
        sX[bl * kBStride + kk] = make_float2(xr, xi);
    }
    __syncthreads();
    // fragment: K step chunk*8 + half*4 + s, column tile `tile`, lane (plane tile*32 + lane%32)
    const int s = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float2 *x = sX + (lane & 31) * kBStride + s * 8 + (lane >> 5) * 4;
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; q++) w[q] = pack_h2(x[q].x, x[q].y);
    B[((size_t)chunk * 8 + half * 4 + s) * 128 + tile * 64 + lane] = make_uint4(w[0], w[1], w[2], w[3]);
}

// ---------------------------------------------------------------------------
// GEMM: a workgroup = 4 waves = 128 rows x 128 columns; a wave = 32 rows.
// K advances in chunks of 8 steps (8 emitters each). B chunks (16 KiB) are
// copied global -> LDS without passing through registers and are shared by
// the 4 waves; A fragments stream from HBM straight into registers. Both are
// requested one chunk ahead of the MFMAs that use them (two LDS buffers, two
// A register sets). Chunks whose emitters all lie on the workgroup's single
// facet (or its run of facets with one normal) hold only zeros and are skipped;
// the splits (blockIdx.y) share out each group's remaining chunks evenly.
// The zero-input gate needs each row's lower-facet terms (emitters below its
// own facet) apart from the rest, in one of two ways:
//   kB (rows of S padded so that each 128-row group is single-facet): the
//     group's chunks below kB are lower-facet, so one accumulator is stored to
//     partL when the walk passes kB and to partU at the end;
//   DUAL (any facet layout): two accumulators, A split per row at its own
//     facet start (rowFs), each stored to partL / partU; chunks wholly below
//     or above every row of the wave take one product, the others two (the
//     product is HBM-bound: the matrix cores have room for the second).
constexpr int kKC = 8;
static_assert(kKC == 8, "K steps and split sizes are padded to multiples of 8");

// Workgroup barrier without __syncthreads' memory fence, whose wait for every
// outstanding global load would also hold the A chunk requested ahead. What the
// barrier orders here needs no fence: each wave waits (s_waitcnt) for its own
// B loads into LDS before it, and its LDS reads of the other buffer have been
// consumed by issued MFMAs. The empty asm keeps the compiler from moving memory
// accesses across it.
__device__ inline void workgroup_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

__device__ void reduce_group_last(const TransferGemm &t, int grp);

// ROLE 0: one accumulator (the kB gate, or none); 2: DUAL. NT: S read non-temporally.
template <int ROLE, bool NT>
__global__ __launch_bounds__(256, ROLE == 2 ? 1 : 2) void k_transfer_gemm(TransferGemm t) {
    constexpr bool DUAL = ROLE == 2;
    // two separately named buffers, so that the compiler can tell the LDS reads of
    // one from the asynchronous global -> LDS copies into the other
    __shared__ uint4 sB0[kKC * 2 * 64], sB1[kKC * 2 * 64];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    // Workgroup L = group * nsplit + split. Workgroups are dealt round-robin to the
    // 8 XCDs, so with nsplit dividing 8 every XCD runs one split of every group it
    // holds: its L2 caches that split's part of B rather than all of it.
    const int grp = blockIdx.x / t.nsplit;
    const int split = blockIdx.x - grp * t.nsplit;
    const int rt = grp * 4 + wave;
    const bool rtValid = rt < t.nRt;
    // this split's run of the group's non-zero chunks (chunk-aligned K steps
    // [skip.x, skip.y) hold only zeros and are never read)
    const SplitWalk sw = split_walk(t.wgSkip[grp], t.nKs, t.nsplit, split);
    const int nch = sw.o1 - sw.o0;
    const uint4 *Arow = t.A + (size_t)(rtValid ? rt : 0) * t.nKs * 64 + lane;
    const uint4 *Bt = t.B + threadIdx.x;     // thread x stages elements x + 256 r of a chunk
    // in-product gate: K steps from kBnd on are the group's higher-facet emitters
    const int kBnd = (!DUAL && t.kB) ? t.kB[grp] : 0x7fffffff;
    int kLast = -1;  // the last chunk summed
    // DUAL: this lane's row's facet start, and the wave's least and greatest
    int fsRow = 0, fsMin = 0, fsMax = 0;
    if (DUAL) {
        const int row = rt * 32 + (lane & 31);
        const bool rv = rtValid && row < t.n;
        fsRow = rv ? t.rowFs[row] : 0;
        int lo = rv ? fsRow : 0x7fffffff, hi = rv ? fsRow : 0;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            lo = min(lo, __shfl_xor(lo, o));
            hi = max(hi, __shfl_xor(hi, o));
        }
        fsMin = __builtin_amdgcn_readfirstlane(lo);
        fsMax = __builtin_amdgcn_readfirstlane(hi);
    }

    f32x16 acc[4], accU[DUAL ? 4 : 1];
#pragma unroll
    for (int c = 0; c < 4; c++) acc[c] = f32x16{};
#pragma unroll
    for (int c = 0; c < (DUAL ? 4 : 1); c++) accU[c] = f32x16{};
    // C layout: element i of lane l is (row (i/4)*8 + (l/32)*4 + i%4, column l%32).
    auto store_part = [&](float2 *part, const f32x16 *ac) {
        if (!rtValid) return;
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const int b = c * 32 + (lane & 31);
            if (b >= t.nplanes) continue;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int row = rt * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
                if (row >= t.n) continue;
                part[((size_t)split * 64 + b) * t.n + row] = make_float2(ac[c][i], ac[c + 2][i]);
            }
        }
    };
    // one K step's four products into an accumulator set: the S words (Re, Im)
    // as (Re, -Im) into the real parts (ac[0..1]) and as (Im, Re) into the
    // imaginary parts (ac[2..3]), each against the plane tile's B fragment
    auto mfma_step = [&](f32x16 *ac, uint4 a, const uint4 *sb, int s) {
        const uint4 ar = make_uint4(a.x ^ 0x80000000u, a.y ^ 0x80000000u, a.z ^ 0x80000000u, a.w ^ 0x80000000u);
        const uint4 ai = make_uint4(__builtin_amdgcn_alignbit(a.x, a.x, 16), __builtin_amdgcn_alignbit(a.y, a.y, 16),
                                    __builtin_amdgcn_alignbit(a.z, a.z, 16), __builtin_amdgcn_alignbit(a.w, a.w, 16));
        const half8 avr = __builtin_bit_cast(half8, ar), avi = __builtin_bit_cast(half8, ai);
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const half8 bv = __builtin_bit_cast(half8, sb[(s * 2 + c) * 64 + lane]);
            ac[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(avr, bv, ac[c], 0, 0, 0);
            ac[c + 2] = __builtin_amdgcn_mfma_f32_32x32x16_f16(avi, bv, ac[c + 2], 0, 0, 0);
        }
    };
    // DUAL: the chunk's steps, each split at the row's facet start (lane (j, h)
    // holds emitters 8 k + 4 h .. + 3 of row j, one complex FP16 word each)
    auto dual_chunk = [&](const uint4 *a, int kb, const uint4 *bb) {
        if ((kb + kKC) * 8 <= fsMin) {  // below every row's facet: lower-facet terms
#pragma unroll
            for (int s = 0; s < kKC; s++) mfma_step(acc, a[s], bb, s);
        } else if (kb * 8 >= fsMax) {  // at or above every row's facet start
#pragma unroll
            for (int s = 0; s < kKC; s++) mfma_step(accU, a[s], bb, s);
        } else {
#pragma unroll
            for (int s = 0; s < kKC; s++) {
                const int j0 = (kb + s) * 8 + 4 * (lane >> 5);
                const uint4 q = a[s];
                const uint4 lo = make_uint4(j0 < fsRow ? q.x : 0u, j0 + 1 < fsRow ? q.y : 0u,
                                            j0 + 2 < fsRow ? q.z : 0u, j0 + 3 < fsRow ? q.w : 0u);
                const uint4 hi = make_uint4(q.x ^ lo.x, q.y ^ lo.y, q.z ^ lo.z, q.w ^ lo.w);
                mfma_step(acc, lo, bb, s);
                mfma_step(accU, hi, bb, s);
            }
        }
    };

    auto load_a = [](const uint4 *p) { return NT ? load_nt(p) : *p; };
    // K step of the split's i-th chunk; past the end the last chunk again, so that
    // the loop's requests ahead are unconditional (and those re-reads hit L2)
    auto kat = [&](int i) { return sw.k(sw.o0 + min(i, nch - 1)); };
    // B goes global -> LDS directly (global_load_lds_dwordx4: lane l of wave w
    // writes element 256 r + 64 w + l of the chunk in request r) into two buffers
    // used in turn; a chunk's 8 K steps are contiguous in B.
    typedef __attribute__((address_space(3))) void lds_void;
    auto stage_b = [&](int k, uint4 *bb) {
#pragma unroll
        for (int r = 0; r < kKC / 2; r++)
            __builtin_amdgcn_global_load_lds((const void *)(Bt + (size_t)k * 128 + r * 256),
                                             (lds_void *)&bb[r * 256 + wave * 64], 16, 0, 0);
    };
    // s_waitcnt immediate (gfx9 encoding): vector-memory count <= kKC, the export and
    // LDS/scalar counts at their maxima (not waited on)
    constexpr int kWaitAllButNextA = (kKC & 0xf) | ((kKC >> 4) << 14) | (0x7 << 4) | (0xf << 8);
    uint4 a0[kKC], a1[kKC], a2[kKC];
    // One accumulator set: one chunk per phase, A requested two chunks ahead into
    // three register sets.
    // Phase I: wait until this chunk's A and B have landed (the vector-memory
    // count completes in issue order, and the one request issued after this
    // chunk's B, the next chunk's A, may stay in flight), barrier (every wave's
    // part of B in LDS; every wave done reading the other buffer), stage the next
    // chunk's B into the other buffer, request A two chunks ahead, then run this
    // chunk's products while two chunks of A are in flight. Every request is
    // unconditional (past the end the last chunk again), which keeps the register
    // sets fixed across the loop; the loop stops after the last chunk's phase.
#define SR_GEMM_PHASE(ACUR, AFAR, BUF, NBUF, I)                                                      \
    {                                                                                                \
        const int kc = kat(I), kn = kat((I) + 1), kf = kat((I) + 2);                                 \
        __builtin_amdgcn_s_waitcnt(kWaitAllButNextA); /* all but the next chunk's A: A, B of kc */  \
        workgroup_barrier(); /* BUF complete; every wave is done reading NBUF */                     \
        stage_b(kn, NBUF);                                                                           \
        _Pragma("unroll") for (int s = 0; s < kKC; s++) AFAR[s] = load_a(Arow + (size_t)(kf + s) * 64); \
        _Pragma("unroll") for (int s = 0; s < kKC; s++) mfma_step(acc, ACUR[s], BUF, s);            \
        kLast = kc;                                                                                  \
        if (kc < kBnd && (I) + 1 < nch && kn >= kBnd) { /* lower-facet part done */                  \
            store_part(t.partL, acc);                                                                \
            _Pragma("unroll") for (int c = 0; c < 4; c++) acc[c] = f32x16{};                         \
        }                                                                                            \
        if ((I) + 1 >= nch) break;                                                                   \
    }
    // DUAL (two accumulator sets): one chunk ahead, two register sets
#define SR_GEMM_PHASE1(ACUR, ANEXT, BUF, NBUF, I)                                                    \
    {                                                                                                \
        const int kc = kat(I), kn = kat((I) + 1);                                                    \
        _Pragma("unroll") for (int s = 0; s < kKC; s++) ANEXT[s] = load_a(Arow + (size_t)(kn + s) * 64); \
        __builtin_amdgcn_s_waitcnt(kWaitAllButNextA); /* all but those kKC loads: A, B of kc */      \
        workgroup_barrier(); /* BUF complete; every wave is done reading NBUF */                     \
        stage_b(kn, NBUF);                                                                           \
        if ((I) < nch) dual_chunk(ACUR, kc, BUF);                                                    \
    }
    if (DUAL && nch > 0) {
        stage_b(kat(0), sB0);
#pragma unroll
        for (int s = 0; s < kKC; s++) a0[s] = load_a(Arow + (size_t)(kat(0) + s) * 64);
        for (int i = 0; i < nch; i += 2) {
            SR_GEMM_PHASE1(a0, a1, sB0, sB1, i)
            SR_GEMM_PHASE1(a1, a0, sB1, sB0, i + 1)
        }
        __builtin_amdgcn_s_waitcnt(0);
    }
#undef SR_GEMM_PHASE1
    if (!DUAL && nch > 0) {
        // issue order A(0), B(0), A(1): the first phase's wait leaves A(1) in flight
#pragma unroll
        for (int s = 0; s < kKC; s++) a0[s] = load_a(Arow + (size_t)(kat(0) + s) * 64);
        stage_b(kat(0), sB0);
#pragma unroll
        for (int s = 0; s < kKC; s++) a1[s] = load_a(Arow + (size_t)(kat(1) + s) * 64);
        for (int i = 0;; i += 6) {
            SR_GEMM_PHASE(a0, a2, sB0, sB1, i)
            SR_GEMM_PHASE(a1, a0, sB1, sB0, i + 1)
            SR_GEMM_PHASE(a2, a1, sB0, sB1, i + 2)
            SR_GEMM_PHASE(a0, a2, sB1, sB0, i + 3)
            SR_GEMM_PHASE(a1, a0, sB0, sB1, i + 4)
            SR_GEMM_PHASE(a2, a1, sB1, sB0, i + 5)
        }
        // the last phase's requests (re-reads of the last chunk) land before the
        // workgroup ends: its LDS may go to another workgroup next
        __builtin_amdgcn_s_waitcnt(0);
    }
#undef SR_GEMM_PHASE
    if (DUAL) {  // both partials of every split, zero where it summed nothing
        store_part(t.partL, acc);
        store_part(t.partU, accU);
    } else if (!t.kB) {
        store_part(t.partL, acc);  // every split's partial, zero when it summed nothing
    } else if (kLast >= 0) {
        store_part(kLast < kBnd ? t.partL : t.partU, acc);
    }
    if (t.grpCount) reduce_group_last(t, grp);
}

// Rows whose own input pressure is exactly zero in some plane (the zero-input
// gate of OptiX.cu:241-245 drops their higher-facet terms there).
__global__ void k_gated_rows(const float2 *__restrict__ Pin, int n, int nplanes, int *__restrict__ list,
                             int *__restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool z = false;
    if (i < n)
#pragma unroll 8
        for (int b = 0; b < nplanes; b++) {
            const float2 v = Pin[(size_t)b * n + i];
            z |= (v.x == 0.0f && v.y == 0.0f);
        }
    const unsigned long long m = __ballot(z);
    if (!m) return;
    int base = 0;
    if (__lane_id() == 0) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, 0);
    if (z) list[base + __popcll(m & ((1ull << __lane_id()) - 1ull))] = i;
}

// One (plane, row of S) of the gated reduction: sums the splits' lower- and
// higher-facet partials, applies the zero-input gate and the column scale, stores
// P_out and adds it into P. Threads run over the rows of S (padding included) so
// that the partials' addresses and the group's split classes (cls, uniform over
// a wave: waves never straddle a 128-row group) need no lookup first; the active
// row comes from rowAct alongside. Returns the stored value, zero for padding.
__device__ inline float2 reduce_gate_row(const ReduceGateTask &t, int b, int sr, int &r) {
    r = -1;
    if (sr >= t.nPadRows) return make_float2(0.f, 0.f);
    r = t.rowAct ? t.rowAct[sr] : sr;
    const int grp = __builtin_amdgcn_readfirstlane(sr >> 7);
    const unsigned char *cls = t.cls + (size_t)grp * t.nsplit;
    float2 L = make_float2(0.f, 0.f), U = make_float2(0.f, 0.f);
    // One load per split, unconditional so that the splits' loads are in flight
    // together: its lower-facet partial if it wrote one, else its higher-facet one
    // (an unwritten partial is in bounds and selected away). The split that spans
    // the group's facet boundary wrote both: its higher-facet partial follows.
#pragma unroll 8
    for (int s = 0; s < t.nsplit; s++) {
        const size_t o = ((size_t)s * 64 + b) * t.nPadRows + sr;
        const unsigned c = cls[s];
        const float2 v = ((c & 1u) ? t.partL : t.partU)[o];
        const bool toL = (c & 1u) != 0u, toU = c == 2u;
        L.x += toL ? v.x : 0.f;
        L.y += toL ? v.y : 0.f;
        U.x += toU ? v.x : 0.f;
        U.y += toU ? v.y : 0.f;
    }
    for (int s = 0; s < t.nsplit; s++)
        if (cls[s] == 3u) {
            const float2 u = t.partU[((size_t)s * 64 + b) * t.nPadRows + sr];
            U.x += u.x;
            U.y += u.y;
        }
    if (r < 0) return make_float2(0.f, 0.f);
    const float2 gv = t.gval[(size_t)b * t.n + r];
    const float g = (gv.x == 0.0f && gv.y == 0.0f) ? 0.0f : 1.0f;
    const float sc = t.colInvScale[b];
    const float2 v = make_float2((L.x + g * U.x) * sc, (L.y + g * U.y) * sc);
    t.out[(size_t)b * t.n + r] = v;
    float2 &P = t.pcell[(size_t)b * t.pcellStride + t.rowCell[r]];
    P.x += v.x;
    P.y += v.y;
    return v;
}

// The reduction inside the product, run by each of its workgroups after its
// partial-sum stores: the group's last workgroup to finish (its counter, zero
// before the launch) reduces the group's 128 rows of S for every plane
// (reduce_gate_row: the splits in split order, as k_reduce_gate), each wave's
// 64 rows also giving one column maximum per plane, and sets the counter back
// to zero for the next launch. The fences order every workgroup's partial sums
// before its count, and the last workgroup's reads after the counts.
__device__ void reduce_group_last(const TransferGemm &t, int grp) {
    __shared__ unsigned sLast;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) sLast = atomicAdd(&t.grpCount[grp], 1u) == (unsigned)t.nsplit - 1u;
    __syncthreads();
    if (!sLast) return;  // uniform
    __threadfence();
    const ReduceGateTask &rt = t.red;
    const int sr = grp * 128 + (threadIdx.x & 127);
    const int lane = threadIdx.x & 63;
    for (int b = threadIdx.x >> 7; b < rt.nrhs; b += 2) {  // uniform over each wave
        int r;
        const float2 v = reduce_gate_row(rt, b, sr, r);
        if (!rt.colMaxNext) continue;
        float m = 0.f;
        if (r >= 0) {
            const float a = rt.actP[r].w;
            m = fmaxf(fabsf(a * v.x), fabsf(a * v.y));
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if (lane == 0 && m > 0.f) atomicMax(&rt.colMaxNext[b * kColStride], __float_as_uint(m));
    }
    if (threadIdx.x == 0) t.grpCount[grp] = 0u;
}

// Gated reduction: a workgroup = 256 rows of S for one plane (blockIdx.y).
__global__ __launch_bounds__(256) void k_reduce_gate(ReduceGateTask t) {
    const int b = blockIdx.y;
    int r;
    const float2 v = reduce_gate_row(t, b, blockIdx.x * 256 + threadIdx.x, r);
    if (!t.colMaxNext) return;  // uniform
    __shared__ float red[4];
    float m = 0.f;
    if (r >= 0) {
        const float a = t.actP[r].w;
        m = fmaxf(fabsf(a * v.x), fabsf(a * v.y));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        if (m > 0.f) atomicMax(&t.colMaxNext[b * kColStride], __float_as_uint(m));
    }
}


// ---------------------------------------------------------------------------

int transfer_near(const TransferGeom &g, const TransferPhys &ph, int nQuad, int nMlp, uint4 *A, hipStream_t st) {
    if (nQuad > ph.nearCap || nMlp > ph.nearCap) { set_error("near lists overflow"); return -2; }
    if (nMlp > 0 && !ph.mlp) { set_error("learned near field without a network"); return -2; }
    const long quadBlocks = ((long)nQuad * 16 + 255) / 256;
    const int mlpBlocks = nMlp > 0 ? std::min((nMlp + 127) / 128, 2048) : 0;
    if (quadBlocks + mlpBlocks == 0) return 0;
    k_near<<<(unsigned)(quadBlocks + mlpBlocks), 256, 0, st>>>(g, ph, nQuad, nMlp, (int)quadBlocks, A);
    SR_HIP(hipGetLastError());
    return 0;
}

int transfer_col_max(const float4 *p, const float2 *Pin, long stride, int n, int nplanes, unsigned *colMax,
                     int *nextState, int nextWords, hipStream_t st) {
    if (nplanes <= 0 || n <= 0) return 0;
    const dim3 grid(nplanes, std::max(1, std::min(16, n / 1024)));
    k_col_max<<<grid, 256, 0, st>>>(p, Pin, stride, n, colMax, nextState, nextWords);
    SR_HIP(hipGetLastError());
    return 0;
}

int transfer_build_b(const float4 *p, const float2 *Pin, long stride, int n, int nKs, int nplanes,
                     const unsigned *colMax, float sS, float *scale, float *invScale, uint4 *B, hipStream_t st,
                     int *clearState, int clearWords, const int *kAct) {
    if (nKs % 8) { set_error("B build: K steps must be padded to whole chunks of 8"); return -2; }
    const long blocks = std::max({(long)nKs / 8 * 4, ((long)nplanes + 255) / 256,
                                  clearState ? ((long)clearWords + 255) / 256 : 0L});
    if (blocks <= 0) return 0;
    k_build_b<<<(unsigned)blocks, 256, 0, st>>>(p, Pin, stride, n, nKs, nplanes, colMax, sS, scale, invScale, B,
                                                clearState, clearWords, kAct);
    SR_HIP(hipGetLastError());
    return 0;
}

int transfer_gemm(const TransferGemm &t, hipStream_t st) {
    if (t.n <= 0) return 0;
    if (t.nKs % kKC || t.nRt % 4 || t.nsplit < 1) {
        set_error("transfer GEMM: K steps and row tiles must be padded (8, 4), at least one split");
        return -2;
    }
    if (t.rowFs && (t.kB || !t.partU)) {
        set_error("transfer GEMM: two accumulators need partU and exclude kB");
        return -2;
    }
    if (t.grpCount && (t.red.nPadRows != t.n || t.red.nsplit != t.nsplit || t.red.nrhs != t.nplanes ||
                       (t.red.colMaxNext && !t.red.actP) || (!t.red.rowAct && t.red.nPadRows != t.red.n))) {
        set_error("transfer GEMM: the reduction inside it needs the product's rows, splits and planes");
        return -2;
    }
    const long grid = (long)((t.nRt + 3) / 4) * t.nsplit;
    if (grid > 0x7fffffffL) { set_error("transfer GEMM: grid too large"); return -2; }
    if (t.rowFs) {
        if (t.streamS) k_transfer_gemm<2, true><<<(unsigned)grid, 256, 0, st>>>(t);
        else k_transfer_gemm<2, false><<<(unsigned)grid, 256, 0, st>>>(t);
    } else {
        if (t.streamS) k_transfer_gemm<0, true><<<(unsigned)grid, 256, 0, st>>>(t);
        else k_transfer_gemm<0, false><<<(unsigned)grid, 256, 0, st>>>(t);
    }
    SR_HIP(hipGetLastError());
    return 0;
}

int transfer_gated_rows(const float2 *Pin, int n, int nplanes, int *list, int *count, hipStream_t st) {
    if (n <= 0) return 0;
    k_gated_rows<<<(n + 255) / 256, 256, 0, st>>>(Pin, n, nplanes, list, count);
    SR_HIP(hipGetLastError());
    return 0;
}

int transfer_reduce_gate(const ReduceGateTask &t, hipStream_t st) {
    if (t.n <= 0 || t.nrhs <= 0) return 0;
    if (t.colMaxNext && !t.actP) { set_error("gated reduction: column maxima need the fragment weights"); return -2; }
    if (!t.rowAct && t.nPadRows != t.n) { set_error("gated reduction: padded rows need the row map"); return -2; }
    k_reduce_gate<<<dim3((unsigned)((t.nPadRows + 255) / 256), (unsigned)t.nrhs), 256, 0, st>>>(t);
    SR_HIP(hipGetLastError());
    return 0;
}



}  // namespace sr
