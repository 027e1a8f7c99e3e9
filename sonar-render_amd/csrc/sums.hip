// Discrete Kirchhoff gathers, plain FP32 form, for gfx950.
//
// Reference sums (src/OptiX.cu):
//   point->fragment   :35-110   P_B[j] += cos_inc * P_s * e^{ikr} / r
//   fragment->point   :112-190  P_q    += A_i cos_s P_i (-e^{ikr}/(4 pi) (ik/r) + 1/r^2)
//   fragment->fragment:192-406  P_out  += A_i cos_s cos_i P_in 0.9 G, G as above or ik/(4 pi) for r < delta
// The reference runs one thread per facet and scatters with FP64 CAS atomics.
// Here every receiver is owned by one lane (no atomics, deterministic order):
// a wave holds 64 receivers, a workgroup 4 waves, and emitters stream through
// LDS in tiles of 256 that every lane reads by broadcast. Which pairs
// contribute (cosine culls and the shadow ray) is one bit per pair, computed
// beforehand by the visibility kernel; a zero 64-bit word skips the emitter for
// the whole wave. Long emitter lists are split across workgroups (split-K) and
// the partial sums are reduced in a fixed order.
//
// Phase: sin/cos of k r are taken in revolutions, x = r * k/(2 pi), reduced to
// [-1/2, 1/2] before v_sin_f32 / v_cos_f32. The split hi+lo product keeps the
// fractional part accurate for r up to km.

#include <hip/hip_runtime.h>
#include <cstdint>
#include <algorithm>

#include "sums.h"
#include "pair_kernel.h"
#include "p2f_term.h"
#include "launch.h"

namespace sr {

constexpr int kTile = 256;

template <int KIND, int NB, bool ATT, bool SPLITG>
__device__ __attribute__((always_inline)) inline void tile_loop(const Phys &ph, int n, const float4 *s_p, const float4 *s_n, const float2 (*s_x)[kTile],
                                 const uint64_t *s_m, int lane, float4 rp, float4 rn, int rfacet, float2 *accL,
                                 float2 *accU, float2 *compL) {
    for (int k = 0; k < n; k++) {
        uint64_t word = s_m[k];
        if (word == 0) continue;  // wave-uniform
        bool bit = (word >> lane) & 1ull;
        float4 ep = s_p[k], en = s_n[k];
        float Kre, Kim;
        pair_kernel<KIND, ATT>(ph, rp, rn, ep, en, bit, Kre, Kim);
        if (KIND == KIND_F2P) {
            // compensated: K * (A P) for one plane per lane register pair
#pragma unroll
            for (int b = 0; b < NB; b++) {
                float2 X = s_x[b][k];
                float cre = Kre * X.x - Kim * X.y;
                float cim = Kre * X.y + Kim * X.x;
                two_sum_acc(accL[b].x, compL[b].x, cre);
                two_sum_acc(accL[b].y, compL[b].y, cim);
            }
        } else if (SPLITG) {
            bool up = __float_as_int(en.w) > rfacet;
            float KUre = up ? Kre : 0.0f, KUim = up ? Kim : 0.0f;
            float KLre = up ? 0.0f : Kre, KLim = up ? 0.0f : Kim;
#pragma unroll
            for (int b = 0; b < NB; b++) {
                float2 X = s_x[b][k];
                cmac<NB>(accL[b], KLre, KLim, X);
                cmac<NB>(accU[b], KUre, KUim, X);
            }
        } else {
#pragma unroll
            for (int b = 0; b < NB; b++) cmac<NB>(accL[b], Kre, Kim, s_x[b][k]);
        }
    }
}

template <int NB>
struct GatherLds {
    float4 p[kTile];
    float4 n[kTile];
    float2 x[NB][kTile];
    uint64_t m[4][kTile];
};

// One workgroup's share of a gather: receiver blocks 4*wg .. 4*wg+3, emitter range `split`.
// Called by all 256 threads of the workgroup (it synchronises internally).
template <int KIND, int NB, bool ATT>
__device__ __attribute__((always_inline)) inline void gather_block(const GatherTask &t, const Phys &ph, int wg,
                                                                   int split, GatherLds<NB> &L) {
    float4 *s_p = L.p;
    float4 *s_n = L.n;
    float2 (*s_x)[kTile] = L.x;
    uint64_t (*s_m)[kTile] = L.m;

    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int nrb = (t.nr + 63) >> 6;
    const int rb = wg * 4 + wave;
    const int e0 = split * t.emPerSplit;
    const int e1 = min(t.ne, e0 + t.emPerSplit);
    const int r = rb * 64 + lane;
    const bool rvalid = (rb < nrb) && (r < t.nr);

    float4 rp = rvalid ? t.rp[r] : make_float4(0.f, 0.f, 0.f, 0.f);
    float4 rn = (rvalid && t.rn) ? t.rn[r] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int rfacet = __float_as_int(rn.w);

    // Does this wave hold a receiver whose own input pressure is exactly zero?
    bool needSplit = false;
    if (KIND == KIND_F2F && t.gateMode == GATE_UPPER) {
        bool z = false;
        if (rvalid)
            for (int b = 0; b < t.nrhs; b++) {
                float2 g = t.gval[b * t.gvalStride + r];
                z |= (g.x == 0.0f && g.y == 0.0f);
            }
        needSplit = __ballot(z) != 0;
    }

    float2 accL[NB], accU[NB], compL[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        accL[b] = make_float2(0.f, 0.f);
        accU[b] = make_float2(0.f, 0.f);
        compL[b] = make_float2(0.f, 0.f);
    }

    for (int base = e0; base < e1; base += kTile) {
        const int n = min(kTile, e1 - base);
        __syncthreads();
        for (int k = threadIdx.x; k < n; k += 256) {
            const int e = base + k;
            float4 p = t.ep[e];
            s_p[k] = p;
            s_n[k] = t.en ? t.en[e] : make_float4(0.f, 0.f, 0.f, 0.f);
            const int ci = t.ecell ? t.ecell[e] : e;
#pragma unroll
            for (int b = 0; b < NB; b++) {
                float2 v = (b < t.nrhs) ? t.eval[b * t.evalStride + ci] : make_float2(0.f, 0.f);
                s_x[b][k] = make_float2(p.w * v.x, p.w * v.y);
            }
        }
        for (int w = 0; w < 4; w++) {
            const int rbw = wg * 4 + w;
            for (int k = threadIdx.x; k < n; k += 256)
                s_m[w][k] = (rbw < nrb) ? t.mask[(size_t)rbw * t.ne + base + k] : 0ull;
        }
        __syncthreads();
        if (rb < nrb) {
            if (needSplit)
                tile_loop<KIND, NB, ATT, true>(ph, n, s_p, s_n, s_x, s_m[wave], lane, rp, rn, rfacet, accL, accU,
                                               compL);
            else
                tile_loop<KIND, NB, ATT, false>(ph, n, s_p, s_n, s_x, s_m[wave], lane, rp, rn, rfacet, accL, accU,
                                                compL);
        }
    }

    if (!rvalid) return;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        if (b >= t.nrhs) continue;
        const size_t o = ((size_t)split * NB + b) * t.nr + r;
        if (KIND == KIND_F2P) {
            t.partL[o] = accL[b];
            t.partU[o] = compL[b];
        } else {
            t.partL[o] = accL[b];
            if (t.partU) t.partU[o] = accU[b];
        }
    }
}

template <int KIND, int NB, bool ATT>
__global__ __launch_bounds__(256) void k_gather(GatherTask t, Phys ph) {
    __shared__ GatherLds<NB> L;
    gather_block<KIND, NB, ATT>(t, ph, blockIdx.x, blockIdx.y, L);
}

constexpr int kGatherClaim = 4;  // items a workgroup claims at once from the work counter

// Persistent form: workgroups pull (task, receiver-block group, emitter range)
// items from a shared counter until the list is exhausted; every task may be a
// different scene with its own sizes and wavenumber.
template <int KIND, int NB, bool ATT>
__global__ __launch_bounds__(256) void k_gather_multi(const GatherTask *__restrict__ tasks,
                                                      const Phys *__restrict__ phys, const int4 *__restrict__ items,
                                                      int nitems, int *counter) {
    __shared__ GatherLds<NB> L;
    __shared__ int s_item;
    for (;;) {
        // kGatherClaim consecutive items per claim (fewer atomics queued on the counter)
        if (threadIdx.x == 0) s_item = atomicAdd(counter, kGatherClaim);
        __syncthreads();
        const int it0 = s_item;
        __syncthreads();
        if (it0 >= nitems) break;
        const int it1 = min(nitems, it0 + kGatherClaim);
        for (int it = it0; it < it1; it++) {
            const int4 w = items[it];
            const GatherTask t = tasks[w.x];
            const Phys ph = phys[w.x];
            gather_block<KIND, NB, ATT>(t, ph, w.y, w.z, L);
        }
    }
}

// ---------------------------------------------------------------------------
// Fragment->fragment, 16..64 incidences: FP32 pair kernel on the VALU, complex
// multiply-accumulate on the FP32 matrix cores (v_mfma_f32_32x32x2_f32).
//
// With X = A_e P_in[b][e], the complex product takes three real products
// (Gauss): P1 = K_r X_r, P2 = K_i X_i, P3 = (K_r + K_i)(X_r + X_i), summed over
// the emitters, then Re = P1 - P2 and Im = P3 - P1 - P2 - 6 matrix operations
// per emitter pair instead of 8. A wave owns 32 receivers and 64 planes (6
// accumulator tiles of 32 x 32: three products x two halves of the planes);
// its 64 lanes evaluate 2 emitters x 32 receivers at once, and the K dimension
// of each v_mfma_f32_32x32x2_f32 is those two emitters, so every lane's own
// kernel value is directly its A operand:
//   A operand: lane l holds A[row l % 32][k l / 32]; B: lane l holds B[k l / 32][col l % 32];
//   C: element i of lane l is (row (i/4)*8 + (l/32)*4 + i%4, col l % 32).
// Emitter tiles (pressures of all planes, positions, normals, visibility
// words) are copied global -> LDS with global_load_lds (no registers), double
// buffered one tile ahead; each lane picks its B values from the raw pressure
// rows, and the emitter area scales K. A pair of emitters whose visibility
// bits are all zero for the wave's 32 receivers is skipped. The zero-input
// gate is not applied here: the rows it affects are recomputed afterwards as a
// compact receiver set with the lower-facet emitters only.
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void lds_void;
constexpr int kMT = 32;       // emitters per LDS tile
constexpr int kXPitch = 66;   // floats per plane row of a tile (32 float2 + 2: conflict-free columns)

// One tile of emitter data, copied global -> LDS by global_load_lds_dword
// (no registers involved): lane l of a wave-instruction writes dword l.
struct MfmaTile {
    float x[64][kXPitch];  // P_in[plane][emitter] as (re, im) pairs
    float4 p[kMT], n[kMT];
    uint64_t m[2][kMT];    // visibility words of the workgroup's two 64-receiver blocks
};

// Requests tile [base, base + kMT) into T. Emitter indices are clamped into
// range (the values past the tile's end are never used: their bits read as 0).
// GATHER: the pressures are read through t.ecell (emitter -> column of eval).
template <bool GATHER>
__device__ __attribute__((always_inline)) inline void mfma_tile_dma(const GatherTask &t, int base, int rb0, int nrb,
                                                                    int wave, int lane, MfmaTile &T) {
    const int last = t.ne - 1;
    // x: plane q = 16 w + i, one dword per lane: emitter lane / 2, component lane & 1
    {
        int e = min(base + (lane >> 1), last);
        if (GATHER && t.ecell) e = t.ecell[e];
        for (int i = 0; i < 16; i++) {
            const int q = wave * 16 + i;
            if (q >= t.nrhs) break;  // wave-uniform; the columns of absent planes are never written out
            const float *g = reinterpret_cast<const float *>(t.eval + (size_t)q * t.evalStride + e) + (lane & 1);
            __builtin_amdgcn_global_load_lds(g, (lds_void *)&T.x[q][0], 4, 0, 0);
        }
    }
    // positions (wave 0, 1) and normals (wave 2, 3): 16 float4 = 64 dwords per instruction
    {
        const float4 *src = wave < 2 ? t.ep : t.en;
        const int k = (wave & 1) * 16 + (lane >> 2);
        const float *g = reinterpret_cast<const float *>(src + min(base + k, last)) + (lane & 3);
        float4 *dst = wave < 2 ? &T.p[(wave & 1) * 16] : &T.n[(wave & 1) * 16];
        __builtin_amdgcn_global_load_lds(g, (lds_void *)dst, 4, 0, 0);
    }
    // visibility words: block w of the two, 32 words = 64 dwords, by waves 0 and 1
    if (wave < 2) {
        const int rb = min(rb0 + wave, nrb - 1);  // rows past the end: the wave is inactive
        const int k = lane >> 1;
        const uint32_t *g =
            reinterpret_cast<const uint32_t *>(t.mask + (size_t)rb * t.ne + min(base + k, last)) + (lane & 1);
        __builtin_amdgcn_global_load_lds(g, (lds_void *)&T.m[wave][0], 4, 0, 0);
    }
}

// KIND_F2F: receivers are fragments (both cosines); KIND_F2P: receivers are
// field points (emitter cosine only, the phase split for long ranges).
// ROLE 1: the compact set of gated rows (identical code, its own name in profiles).
template <int KIND, bool ATT, int ROLE>
__global__ __launch_bounds__(256, 3) void k_f2f_mfma(GatherTask t, Phys ph) {
    __shared__ MfmaTile TA, TB;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int h = lane >> 5, j = lane & 31;
    const int rt = blockIdx.x * 4 + wave;  // 32-receiver tile of this wave
    const int split = blockIdx.y;
    const int row = rt * 32 + j;
    const bool rvalid = row < t.nr;
    const bool waveActive = rt * 32 < t.nr;
    const int nrb = (t.nr + 63) >> 6;
    const int rb0 = blockIdx.x * 2;
    const float4 rp = rvalid ? t.rp[row] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 rn = (rvalid && KIND == KIND_F2F) ? t.rn[row] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int mslot = wave >> 1;
    const int mshift = (wave & 1) * 32 + j;
    const int e0 = split * t.emPerSplit;
    const int e1 = min(t.ne, e0 + t.emPerSplit);

    // three real products per complex term (Gauss): P1 = sum Kr Xr, P2 = sum Ki Xi,
    // P3 = sum (Kr + Ki)(Xr + Xi); Re = P1 - P2, Im = P3 - P1 - P2. Tiles: [product][planes 0-31, 32-63]
    f32x16 acc[6];
#pragma unroll
    for (int c = 0; c < 6; c++) acc[c] = f32x16{};

    // The two tile buffers are separate LDS objects and every phase names them
    // statically, so the compiler sees that reads of one never wait on copies
    // into the other.
    unsigned steps = 0;  // emitter-pair steps executed (wave-uniform)
    auto compute = [&](const MfmaTile &T, int ne) {
        // emitter pairs with any visible receiver of this wave: lane q < kMT/2 tests pair q
        bool act = false;
        if (waveActive && lane < kMT / 2 && 2 * lane < ne) {
            uint64_t w = T.m[mslot][2 * lane];
            if (2 * lane + 1 < ne) w |= T.m[mslot][2 * lane + 1];
            act = ((w >> ((wave & 1) * 32)) & 0xffffffffull) != 0ull;
        }
        uint64_t pairs = __ballot(act);
        steps += (unsigned)__popcll(pairs);
        while (pairs) {  // a straight-line body keeps the accumulators in place
            const int k = 2 * __builtin_ctzll(pairs);
            pairs &= pairs - 1;
            const int ke = k + h;  // this lane's emitter (kMT is even: ke < kMT)
            const bool bit = ke < ne && ((T.m[mslot][ke] >> mshift) & 1ull);
            const float4 ep = T.p[ke];
            float Kre, Kim;
            pair_kernel<KIND, ATT>(ph, rp, rn, ep, T.n[ke], bit, Kre, Kim);
            Kre *= ep.w;  // emitter area
            Kim *= ep.w;
            // K = the two emitters: lane (j, h) holds A[j][h] (its own emitter k + h) and
            // B[h][col j] = X of emitter k + h, plane j (lo) and 32 + j (hi)
            const float2 xl = *reinterpret_cast<const float2 *>(&T.x[j][2 * ke]);
            const float2 xh = *reinterpret_cast<const float2 *>(&T.x[32 + j][2 * ke]);
            const float ks = Kre + Kim;
            acc[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(Kre, xl.x, acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_32x32x2f32(Kre, xh.x, acc[1], 0, 0, 0);
            acc[2] = __builtin_amdgcn_mfma_f32_32x32x2f32(Kim, xl.y, acc[2], 0, 0, 0);
            acc[3] = __builtin_amdgcn_mfma_f32_32x32x2f32(Kim, xh.y, acc[3], 0, 0, 0);
            acc[4] = __builtin_amdgcn_mfma_f32_32x32x2f32(ks, xl.x + xl.y, acc[4], 0, 0, 0);
            acc[5] = __builtin_amdgcn_mfma_f32_32x32x2f32(ks, xh.x + xh.y, acc[5], 0, 0, 0);
        }
    };
    if (e0 < e1) mfma_tile_dma<KIND == KIND_F2P>(t, e0, rb0, nrb, wave, lane, TA);
    int base = e0;
    while (base < e1) {
        __builtin_amdgcn_s_waitcnt(0);  // this wave's share of TA has landed
        __syncthreads();                // TA complete; every wave is done with TB
        if (base + kMT < e1) mfma_tile_dma<KIND == KIND_F2P>(t, base + kMT, rb0, nrb, wave, lane, TB);
        compute(TA, min(kMT, e1 - base));
        base += kMT;
        if (base >= e1) break;
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (base + kMT < e1) mfma_tile_dma<KIND == KIND_F2P>(t, base + kMT, rb0, nrb, wave, lane, TA);
        compute(TB, min(kMT, e1 - base));
        base += kMT;
    }
    if (t.workCount && lane == 0 && steps) atomicAdd(t.workCount, (unsigned long long)steps);
    if (!waveActive) return;
    // column j of tiles 0/2/4 is plane j, of tiles 1/3/5 plane 32 + j
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const int b = c * 32 + j;
        if (b >= t.nrhs) continue;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int r = rt * 32 + (i >> 2) * 8 + h * 4 + (i & 3);
            if (r >= t.nr) continue;
            const float p1 = acc[c][i], p2 = acc[c + 2][i], p3 = acc[c + 4][i];
            t.partL[((size_t)split * 64 + b) * t.nr + r] = make_float2(p1 - p2, p3 - p1 - p2);
        }
    }
}

// ---- gated rows of a matrix-core bounce, as a compact receiver set ----
// Row i of the set is active row list[i]; its visibility words keep only the
// emitters on lower facets (e < first index of the row's facet), which is
// the zero-input gate of OptiX.cu:241-245. The set then goes through
// k_f2f_mfma unchanged, and k_gated_replace swaps the result in for the
// (row, plane) entries whose input is zero.
__global__ void k_gated_compact_rows(const int *__restrict__ list, int count, const float4 *__restrict__ p,
                                     const float4 *__restrict__ nrm, float4 *__restrict__ pc, float4 *__restrict__ nc,
                                     int *__restrict__ fsc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int r = list[i];
    const float4 rn = nrm[r];
    const int fr = __float_as_int(rn.w);
    int lo = 0, hi = r;  // first active index on facet fr (facets ascend)
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (__float_as_int(nrm[mid].w) < fr) lo = mid + 1; else hi = mid;
    }
    pc[i] = p[r];
    nc[i] = rn;
    fsc[i] = lo;
}

// word (g, e) of the compact set: bit i = visibility bit of (list[64 g + i], e) and e < fs
__global__ void k_gated_compact_mask(const int *__restrict__ list, const int *__restrict__ fsc, int count,
                                     const uint64_t *__restrict__ mask, int ne, uint64_t *__restrict__ mc) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ng = (count + 63) / 64;
    if (t >= (long)ng * ne) return;
    const int g = (int)(t / ne), e = (int)(t % ne);
    uint64_t w = 0;
    const int n = min(64, count - 64 * g);
    for (int i = 0; i < n; i++) {
        const int r = list[64 * g + i];
        if (e < fsc[64 * g + i]) w |= ((mask[(size_t)(r >> 6) * ne + e] >> (r & 63)) & 1ull) << i;
    }
    mc[t] = w;
}

// P_out[b][r] <- L, P[b][cell(r)] += L - (old P_out) where P_in[b][r] == 0; L from the partials [split][64][count].
__global__ void k_gated_replace(const float2 *__restrict__ part, int nsplit, int count, const int *__restrict__ list,
                                const float2 *__restrict__ Pin, long stride, int np, float2 *__restrict__ Pout,
                                float2 *__restrict__ P, long pStride, const int *__restrict__ rowCell,
                                const float *__restrict__ colScale) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)count * np) return;
    const int b = (int)(k / count), i = (int)(k % count);
    const int r = list[i];
    const float2 g = Pin[(size_t)b * stride + r];
    if (!(g.x == 0.0f && g.y == 0.0f)) return;
    float2 L = make_float2(0.f, 0.f);
#pragma unroll 8
    for (int s = 0; s < nsplit; s++) {
        const float2 v = part[((size_t)s * 64 + b) * count + i];
        L.x += v.x;
        L.y += v.y;
    }
    if (colScale) {
        L.x *= colScale[b];
        L.y *= colScale[b];
    }
    float2 &o = Pout[(size_t)b * stride + r];
    float2 &c = P[(size_t)b * pStride + rowCell[r]];
    c.x += L.x - o.x;
    c.y += L.y - o.y;
    o = L;
}

__device__ inline void reduce_elem(const ReduceTask &t, long i) {
    if (i >= (long)t.nrhs * t.nr) return;
    const int b = (int)(i / t.nr), r = (int)(i % t.nr);
    float2 L = make_float2(0.f, 0.f), U = make_float2(0.f, 0.f);
    for (int s = 0; s < t.nsplit; s++) {
        const size_t o = ((size_t)s * t.nbStride + b) * t.nr + r;
        float2 a = t.partL[o];
        L.x += a.x; L.y += a.y;
        if (t.partU) { float2 u = t.partU[o]; U.x += u.x; U.y += u.y; }
    }
    float g = 1.0f;
    if (t.gateMode != GATE_NONE) {
        float2 gv = t.gval[b * t.gvalStride + r];
        g = (gv.x == 0.0f && gv.y == 0.0f) ? 0.0f : 1.0f;
    }
    float2 v;
    if (t.gateMode == GATE_ALL) v = make_float2(g * L.x, g * L.y);
    else v = make_float2(L.x + g * U.x, L.y + g * U.y);
    if (t.colInvScale) {
        const float sc = t.colInvScale[b];
        v.x *= sc;
        v.y *= sc;
    }
    if (t.out) {
        float2 &o = t.out[b * t.outStride + r];
        if (t.outAdd) { o.x += v.x; o.y += v.y; } else o = v;
    }
    if (t.pcell) {
        float2 &P = t.pcell[b * t.pcellStride + t.rowCell[r]];
        P.x += v.x; P.y += v.y;
    }
}

__device__ inline void reduce_points_elem(const float2 *hi, const float2 *lo, int nsplit, int nbStride, int nrhs,
                                          int nq, double2 *fld, long fldStride, long i) {
    if (i >= (long)nrhs * nq) return;
    const int b = (int)(i / nq), q = (int)(i % nq);
    double sr = 0.0, si = 0.0;
    for (int s = 0; s < nsplit; s++) {
        const size_t o = ((size_t)s * nbStride + b) * nq + q;
        sr += (double)hi[o].x + (double)lo[o].x;
        si += (double)hi[o].y + (double)lo[o].y;
    }
    double2 &f = fld[b * fldStride + q];
    f.x += sr;
    f.y += si;
}

__device__ inline void prime_elem(const float2 *P, long pStride, const int *actCell, int nact, int nrhs, float2 *Pout,
                                  long outStride, long i) {
    if (i >= (long)nact * nrhs) return;
    const int b = (int)(i / nact), k = (int)(i % nact);
    Pout[b * outStride + k] = P[b * pStride + actCell[k]];
}

__global__ void k_any_zero(const float2 *v, long n, int *flag) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool z = false;
    if (i < n) {
        const float2 x = v[i];
        z = (x.x == 0.0f && x.y == 0.0f);
    }
    if (__ballot(z) && __lane_id() == 0) atomicOr(flag, 1);  // one atomic per wave
}

__global__ void k_reduce(ReduceTask t) { reduce_elem(t, (long)blockIdx.x * blockDim.x + threadIdx.x); }

// A workgroup per 64 consecutive field points of one plane (a lane each, so
// every load of a split's partial is one coalesced 512 bytes): wave w sums splits
// w, w + 16, ... in order in FP64, then the 16 wave sums are added in wave order.
constexpr int kRpWaves = 16;
__global__ __launch_bounds__(64 * kRpWaves) void k_reduce_points(const float2 *hi, const float2 *lo, int nsplit,
                                                                 int nbStride, int nq, double2 *fld, long fldStride) {
    __shared__ double2 part[kRpWaves][64];
    const int qBlocks = (nq + 63) / 64;
    const int b = blockIdx.x / qBlocks;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int q = (blockIdx.x - b * qBlocks) * 64 + lane;
    double sr = 0.0, si = 0.0;
    if (q < nq) {
#pragma unroll 4
        for (int s = w; s < nsplit; s += kRpWaves) {
            const size_t o = ((size_t)s * nbStride + b) * nq + q;
            const float2 h = hi[o];
            sr += (double)h.x;
            si += (double)h.y;
            if (lo) {
                const float2 l = lo[o];
                sr += (double)l.x;
                si += (double)l.y;
            }
        }
    }
    part[w][lane] = make_double2(sr, si);
    __syncthreads();
    if (w == 0 && q < nq) {
        for (int k = 1; k < kRpWaves; k++) {
            sr += part[k][lane].x;
            si += part[k][lane].y;
        }
        double2 &f = fld[b * fldStride + q];
        f.x += sr;
        f.y += si;
    }
}


__global__ void k_prime(const float2 *P, long pStride, const int *actCell, int nact, int nrhs, float2 *Pout,
                        long outStride) {
    prime_elem(P, pStride, actCell, nact, nrhs, Pout, outStride, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- multi-task kernels: block = (task, 256 consecutive elements of it) ----
__global__ void k_reduce_multi(const ReduceTask *__restrict__ tasks, const int2 *__restrict__ items) {
    const int2 w = items[blockIdx.x];
    reduce_elem(tasks[w.x], (long)w.y + threadIdx.x);
}
__global__ void k_reduce_points_multi(const PointsReduceTask *__restrict__ tasks, const int2 *__restrict__ items) {
    const int2 w = items[blockIdx.x];
    const PointsReduceTask t = tasks[w.x];
    reduce_points_elem(t.hi, t.lo, t.nsplit, t.nbStride, t.nrhs, t.nq, t.fld, t.fldStride, (long)w.y + threadIdx.x);
}
template <bool ATT>
__global__ __launch_bounds__(256) void k_p2f_multi(const P2FTask *__restrict__ tasks, const Phys *__restrict__ phys,
                                                   const int2 *__restrict__ items) {
    const int2 w = items[blockIdx.x];
    p2f_elem<ATT>(tasks[w.x], phys[w.x], w.y + threadIdx.x);
}
__global__ void k_prime_multi(const PrimeTask *__restrict__ tasks, const int2 *__restrict__ items) {
    const int2 w = items[blockIdx.x];
    const PrimeTask t = tasks[w.x];
    prime_elem(t.P, t.pStride, t.actCell, t.nact, t.nrhs, t.Pout, t.outStride, (long)w.y + threadIdx.x);
}
__global__ void k_zero_multi(const ZeroTask *__restrict__ tasks, const int2 *__restrict__ items) {
    const int2 w = items[blockIdx.x];
    const ZeroTask t = tasks[w.x];
    const long i = (long)w.y + threadIdx.x;
    if (i < t.n) t.p[i] = 0.0f;
}
__global__ void k_counters_zero(int *c, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = 0;
}
struct ZeroRanges {
    float2 *p[8];
    long n[8];
};
// blockIdx.y: the range; a grid-stride loop over its elements
__global__ __launch_bounds__(256) void k_zero_ranges(ZeroRanges z) {
    float2 *p = z.p[blockIdx.y];
    const long n = z.n[blockIdx.y];
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) p[i] = make_float2(0.f, 0.f);
}

// ---------------------------------------------------------------------------

int choose_splits(int nrb, int ne, int minPerSplit) {
    const int targetWaves = 8192;  // 256 CUs x 4 SIMDs x 8 waves
    int s = std::max(1, targetWaves / std::max(1, nrb));
    s = std::min(s, std::max(1, ne / std::max(1, minPerSplit)));
    return std::max(1, std::min(s, 4096));
}

template <int KIND, int NB>
static int gather_dispatch_att(const GatherTask &t, const Phys &ph, hipStream_t st) {
    const int nrb = (t.nr + 63) / 64;
    dim3 grid((nrb + 3) / 4, t.nsplit);
    if (ph.alpha != 0.0f) k_gather<KIND, NB, true><<<grid, 256, 0, st>>>(t, ph);
    else k_gather<KIND, NB, false><<<grid, 256, 0, st>>>(t, ph);
    SR_HIP(hipGetLastError());
    return 0;
}

template <int KIND>
static int gather_dispatch_nb(const GatherTask &t, const Phys &ph, int nb, hipStream_t st) {
    switch (nb) {
    case 1: return gather_dispatch_att<KIND, 1>(t, ph, st);
    case 2: return gather_dispatch_att<KIND, 2>(t, ph, st);
    case 4: return gather_dispatch_att<KIND, 4>(t, ph, st);
    case 8: return gather_dispatch_att<KIND, 8>(t, ph, st);
    default: return gather_dispatch_att<KIND, 16>(t, ph, st);
    }
}

static int nb_bucket(int nrhs) {
    if (nrhs <= 1) return 1;
    if (nrhs <= 2) return 2;
    if (nrhs <= 4) return 4;
    if (nrhs <= 8) return 8;
    return 16;
}

int gather_launch(const GatherTask &t, const Phys &ph, int kind, hipStream_t st) {
    if (t.nr <= 0 || t.ne <= 0) return 0;
    if (t.nrhs < 1 || t.nrhs > 16) { set_error("gather: 1..16 planes per launch"); return -1; }
    const int nb = nb_bucket(t.nrhs);
    switch (kind) {
    case KIND_F2F: return gather_dispatch_nb<KIND_F2F>(t, ph, nb, st);
    case KIND_F2F_FIELD: return gather_dispatch_nb<KIND_F2F_FIELD>(t, ph, nb, st);
    case KIND_F2P: return gather_dispatch_nb<KIND_F2P>(t, ph, nb, st);
    default: set_error("gather: unknown kind"); return -1;
    }
}

int gather_nb_stride(int nrhs) { return nb_bucket(nrhs); }

int gated_compact_launch(const int *list, int count, const float4 *p, const float4 *nrm, const uint64_t *mask, int ne,
                         float4 *pc, float4 *nc, int *fsc, uint64_t *mc, hipStream_t st) {
    if (count <= 0) return 0;
    k_gated_compact_rows<<<(count + 255) / 256, 256, 0, st>>>(list, count, p, nrm, pc, nc, fsc);
    SR_HIP(hipGetLastError());
    if (!mask) return 0;  // row bounds only
    const long words = (long)((count + 63) / 64) * ne;
    k_gated_compact_mask<<<(unsigned)((words + 255) / 256), 256, 0, st>>>(list, fsc, count, mask, ne, mc);
    SR_HIP(hipGetLastError());
    return 0;
}

int gated_replace_launch(const float2 *part, int nsplit, int count, const int *list, const float2 *Pin, long stride,
                         int np, float2 *Pout, float2 *P, long pStride, const int *rowCell, const float *colScale,
                         hipStream_t st) {
    const long m = (long)count * np;
    if (m <= 0) return 0;
    k_gated_replace<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(part, nsplit, count, list, Pin, stride, np, Pout, P,
                                                                 pStride, rowCell, colScale);
    SR_HIP(hipGetLastError());
    return 0;
}

int f2f_mfma_launch(const GatherTask &t, const Phys &ph, int kind, hipStream_t st, bool compactSet) {
    if (t.nr <= 0 || t.ne <= 0) return 0;
    if ((t.ecell && kind != KIND_F2P) || t.nrhs < 1 || t.nrhs > 64 || t.emPerSplit <= 0 ||
        (long)t.nsplit * t.emPerSplit < t.ne) {
        set_error("f2f_mfma: needs compact emitters (field points: or an emitter->cell map), 1..64 planes and a "
                  "covering emitter split");
        return -2;
    }
    if (kind != KIND_F2F && kind != KIND_F2P) { set_error("f2f_mfma: fragment or field-point receivers only"); return -2; }
    if (kind == KIND_F2F && !t.rn) { set_error("f2f_mfma: fragment receivers need normals"); return -2; }
    dim3 grid((t.nr + 127) / 128, t.nsplit);
    if (kind == KIND_F2F && compactSet) {
        if (ph.alpha != 0.0f) k_f2f_mfma<KIND_F2F, true, 1><<<grid, 256, 0, st>>>(t, ph);
        else k_f2f_mfma<KIND_F2F, false, 1><<<grid, 256, 0, st>>>(t, ph);
    } else if (kind == KIND_F2F) {
        if (ph.alpha != 0.0f) k_f2f_mfma<KIND_F2F, true, 0><<<grid, 256, 0, st>>>(t, ph);
        else k_f2f_mfma<KIND_F2F, false, 0><<<grid, 256, 0, st>>>(t, ph);
    } else {
        if (ph.alpha != 0.0f) k_f2f_mfma<KIND_F2P, true, 0><<<grid, 256, 0, st>>>(t, ph);
        else k_f2f_mfma<KIND_F2P, false, 0><<<grid, 256, 0, st>>>(t, ph);
    }
    SR_HIP(hipGetLastError());
    return 0;
}

int reduce_launch(const ReduceTask &t, hipStream_t st) {
    const long n = (long)t.nrhs * t.nr;
    if (n <= 0) return 0;
    k_reduce<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(t);
    SR_HIP(hipGetLastError());
    return 0;
}

int reduce_points_launch(const float2 *hi, const float2 *lo, int nsplit, int nbStride, int nrhs, int nq,
                         double2 *fld, long fldStride, hipStream_t st) {
    const long blocks = (long)nrhs * ((nq + 63) / 64);
    if (blocks <= 0) return 0;
    k_reduce_points<<<(unsigned)blocks, 64 * kRpWaves, 0, st>>>(hi, lo, nsplit, nbStride, nq, fld, fldStride);
    SR_HIP(hipGetLastError());
    return 0;
}


int prime_launch(const float2 *P, long pStride, const int *actCell, int nact, int nrhs, float2 *Pout, long outStride,
                 hipStream_t st) {
    const long n = (long)nact * nrhs;
    if (n <= 0) return 0;
    k_prime<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(P, pStride, actCell, nact, nrhs, Pout, outStride);
    SR_HIP(hipGetLastError());
    return 0;
}

int gather_multi_launch(const GatherTask *tasks, const Phys *phys, const int4 *items, int nitems, int *counter,
                        int kind, int nb, bool att, int grid, hipStream_t st) {
    if (nitems <= 0) return 0;
#define SR_GM(K, N, A) k_gather_multi<K, N, A><<<grid, 256, 0, st>>>(tasks, phys, items, nitems, counter)
#define SR_GM_NB(K, A)                   \
    switch (nb) {                        \
    case 1: SR_GM(K, 1, A); break;       \
    case 2: SR_GM(K, 2, A); break;       \
    case 4: SR_GM(K, 4, A); break;       \
    case 8: SR_GM(K, 8, A); break;       \
    default: SR_GM(K, 16, A); break;     \
    }
    if (kind == KIND_F2F) { if (att) { SR_GM_NB(KIND_F2F, true) } else { SR_GM_NB(KIND_F2F, false) } }
    else if (kind == KIND_F2F_FIELD) { if (att) { SR_GM_NB(KIND_F2F_FIELD, true) } else { SR_GM_NB(KIND_F2F_FIELD, false) } }
    else { if (att) { SR_GM_NB(KIND_F2P, true) } else { SR_GM_NB(KIND_F2P, false) } }
#undef SR_GM_NB
#undef SR_GM
    SR_HIP(hipGetLastError());
    return 0;
}

int p2f_multi_launch(const P2FTask *tasks, const Phys *phys, const int2 *items, int nitems, bool att, hipStream_t st) {
    if (nitems <= 0) return 0;
    if (att) k_p2f_multi<true><<<nitems, 256, 0, st>>>(tasks, phys, items);
    else k_p2f_multi<false><<<nitems, 256, 0, st>>>(tasks, phys, items);
    SR_HIP(hipGetLastError());
    return 0;
}
int reduce_multi_launch(const ReduceTask *tasks, const int2 *items, int nitems, hipStream_t st) {
    if (nitems <= 0) return 0;
    k_reduce_multi<<<nitems, 256, 0, st>>>(tasks, items);
    SR_HIP(hipGetLastError());
    return 0;
}
int reduce_points_multi_launch(const PointsReduceTask *tasks, const int2 *items, int nitems, hipStream_t st) {
    if (nitems <= 0) return 0;
    k_reduce_points_multi<<<nitems, 256, 0, st>>>(tasks, items);
    SR_HIP(hipGetLastError());
    return 0;
}
int prime_multi_launch(const PrimeTask *tasks, const int2 *items, int nitems, hipStream_t st) {
    if (nitems <= 0) return 0;
    k_prime_multi<<<nitems, 256, 0, st>>>(tasks, items);
    SR_HIP(hipGetLastError());
    return 0;
}
int zero_multi_launch(const ZeroTask *tasks, const int2 *items, int nitems, hipStream_t st) {
    if (nitems <= 0) return 0;
    k_zero_multi<<<nitems, 256, 0, st>>>(tasks, items);
    SR_HIP(hipGetLastError());
    return 0;
}
int counters_zero_launch(int *counters, int n, hipStream_t st) {
    k_counters_zero<<<(n + 255) / 256, 256, 0, st>>>(counters, n);
    SR_HIP(hipGetLastError());
    return 0;
}

int zero_ranges_launch(float2 *const *p, const long *n, int count, hipStream_t st) {
    for (int c0 = 0; c0 < count; c0 += 8) {
        ZeroRanges z{};
        const int k = std::min(8, count - c0);
        long mx = 0;
        for (int i = 0; i < k; i++) {
            z.p[i] = p[c0 + i];
            z.n[i] = p[c0 + i] ? n[c0 + i] : 0;
            mx = std::max(mx, z.n[i]);
        }
        if (mx <= 0) continue;
        const unsigned gx = (unsigned)std::min<long>(512, (mx + 255) / 256);
        k_zero_ranges<<<dim3(gx, k), 256, 0, st>>>(z);
        SR_HIP(hipGetLastError());
    }
    return 0;
}

int any_zero_launch(const float2 *v, long n, int *flag, hipStream_t st) {
    if (n <= 0) return 0;
    k_any_zero<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(v, n, flag);
    SR_HIP(hipGetLastError());
    return 0;
}

}  // namespace sr
