// Builds of the transfer matrix S (the stored pair kernel of the transfer form).
//
// S_ij = cos_s cos_i 0.9 G(r_ij) [visible], the reference's pair kernel of
// src/OptiX.cu:297-354 without the emitter's area, as complex FP16 in the
// product's fragment layout (transfer.hip). This translation unit is compiled
// with -ffp-contract=off: the pair kernel's fused multiply-adds are written out
// (fmaf), and the visibility test evaluated inside the symmetric build runs the
// visibility kernels' FP32 arithmetic exactly as they do (vis_fp32.h).

#include <hip/hip_runtime.h>
#include <cstdint>
#include <algorithm>

#include "launch.h"
#include "transfer.h"
#include "transfer_dev.h"
#include "near_eval.h"
#include "vis_fp32.h"

namespace sr {

// The emitter (active cell) at K position k, -1 for a padding position.
__device__ inline int emitter_at(const TransferGeom &g, int k) {
    return g.rowAct ? g.rowAct[k] : (k < g.n ? k : -1);
}

// Near pairs are staged per wave in LDS and flushed to the global lists with
// one atomic per batch (called by the whole wave; counts are wave-uniform).
constexpr int kNearStage = 256;

// entries: this wave's LDS staging area; n: its fill (wave-uniform register)
__device__ inline void near_flush(int2 *entries, int &n, int2 *list, int *count, int cap) {
    if (!n) return;
    int base = 0;
    if (__lane_id() == 0) base = atomicAdd(count, n);
    base = __shfl(base, 0);
    for (int k = __lane_id(); k < n; k += 64)
        if (base + k < cap) list[base + k] = entries[k];
    n = 0;
}

__device__ inline void near_stage(int2 *entries, int &n, bool pred, int2 *list, int *count, int cap, int row, int j) {
    const uint64_t m = __ballot(pred);
    if (!m) return;
    const int cnt = __popcll(m);
    if (n + cnt > kNearStage) near_flush(entries, n, list, count, cap);
    if (pred) entries[n + __popcll(m & ((1ull << __lane_id()) - 1ull))] = make_int2(row, j);
    n += cnt;
}

// ---------------------------------------------------------------------------
// S build: a wave = 64 consecutive rows of S (two 32-row tiles), one lane per
// row, all 8 emitters of each K step. The step's emitter records and
// visibility words are then the same for every lane: the wave stages them in
// LDS once and reads each with a broadcast. Each lane stores its row's two
// fragments (emitters 0-3, 4-7) into the MFMA A layout of its tile. Near pairs
// (modes 1, 2) are left zero and listed for the near-field kernels.

struct EmRec {
    float4 p, n;      // point (area), normal (facet id bits)
    uint64_t w0, w1;  // visibility words of the wave's two receiver blocks (fused build: w0 the plane bits)
};
constexpr int kBuildEm = 16 * 8;  // emitters per wave: at most 16 K steps

// Difference vector, squared length and the two cosine dot products of a pair.
// Exchanging receiver and emitter negates v exactly, which leaves r2 as it is
// and turns de and dr into each other's exact negations (fmaf(-a, b, -c) is
// -fmaf(a, b, c)).
struct PairGeom {
    float vx, vy, vz, r2, de, dr;
};
__device__ __forceinline__ PairGeom pair_geom(const float4 &rp, const float4 &rn, const float4 &ep,
                                              const float4 &en) {
    PairGeom q;
    q.vx = rp.x - ep.x;
    q.vy = rp.y - ep.y;
    q.vz = rp.z - ep.z;
    q.r2 = fmaf(q.vz, q.vz, fmaf(q.vy, q.vy, q.vx * q.vx));
    q.de = fmaf(q.vz, en.z, fmaf(q.vy, en.y, q.vx * en.x));
    q.dr = fmaf(q.vz, rn.z, fmaf(q.vy, rn.y, q.vx * rn.x));
    return q;
}

// The FP16 word of the pair (vis: visible), and whether it is a near pair (its
// word left zero, listed for the near-field kernels) and which rule takes it.
// Every lane evaluates the far-field form and selects (the excluded cases'
// values, even non-finite ones at r = 0, are never selected). Symmetric in the
// pair's two ends (pair_geom; de dr is one product): the word of (j, i) is the
// word of (i, j), bit for bit.
template <bool ATT, int NEAR>
__device__ __forceinline__ uint32_t pair_value(const TransferPhys &ph, const PairGeom &q, bool vis, bool &isNear,
                                               bool &isQuad) {
    const float rinv = __builtin_amdgcn_rsqf(q.r2);
    const float r = q.r2 * rinv;
    const float rinv2 = rinv * rinv;
    const float wgt = -(q.de * q.dr) * rinv2 * ph.damp * ph.scale;
    // FP64 reduction of k r to one revolution, then FP32 sin/cos.
    const double rev = (double)r * ph.krevD;
    const float x = (float)(rev - rint(rev));
    const float c = __builtin_amdgcn_cosf(x), sn = __builtin_amdgcn_sinf(x);
    float a = rinv * kInv4PiT;
    if (ATT) a *= __builtin_amdgcn_exp2f(-ph.alphaLog2e * r);
    float Gre, Gim;
    if (ATT) {
        Gre = fmaf(a, fmaf(ph.alpha, c, ph.kr * sn), rinv2);
        Gim = -a * fmaf(ph.kr, c, -ph.alpha * sn);
    } else {
        Gre = fmaf(a * ph.kr, sn, rinv2);
        Gim = -(a * ph.kr) * c;
    }
    // near pairs: evaluated afterwards from the near lists (k_near)
    isNear = NEAR != 0 && vis && r < ph.nearRadius;
    // the learned range starts at r >= delta
    isQuad = isNear && (NEAR == 1 || r < ph.quadRadius);
    const bool reg = r < ph.deltaReg;
    float re = reg ? wgt * (-ph.alpha * kInv4PiT) : wgt * Gre;
    float im = reg ? wgt * (ph.kr * kInv4PiT) : wgt * Gim;
    const bool keep = vis && !isNear;
    re = keep ? re : 0.f;
    im = keep ? im : 0.f;
    return pack_h2(re, im);
}

template <bool ATT, int NEAR>
__global__ __launch_bounds__(256) void k_build_s(TransferGeom g, TransferPhys ph, const uint64_t *__restrict__ mask,
                                                 uint4 *__restrict__ A, int ksPerWave) {
    __shared__ int2 stage[NEAR != 0 ? 4 : 1][2][NEAR != 0 ? kNearStage : 1];
    __shared__ EmRec sEmAll[4][kBuildEm];
    int2 *stQ = stage[NEAR != 0 ? (threadIdx.x >> 6) : 0][0], *stM = stage[NEAR != 0 ? (threadIdx.x >> 6) : 0][1];
    int nQ = 0, nM = 0;
    // wave-uniform by construction; readfirstlane lets the compiler see it
    const int wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    const int nKsChunks = (g.nKs + ksPerWave - 1) / ksPerWave;
    const int rp64 = wave / nKsChunks;  // 64-row block of S
    const int kc = wave % nKsChunks;
    if (rp64 >= g.nRt / 2) return;  // uniform (nRt is a multiple of 4)
    const int srow = rp64 * 64 + lane;  // row of S
    const int row = g.rowAct ? g.rowAct[srow] : srow;  // its active cell (-1: padding)
    const bool rvalid = row >= 0 && row < g.n;
    const float4 rp = rvalid ? g.p[row] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 rn = rvalid ? g.nrm[row] : make_float4(0.f, 0.f, 0.f, 0.f);
    // The wave's rows are consecutive active cells (one facet's, or the identity
    // map), so their visibility words come from two 64-row blocks at most: b0
    // (lane 0's, or the first valid lane's) and b0 + 1.
    const int ks0 = kc * ksPerWave, ks1 = min(g.nKs, ks0 + ksPerWave);
    const int2 skip = g.wgSkip ? g.wgSkip[rp64 >> 1] : make_int2(0, 0);  // never read by the GEMM
    const int tile = srow >> 5, tl = srow & 31;
    const uint64_t valid = __ballot(rvalid);
    if (!valid) {  // padding rows only: zero fragments
        for (int ks = ks0; ks < ks1; ks++) {
            if (ks >= skip.x && ks < skip.y) continue;
            uint4 *dst = A + ((size_t)tile * g.nKs + ks) * 64 + tl;
            dst[0] = make_uint4(0u, 0u, 0u, 0u);
            dst[32] = make_uint4(0u, 0u, 0u, 0u);
        }
        return;
    }
    const int first = __builtin_amdgcn_readfirstlane(__ffsll((long long)valid) - 1);
    const int b0 = __builtin_amdgcn_readfirstlane(__shfl(row, first) >> 6);
    const int nrbM = (g.n + 63) >> 6;
    const int b1 = min(b0 + 1, nrbM - 1);
    const bool upper = rvalid && (row >> 6) != b0;  // this lane's words are block b0 + 1's
    const int bit = row & 63;
    // This wave's emitters (at most kBuildEm: its K steps' 8 each) staged in LDS
    // with coalesced loads, one per lane; every lane then reads each emitter's
    // record at one address (a broadcast). An emitter past the last is the last
    // one again (never selected). The wave alone writes and then reads its slots,
    // and its LDS accesses run in order.
    EmRec *sEm = sEmAll[threadIdx.x >> 6];
    for (int u = lane; u < (ks1 - ks0) * 8; u += 64) {
        const int jj = emitter_at(g, ks0 * 8 + u);
        const int jc = jj >= 0 ? jj : 0;
        sEm[u].p = g.p[jc];
        sEm[u].n = g.nrm[jc];
        // a padding position: no visibility (never selected)
        sEm[u].w0 = jj >= 0 ? mask[(size_t)b0 * g.n + jj] : 0ull;
        sEm[u].w1 = jj >= 0 ? mask[(size_t)b1 * g.n + jj] : 0ull;
    }
    for (int ks = ks0; ks < ks1; ks++) {
        if (ks >= skip.x && ks < skip.y) continue;  // uniform
        uint32_t w[8];
        unsigned nearBits = 0, quadBits = 0;  // per emitter: listed near pair, by quadrature
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int jl = ks * 8 + q - ks0 * 8;  // the emitter's slot in this wave's staged records
            const uint64_t word = upper ? sEm[jl].w1 : sEm[jl].w0;
            bool isNear, isQuad;
            w[q] = pair_value<ATT, NEAR>(ph, pair_geom(rp, rn, sEm[jl].p, sEm[jl].n),
                                         rvalid && ((word >> bit) & 1ull), isNear, isQuad);
            nearBits |= (unsigned)isNear << q;
            quadBits |= (unsigned)isQuad << q;
        }
        // near pairs are rare: one ballot decides whether this step lists any
        if (NEAR != 0 && __ballot(nearBits != 0u)) {
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int j = emitter_at(g, ks * 8 + q);  // listed as active cells
                const bool nr = (nearBits >> q) & 1u, qd = (quadBits >> q) & 1u;
                near_stage(stQ, nQ, nr && qd, ph.nearQ, ph.nearCount, ph.nearCap, row, j);
                near_stage(stM, nM, nr && !qd, ph.nearM, ph.nearCount + 1, ph.nearCap, row, j);
            }
        }
        // fragment h of row tl of the tile: emitters 4h .. 4h + 3
        uint4 *dst = A + ((size_t)tile * g.nKs + ks) * 64 + tl;
        if (ph.streamS) {
            store_nt(dst, make_uint4(w[0], w[1], w[2], w[3]));
            store_nt(dst + 32, make_uint4(w[4], w[5], w[6], w[7]));
        } else {
            dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
            dst[32] = make_uint4(w[4], w[5], w[6], w[7]);
        }
    }
    if (NEAR != 0) {
        near_flush(stQ, nQ, ph.nearQ, ph.nearCount, ph.nearCap);
        near_flush(stM, nM, ph.nearM, ph.nearCount + 1, ph.nearCap);
    }
}

// ---------------------------------------------------------------------------
// Visibility of a SELF pair inside the symmetric build (FUSED): the test the
// visibility kernels apply to an object's own fragment pairs (vis.hip,
// ORIGIN_LOWER_FACET, both cosine culls, at most kPreTris occluders), on the
// same FP32 arithmetic, so each bit is the visibility word's bit. Per 64 x 64
// block (rows on one facet, emitters on another), before the block's pairs are
// evaluated:
//   Culls (block_culls): with one normal per side, the rows that lie well in
//   front of the emitters' plane keep every pair by the emitter's cull, those
//   behind it cull every pair, and likewise the emitters against the rows'
//   plane; the rows and emitters in between are special.
//   Occluders (special_points): from the plane bits of the rows and emitters
//   (plane_bits), the rows and emitters a candidate triangle (pair_candidates)
//   can involve are special too; no other pair has one.
//   The pairs of the special rows and emitters are then decided exactly, one
//   row (or emitter) at a time across the wave's 64 lanes (special_vis): both
//   culls by cull_margin or, near the threshold, the visibility kernels' exact
//   set-up and comparisons; candidate triangles by the exact ray and the FP32
//   triangle test (mt_hit_dev, the ray leaving the lower facet's end). Their
//   bits go to LDS; every other pair's bit is known from its row and emitter.

__device__ __forceinline__ bool pair_culls(const PairGeom &q, const float4 &rp, const float4 &rn, const float4 &ep,
                                           const float4 &en, bool pairOk) {
    const int cE = cull_margin(q.de, q.r2), cR = cull_margin(-q.dr, q.r2);
    bool kept = pairOk && cE == 1 && cR == 1;
    const bool undecided = pairOk && cE != 0 && cR != 0 && (cE == 2 || cR == 2);
    if (__ballot(undecided) != 0ull) {
        if (undecided) {
            F3 uv;
            float rr;
            ray_setup_dev(F3{ep.x, ep.y, ep.z}, F3{rp.x, rp.y, rp.z}, uv, rr);
            kept = f3dot(uv, F3{en.x, en.y, en.z}) >= 1e-6f &&
                   f3dot(F3{-uv.x, -uv.y, -uv.z}, F3{rn.x, rn.y, rn.z}) >= 1e-6f;
        }
    }
    return kept;
}

// The shadow test of the pair (receiver rp, emitter ep) against its candidate
// triangles cand (staged records tris), the ray leaving the emitter end when
// fromEm, else the receiver end.
__device__ inline bool pair_occluded(unsigned cand, const float4 *tris, const SelfVis &sv, const float4 &ep,
                                     const float4 &rp, bool fromEm) {
    const F3 ep3{ep.x, ep.y, ep.z}, rp3{rp.x, rp.y, rp.z};
    F3 uv;
    float rr;
    ray_setup_dev(ep3, rp3, uv, rr);
    const F3 o = fromEm ? ep3 : rp3;
    const F3 d = fromEm ? uv : F3{-uv.x, -uv.y, -uv.z};
    bool hit = false;
#pragma unroll
    for (int ti = 0; ti < kPreTris; ti++) {
        if (ti < sv.ntri && ((cand >> ti) & 1u) && !hit) {
            const float4 *t4 = tris + 4 * ti;
            hit = mt_hit_dev(tri_unpack(t4[0], t4[1], t4[2], t4[3]), sv.padTri, o, d, sv.tmin, rr - 1e-3f);
        }
    }
    return hit;
}

// The whole visibility bit of one pair, exactly (the special pairs).
__device__ inline bool pair_vis_exact(const SelfVis &sv, const float4 *tris, const float4 &rp, const float4 &rn,
                                      unsigned bR, bool rv, const float4 &ep, const float4 &en, unsigned bE, bool ev) {
    const int rf = __float_as_int(rn.w), ef = __float_as_int(en.w);
    const bool pairOk = rv && ev && ef != rf;
    bool vis = pair_culls(pair_geom(rp, rn, ep, en), rp, rn, ep, en, pairOk);
    const unsigned cand = vis ? pair_candidates(bE, bR, sv.ntri) : 0u;
    if (__ballot(cand != 0u) != 0ull) {
        bool hit = false;
        if (cand != 0u) hit = pair_occluded(cand, tris, sv, ep, rp, ef < rf);
        vis = vis && !hit;
    }
    return vis;
}

// The block's points that can have a candidate triangle: per triangle, the side
// (beyond the margin) that more of the block's points not lying on it are on;
// a point is special when, for some triangle it does not lie on, it is not
// beyond the margin on that side. Two points that are not special have no
// candidate triangle (each triangle is one's own, or both are beyond the margin
// on its majority side), so every pair with one has a special end.
__device__ inline void special_points(unsigned bR, bool rv, unsigned bE, bool ev, int ntri, bool &specR,
                                      bool &specE) {
    specR = specE = false;
    for (int ti = 0; ti < ntri; ti++) {
        const bool nownR = rv && !((bR >> (ti + 8)) & 1u), nownE = ev && !((bE >> (ti + 8)) & 1u);
        const int nAbove = __popcll(__ballot(nownR && ((bR >> ti) & 1u))) + __popcll(__ballot(nownE && ((bE >> ti) & 1u)));
        const int nBelow = __popcll(__ballot(nownR && ((bR >> (ti + 4)) & 1u))) +
                           __popcll(__ballot(nownE && ((bE >> (ti + 4)) & 1u)));
        const int side = nAbove >= nBelow ? ti : ti + 4;
        specR = specR || (nownR && !((bR >> side) & 1u));
        specE = specE || (nownE && !((bE >> side) & 1u));
    }
}

__device__ inline float wave_min(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// The cosine culls of a block by row and by emitter, when all its rows share one
// normal nR and all its emitters another, nE. With a = nE . rp per row and
// b = nE . ep per emitter (FP32, each within E = 4e-7 scale of the exact dot
// product: |p| <= sqrt(3) scale), a pair's nE . (rp - ep) lies within
// [a - max b - 2E, a - min b + 2E], and |rp - ep| <= D (the largest distances
// of the rows and of the emitters from one row point, added). A row whose lower
// end exceeds 3.5e-6 D keeps every pair by the emitter's cull (every exact
// cosine, its set-up within 7.6 u of the true one, cull_margin, is above
// 2.9e-6); a row whose upper end is below zero culls every pair. Likewise each
// emitter with nR . (ep - rp) for the receiver's cull. Per lane: rowKept /
// rowCulled for its row, emKept / emCulled for its staged emitter; with
// differing normals every valid point is neither (special).
__device__ inline void block_culls(const float4 &rp, const float4 &rn, bool rv, const float4 &ep, const float4 &en,
                                   bool ev, float scale, bool &rowKept, bool &rowCulled, bool &emKept,
                                   bool &emCulled) {
    rowKept = rowCulled = emKept = emCulled = false;
    const int fr = __builtin_amdgcn_readfirstlane(__ffsll((long long)__ballot(rv)) - 1);
    const int fe = __builtin_amdgcn_readfirstlane(__ffsll((long long)__ballot(ev)) - 1);
    const F3 nR{__shfl(rn.x, fr), __shfl(rn.y, fr), __shfl(rn.z, fr)};
    const F3 nE{__shfl(en.x, fe), __shfl(en.y, fe), __shfl(en.z, fe)};
    const bool sameR = !rv || (rn.x == nR.x && rn.y == nR.y && rn.z == nR.z);
    const bool sameE = !ev || (en.x == nE.x && en.y == nE.y && en.z == nE.z);
    if (__ballot(!(sameR && sameE)) != 0ull) return;
    const F3 r3{rp.x, rp.y, rp.z}, e3{ep.x, ep.y, ep.z};
    const F3 p0{__shfl(rp.x, fr), __shfl(rp.y, fr), __shfl(rp.z, fr)};
    const float inf = __builtin_inff();
    const float dR = rv ? sqrtf(f3dot(f3sub(r3, p0), f3sub(r3, p0))) : 0.f;
    const float dE = ev ? sqrtf(f3dot(f3sub(e3, p0), f3sub(e3, p0))) : 0.f;
    const float D = (wave_max(dR) + wave_max(dE)) * 1.0001f;
    const float E2 = 8e-7f * scale, thr = 3.5e-6f * D;
    const float a = f3dot(nE, r3), b = f3dot(nE, e3);  // emitter's cull: nE . (rp - ep)
    const float c = f3dot(nR, e3), d = f3dot(nR, r3);  // receiver's cull: nR . (ep - rp)
    const float minB = wave_min(ev ? b : inf), maxB = wave_max(ev ? b : -inf);
    const float minD = wave_min(rv ? d : inf), maxD = wave_max(rv ? d : -inf);
    rowCulled = rv && (a - minB) + E2 < 0.f;
    rowKept = rv && (a - maxB) - E2 > thr;
    emCulled = ev && (c - minD) + E2 < 0.f;
    emKept = ev && (c - maxD) - E2 > thr;
}

// The special pairs' bits: lane e's word, bit r = visibility of (row r, staged
// emitter e), for every row r in specRows and every pair of the emitters in
// specEms (other bits zero). Lanes hold their row (rp, rn, bR, rv) and their
// staged emitter (ep, en, bE, ev); sEm holds the staged emitters' records.
__device__ inline uint64_t special_vis(const SelfVis &sv, const float4 *tris, const EmRec *sEm, uint64_t specRows,
                                       uint64_t specEms, const float4 &rp, const float4 &rn, unsigned bR, bool rv,
                                       const float4 &ep, const float4 &en, unsigned bE, bool ev) {
    const int lane = __lane_id();
    uint64_t vis = 0;
    // the special rows: lane e takes the pair (row r, emitter e)
    for (uint64_t m = specRows; m; m &= m - 1) {
        const int r = __builtin_amdgcn_readfirstlane(__ffsll((long long)m) - 1);
        const float4 rpR = make_float4(__shfl(rp.x, r), __shfl(rp.y, r), __shfl(rp.z, r), 0.f);
        const float4 rnR = make_float4(__shfl(rn.x, r), __shfl(rn.y, r), __shfl(rn.z, r), __shfl(rn.w, r));
        const bool v = pair_vis_exact(sv, tris, rpR, rnR, __shfl(bR, r), true, ep, en, bE, ev);
        vis |= (uint64_t)v << r;
    }
    // the special emitters: lane r takes the pair (row r, emitter e)
    for (uint64_t m = specEms; m; m &= m - 1) {
        const int e = __builtin_amdgcn_readfirstlane(__ffsll((long long)m) - 1);
        const bool v = pair_vis_exact(sv, tris, rp, rn, bR, rv, sEm[e].p, sEm[e].n, (unsigned)sEm[e].w0, true);
        const uint64_t vm = __ballot(v);
        if (lane == e) vis |= vm;
    }
    return vis;
}

// Symmetric S build (rows and emitters padded alike, no near pairs listed). S is
// symmetric bit for bit (pair_value; the visibility words are symmetric), so a
// wave evaluates one 64 x 64 block, rows of block rb by K chunk cb with rb <= cb,
// stores it, and stores its transpose as block (cb, rb): after each K step the
// lanes' 8 words go through the wave's LDS tile and come back as transposed rows
// (16-byte pieces of 4 consecutive K positions; 8 lanes store one row's 128
// contiguous bytes of a K step). Items: the (rb, cb) pairs whose chunk is not
// skipped, listed on the host. A block with no valid row or emitter is stored
// as zeros without evaluating it. FUSED: the visibility bits are evaluated here
// (block_culls, special_points, special_vis) instead of read from the SELF
// visibility words.
constexpr int kSymT = 72;  // LDS tile row stride in words: 8 transposed rows of 64 K positions, padded
// Five waves per SIMD: the FUSED build's visibility code needs 104 registers per
// lane (four waves); capped at 96 it spills 32 bytes per lane and still builds
// the trihedral's S faster, 0.195 ms against 0.203 (mean of three runs each).
template <bool ATT, bool FUSED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_build_s_sym(TransferGeom g, TransferPhys ph,
                                                     const uint64_t *__restrict__ mask, uint4 *__restrict__ A,
                                                     const int2 *__restrict__ items, int nItems, SelfVis sv) {
    __shared__ EmRec sEmAll[4][64];
    __shared__ uint4 sTAll[4][8 * kSymT / 4];
    __shared__ float4 sTri[FUSED ? 4 * kPreTris : 1];
    __shared__ uint64_t sVisAll[4][FUSED ? 64 : 1];
    if (FUSED) {
        for (int q = threadIdx.x; q < 4 * sv.ntri; q += 256) sTri[q] = sv.tris[q];
        __syncthreads();
    }
    const int item = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (item >= nItems) return;  // uniform
    const int lane = threadIdx.x & 63;
    const int2 bc = items[item];
    const int rb = __builtin_amdgcn_readfirstlane(bc.x), cb = __builtin_amdgcn_readfirstlane(bc.y);
    const int srow = rb * 64 + lane;
    const int row = g.rowAct[srow];  // this lane's receiver (-1: padding)
    const int em = g.rowAct[cb * 64 + lane];  // this lane's staged emitter (-1: padding)
    const bool rvalid = row >= 0;
    const float4 rp = rvalid ? g.p[row] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 rn = rvalid ? g.nrm[row] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint64_t valid = __ballot(rvalid);
    const bool live = valid != 0ull && __ballot(em >= 0) != 0ull;  // uniform
    // The rows are consecutive active cells of one facet: their visibility words
    // come from two 64-row blocks at most (as in k_build_s).
    const int first = live ? __builtin_amdgcn_readfirstlane(__ffsll((long long)valid) - 1) : 0;
    const int b0 = live ? __builtin_amdgcn_readfirstlane(__shfl(row, first) >> 6) : 0;
    const int b1 = min(b0 + 1, ((g.n + 63) >> 6) - 1);
    const bool upper = rvalid && (row >> 6) != b0;
    const int bit = row & 63;
    EmRec *sEm = sEmAll[threadIdx.x >> 6];
    uint32_t *sT = reinterpret_cast<uint32_t *>(sTAll[threadIdx.x >> 6]);
    uint64_t *sVis = sVisAll[threadIdx.x >> 6];
    // FUSED: this lane's row is special or keeps its non-special pairs; per
    // staged emitter (uniform masks): special, keeps its non-special pairs
    bool rowSpec = false, rowOk = false;
    uint64_t emSpec = 0, emOk = 0;
    if (live) {
        const int ec = em >= 0 ? em : 0;
        const float4 ep = g.p[ec], en = g.nrm[ec];
        sEm[lane].p = ep;
        sEm[lane].n = en;
        if (FUSED) {
            const bool ev = em >= 0;
            const unsigned bE = plane_bits(sTri, sv.ntri, sv.scale, F3{ep.x, ep.y, ep.z},
                                           sv.triBase + __float_as_int(en.w));
            const unsigned bR = plane_bits(sTri, sv.ntri, sv.scale, F3{rp.x, rp.y, rp.z},
                                           sv.triBase + __float_as_int(rn.w));
            sEm[lane].w0 = bE;
            __builtin_amdgcn_wave_barrier();  // the staged records precede the reads of them
            bool rowKept, rowCulled, emKept, emCulled, occR, occE;
            block_culls(rp, rn, rvalid, ep, en, ev, sv.scale, rowKept, rowCulled, emKept, emCulled);
            special_points(bR, rvalid, bE, ev, sv.ntri, occR, occE);
            // a pair with neither end special: both culls decided by its row and its
            // emitter, no candidate triangle, and the two ends on different facets
            // (the rows on one facet, the emitters on another: else every row is special)
            const int rf = __float_as_int(rn.w), ef = __float_as_int(en.w);
            const int rf0 = __shfl(rf, first);
            const int ef0 = __shfl(ef, __builtin_amdgcn_readfirstlane(__ffsll((long long)__ballot(ev)) - 1));
            const bool twoFacets = rf0 != ef0 && __ballot(rvalid && rf != rf0) == 0ull &&
                                   __ballot(ev && ef != ef0) == 0ull;
            rowSpec = rvalid && (!twoFacets || (!rowCulled && (!rowKept || occR)));
            rowOk = rvalid && rowKept && twoFacets;
            emSpec = __ballot(ev && !emCulled && (!emKept || occE));
            emOk = __ballot(ev && emKept);
            const uint64_t rowsM = __ballot(rowSpec);
            sVis[lane] = (rowsM | emSpec) ? special_vis(sv, sTri, sEm, rowsM, emSpec, rp, rn, bR, rvalid, ep, en, bE, ev)
                                          : 0ull;
        } else {
            sEm[lane].w0 = em >= 0 ? mask[(size_t)b0 * g.n + em] : 0ull;
            sEm[lane].w1 = em >= 0 ? mask[(size_t)b1 * g.n + em] : 0ull;
        }
    }
    const bool mirror = rb != cb;
    // this lane's two transposed pieces per K step: row q of the step's 8, K positions
    // 8 s2 + 4 h .. + 3 of block rb
    const int q0 = lane & 7, h0 = (lane >> 3) & 1, s0 = lane >> 4;  // piece lane; the other is s0 + 4
    for (int ks = 0; ks < 8; ks++) {
        uint32_t w[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            w[q] = 0u;
            if (live) {
                const int jl = ks * 8 + q;
                const float4 ep = sEm[jl].p, en = sEm[jl].n;
                bool vis;
                if (FUSED) {
                    const bool spec = rowSpec || ((emSpec >> jl) & 1ull);
                    vis = spec ? ((sVis[jl] >> lane) & 1ull) : (rowOk && ((emOk >> jl) & 1ull));
                } else {
                    const uint64_t word = upper ? sEm[jl].w1 : sEm[jl].w0;
                    vis = rvalid && ((word >> bit) & 1ull);
                }
                bool isNear, isQuad;
                w[q] = pair_value<ATT, 0>(ph, pair_geom(rp, rn, ep, en), vis, isNear, isQuad);
            }
        }
        uint4 *dst = A + ((size_t)(srow >> 5) * g.nKs + cb * 8 + ks) * 64 + (srow & 31);
        const uint4 lo = make_uint4(w[0], w[1], w[2], w[3]), hi = make_uint4(w[4], w[5], w[6], w[7]);
        if (ph.streamS) {
            store_nt(dst, lo);
            store_nt(dst + 32, hi);
        } else {
            dst[0] = lo;
            dst[32] = hi;
        }
        if (!mirror) continue;  // uniform
        // The wave's LDS accesses run in order: the last step's reads of the tile
        // precede these writes, which precede the reads below.
#pragma unroll
        for (int q = 0; q < 8; q++) sT[q * kSymT + lane] = w[q];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int s2 = s0 + 4 * u;
            const uint4 v = *reinterpret_cast<const uint4 *>(sT + q0 * kSymT + s2 * 8 + h0 * 4);
            const int trow = cb * 64 + ks * 8 + q0;  // row of S: the emitter
            uint4 *td = A + ((size_t)(trow >> 5) * g.nKs + rb * 8 + s2) * 64 + (trow & 31) + 32 * h0;
            if (ph.streamS) store_nt(td, v);
            else *td = v;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------------------

int transfer_build_s(const TransferGeom &g, const TransferPhys &ph, const uint64_t *mask, uint4 *A, hipStream_t st) {
    // runs of 16 K steps per wave (64 rows each), shorter when that leaves fewer than ~8k waves
    if (g.nRt % 4) { set_error("S build: row tiles must be padded to 4"); return -2; }
    int ksPerWave = kBuildEm / 8;  // the kernel stages that many K steps' emitters per wave
    while (ksPerWave > 4 && (long)(g.nRt / 2) * ((g.nKs + ksPerWave - 1) / ksPerWave) < 8192) ksPerWave /= 2;
    const long waves = (long)(g.nRt / 2) * ((g.nKs + ksPerWave - 1) / ksPerWave);
    const long blocks = (waves + 3) / 4;
    if (blocks <= 0) return 0;
    const bool att = ph.alpha != 0.0f;
    if (ph.nearMode == 1) {
        if (att) k_build_s<true, 1><<<(unsigned)blocks, 256, 0, st>>>(g, ph, mask, A, ksPerWave);
        else k_build_s<false, 1><<<(unsigned)blocks, 256, 0, st>>>(g, ph, mask, A, ksPerWave);
    } else if (ph.nearMode == 2) {
        if (att) k_build_s<true, 2><<<(unsigned)blocks, 256, 0, st>>>(g, ph, mask, A, ksPerWave);
        else k_build_s<false, 2><<<(unsigned)blocks, 256, 0, st>>>(g, ph, mask, A, ksPerWave);
    } else if (att) {
        k_build_s<true, 0><<<(unsigned)blocks, 256, 0, st>>>(g, ph, mask, A, ksPerWave);
    } else {
        k_build_s<false, 0><<<(unsigned)blocks, 256, 0, st>>>(g, ph, mask, A, ksPerWave);
    }
    SR_HIP(hipGetLastError());
    return 0;
}

int transfer_build_s_sym(const TransferGeom &g, const TransferPhys &ph, const uint64_t *mask, uint4 *A,
                         const int2 *items, int nItems, hipStream_t st, const SelfVis *sv) {
    if (!g.rowAct || ph.nearMode != 0) {
        set_error("symmetric S build: needs rows and emitters padded alike and no near pairs");
        return -2;
    }
    if (sv && (sv->ntri < 1 || sv->ntri > kPreTris || !sv->tris)) {
        set_error("symmetric S build with its own visibility: 1 to 4 occluder triangles");
        return -2;
    }
    if (!sv && !mask) { set_error("symmetric S build: no visibility words"); return -2; }
    if (nItems <= 0) return 0;
    const unsigned blocks = (unsigned)((nItems + 3) / 4);
    const SelfVis none{};
    const bool att = ph.alpha != 0.0f;
    if (sv) {
        if (att) k_build_s_sym<true, true><<<blocks, 256, 0, st>>>(g, ph, nullptr, A, items, nItems, *sv);
        else k_build_s_sym<false, true><<<blocks, 256, 0, st>>>(g, ph, nullptr, A, items, nItems, *sv);
    } else {
        if (att) k_build_s_sym<true, false><<<blocks, 256, 0, st>>>(g, ph, mask, A, items, nItems, none);
        else k_build_s_sym<false, false><<<blocks, 256, 0, st>>>(g, ph, mask, A, items, nItems, none);
    }
    SR_HIP(hipGetLastError());
    return 0;
}

}  // namespace sr
