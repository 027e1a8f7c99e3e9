// Shadow-ray visibility on gfx950: LBVH build and wavefront-cooperative
// any-hit traversal that writes one 64-bit visibility word per
// (receiver block of 64, emitter).
//
// Replaces the reference's OptiX GAS build (src/OptiX.cpp:304-350) and its
// optixTrace call sites (src/OptiX.cu:81-93,163-175,321-333). The cosine culls
// of each call site (OptiX.cu:67-78,155-160,297-318) are folded into the same
// word, so a set bit means "this pair contributes".
//
// Compiled with -ffp-contract=off: the ray set-up, culls and the triangle test
// must round exactly as the host oracle does.

#include <hip/hip_runtime.h>
#include <cstdio>
#include <utility>
#include <vector>

#include "sr_types.h"
#include "vis_fp32.h"
#include "launch.h"

namespace sr {

// ---------------------------------------------------------------------------
// LBVH build (Karras 2012: Morton order, radix sort, hierarchy from key
// prefixes, bottom-up box refit).

__device__ inline unsigned expand_bits10(unsigned v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ void k_morton(const float4 *__restrict__ tris, int n, float3 lo, float3 invExt,
                         unsigned *__restrict__ keys, int *__restrict__ ids) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    TriRec t = tri_unpack(tris[4 * i], tris[4 * i + 1], tris[4 * i + 2], tris[4 * i + 3]);
    float cx = (t.bmin.x + t.bmax.x) * 0.5f, cy = (t.bmin.y + t.bmax.y) * 0.5f, cz = (t.bmin.z + t.bmax.z) * 0.5f;
    float fx = fminf(fmaxf((cx - lo.x) * invExt.x, 0.0f), 1.0f);
    float fy = fminf(fmaxf((cy - lo.y) * invExt.y, 0.0f), 1.0f);
    float fz = fminf(fmaxf((cz - lo.z) * invExt.z, 0.0f), 1.0f);
    unsigned ux = (unsigned)fminf(fx * 1024.0f, 1023.0f);
    unsigned uy = (unsigned)fminf(fy * 1024.0f, 1023.0f);
    unsigned uz = (unsigned)fminf(fz * 1024.0f, 1023.0f);
    keys[i] = (expand_bits10(ux) << 2) | (expand_bits10(uy) << 1) | expand_bits10(uz);
    ids[i] = i;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (Morton key, triangle index) pairs, 8 bits a pass.
// A tile of kSortTile keys belongs to one wave. Per pass: every tile counts its
// digits, one workgroup turns the [digit][tile] counts into output offsets, and
// every tile scatters its keys in order: a key's rank among equal digits in its
// 64-key step comes from eight ballots (the lanes with the same digit), plus the
// count of that digit in the tile's earlier steps, which keeps the sort stable.
constexpr int kSortTile = 2048;
constexpr int kSortWaves = 4;

__device__ inline unsigned sort_digit(unsigned key, int shift) { return (key >> shift) & 0xffu; }

__global__ __launch_bounds__(64 * kSortWaves) void k_sort_hist(const unsigned *__restrict__ keys, int n, int shift,
                                                              int tiles, unsigned *__restrict__ hist) {
    __shared__ unsigned cnt[kSortWaves][256];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int d = lane; d < 256; d += 64) cnt[wave][d] = 0u;
    __syncthreads();
    const int tile = blockIdx.x * kSortWaves + wave;
    if (tile < tiles) {
        const int i0 = tile * kSortTile, i1 = min(n, i0 + kSortTile);
        for (int i = i0 + lane; i < i1; i += 64) atomicAdd(&cnt[wave][sort_digit(keys[i], shift)], 1u);
    }
    __syncthreads();
    if (tile < tiles)
        for (int d = lane; d < 256; d += 64) hist[(size_t)d * tiles + tile] = cnt[wave][d];
}

// Exclusive prefix sum over the [digit][tile] counts (one workgroup of 1024).
__global__ __launch_bounds__(1024) void k_sort_scan(unsigned *__restrict__ hist, int total) {
    __shared__ unsigned part[1024];
    const int t = threadIdx.x;
    const int per = (total + 1023) / 1024;
    const int a = min(total, t * per), b = min(total, a + per);
    unsigned sum = 0u;
    for (int i = a; i < b; i++) sum += hist[i];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan of the segment sums
        const unsigned v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    unsigned run = t > 0 ? part[t - 1] : 0u;
    for (int i = a; i < b; i++) {
        const unsigned c = hist[i];
        hist[i] = run;
        run += c;
    }
}

__global__ __launch_bounds__(64 * kSortWaves) void k_sort_scatter(const unsigned *__restrict__ keys,
                                                                 const int *__restrict__ vals, int n, int shift,
                                                                 int tiles, const unsigned *__restrict__ offs,
                                                                 unsigned *__restrict__ keysOut,
                                                                 int *__restrict__ valsOut) {
    __shared__ unsigned base[kSortWaves][256];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tile = blockIdx.x * kSortWaves + wave;
    if (tile >= tiles) return;  // uniform over the wave; no barrier below
    for (int d = lane; d < 256; d += 64) base[wave][d] = offs[(size_t)d * tiles + tile];
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    const int i0 = tile * kSortTile, i1 = min(n, i0 + kSortTile);
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int c = i0; c < i1; c += 64) {
        const int i = c + lane;
        const bool ok = i < i1;
        const unsigned key = ok ? keys[i] : 0u;
        const unsigned d = sort_digit(key, shift);
        unsigned long long same = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const unsigned long long m = __ballot(ok && ((d >> b) & 1u));
            same &= ((d >> b) & 1u) ? m : ~m;
        }
        const unsigned pos = ok ? base[wave][d] + (unsigned)__popcll(same & below) : 0u;
        if (ok) {
            keysOut[pos] = key;
            valsOut[pos] = vals[i];
        }
        __builtin_amdgcn_wave_barrier();
        // the last lane of each digit group advances that digit's base for the next step
        if (ok && (same >> lane) == 1ull) base[wave][d] += (unsigned)__popcll(same);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    }
}

// Sorts keys/vals (n pairs, keys below 2^bits) with ping-pong buffers; the
// sorted pairs end in keys/vals (an even number of passes). hist: 256 x tiles words.
int radix_sort_pairs(unsigned *keys, int *vals, unsigned *keys2, int *vals2, unsigned *hist, int n, int bits,
                     hipStream_t st) {
    if (n <= 1) return 0;
    const int tiles = (n + kSortTile - 1) / kSortTile;
    const int blocks = (tiles + kSortWaves - 1) / kSortWaves;
    int passes = (bits + 7) / 8;
    if (passes & 1) passes++;  // the result lands back in keys/vals
    unsigned *kin = keys, *kout = keys2;
    int *vin = vals, *vout = vals2;
    for (int p = 0; p < passes; p++) {
        const int shift = 8 * p;
        k_sort_hist<<<blocks, 64 * kSortWaves, 0, st>>>(kin, n, shift, tiles, hist);
        SR_HIP(hipGetLastError());
        k_sort_scan<<<1, 1024, 0, st>>>(hist, 256 * tiles);
        SR_HIP(hipGetLastError());
        k_sort_scatter<<<blocks, 64 * kSortWaves, 0, st>>>(kin, vin, n, shift, tiles, hist, kout, vout);
        SR_HIP(hipGetLastError());
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    return 0;
}

__global__ void k_gather_tris(const float4 *__restrict__ src, const int *__restrict__ order, int n,
                              float4 *__restrict__ dst) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = order[i];
    dst[4 * i] = src[4 * s];
    dst[4 * i + 1] = src[4 * s + 1];
    dst[4 * i + 2] = src[4 * s + 2];
    dst[4 * i + 3] = src[4 * s + 3];
}

// Common prefix of (key, index) pairs; -1 outside [0, n).
__device__ inline int prefix_len(const unsigned *keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    unsigned a = keys[i], b = keys[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz((unsigned)i ^ (unsigned)j);
}

// Internal node i -> children (Karras 2012, Alg. 4). Leaves are encoded ~leaf.
__global__ void k_hierarchy(const unsigned *__restrict__ keys, int n, int *__restrict__ child,
                            int *__restrict__ parentInt, int *__restrict__ parentLeaf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (prefix_len(keys, n, i, i + 1) - prefix_len(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = prefix_len(keys, n, i, i - d);
    int lmax = 2;
    while (prefix_len(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (prefix_len(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = prefix_len(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (prefix_len(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? -1 : 0);
    int lo = min(i, j), hi = max(i, j);
    int L = (lo == gamma) ? ~gamma : gamma;
    int R = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    child[2 * i] = L;
    child[2 * i + 1] = R;
    if (L < 0) parentLeaf[~L] = i; else parentInt[L] = i;
    if (R < 0) parentLeaf[~R] = i; else parentInt[R] = i;
}

// Bottom-up refit: the second child to arrive at a node computes its box.
__global__ void k_refit(const float4 *__restrict__ tris, int n, const int *__restrict__ child,
                        const int *__restrict__ parentInt, const int *__restrict__ parentLeaf,
                        float4 *boxes /* (n-1) x 2: min, max */, int *counters, int *depthErr) {
    int leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= n) return;
    int node = parentLeaf[leaf];
    int steps = 0;
    while (node >= 0) {
        __threadfence();
        int prev = atomicAdd(&counters[node], 1);
        if (prev == 0) return;  // the sibling finishes this node
        __threadfence();
        float3 mn = make_float3(1e30f, 1e30f, 1e30f), mx = make_float3(-1e30f, -1e30f, -1e30f);
        for (int c = 0; c < 2; c++) {
            int ch = child[2 * node + c];
            float3 a, b;
            if (ch < 0) {
                TriRec t = tri_unpack(tris[4 * ~ch], tris[4 * ~ch + 1], tris[4 * ~ch + 2], tris[4 * ~ch + 3]);
                a = make_float3(t.bmin.x, t.bmin.y, t.bmin.z);
                b = make_float3(t.bmax.x, t.bmax.y, t.bmax.z);
            } else {
                float4 p = boxes[2 * ch];
                float4 q = boxes[2 * ch + 1];
                a = make_float3(p.x, p.y, p.z);
                b = make_float3(q.x, q.y, q.z);
            }
            mn = make_float3(fminf(mn.x, a.x), fminf(mn.y, a.y), fminf(mn.z, a.z));
            mx = make_float3(fmaxf(mx.x, b.x), fmaxf(mx.y, b.y), fmaxf(mx.z, b.z));
        }
        boxes[2 * node] = make_float4(mn.x, mn.y, mn.z, 0.f);
        boxes[2 * node + 1] = make_float4(mx.x, mx.y, mx.z, 0.f);
        node = parentInt[node];
        if (++steps > 4 * 64 + 64) { atomicExch(depthErr, 1); return; }
    }
}

__global__ void k_pack_nodes(const float4 *__restrict__ tris, int n, const int *__restrict__ child,
                             const float4 *__restrict__ boxes, float4 *__restrict__ nodes) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    float4 cb[2][2];
    for (int c = 0; c < 2; c++) {
        int ch = child[2 * i + c];
        if (ch < 0) {
            TriRec t = tri_unpack(tris[4 * ~ch], tris[4 * ~ch + 1], tris[4 * ~ch + 2], tris[4 * ~ch + 3]);
            cb[c][0] = make_float4(t.bmin.x, t.bmin.y, t.bmin.z, 0.f);
            cb[c][1] = make_float4(t.bmax.x, t.bmax.y, t.bmax.z, 0.f);
        } else {
            cb[c][0] = boxes[2 * ch];
            cb[c][1] = boxes[2 * ch + 1];
        }
    }
    nodes[4 * i + 0] = make_float4(cb[0][0].x, cb[0][0].y, cb[0][0].z, __int_as_float(child[2 * i]));
    nodes[4 * i + 1] = make_float4(cb[0][1].x, cb[0][1].y, cb[0][1].z, __int_as_float(child[2 * i + 1]));
    nodes[4 * i + 2] = make_float4(cb[1][0].x, cb[1][0].y, cb[1][0].z, 0.f);
    nodes[4 * i + 3] = make_float4(cb[1][1].x, cb[1][1].y, cb[1][1].z, 0.f);
}

// ---------------------------------------------------------------------------
// Wavefront-cooperative any-hit.
//
// The 64 lanes of a wave carry 64 rays with (usually) one shared end point.
// The wave walks the tree together: node and triangle records are read with
// wave-uniform addresses (scalar loads), each lane tests its own ray, and a
// child is entered when any still-unoccluded lane's ray touches it. The stack
// is wave-uniform and lives in LDS. A lane retires on its first hit; the walk
// ends when every lane has retired or the stack is empty.

constexpr int kStack = 128;
constexpr int kBruteTris = 8;

__device__ inline unsigned long long ballot64(bool p) { return __ballot(p); }

__device__ inline int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

// The origin terms (origin_terms) of triangle ti of the LDS-staged records, for
// ray origin o; zeros past the last triangle.
__device__ inline float4 staged_origin_terms(const float4 *ldsTris, int ti, int ntri, F3 o) {
    if (ti >= ntri) return make_float4(0.f, 0.f, 0.f, 0.f);
    return origin_terms(tri_unpack(ldsTris[4 * ti], ldsTris[4 * ti + 1], ldsTris[4 * ti + 2], ldsTris[4 * ti + 3]), o);
}

// wave_anyhit<true> for at most kPreTris occluders: the lane's candidate
// triangles cand (pair_candidates; 0 for no ray), the origin terms preE[64 ti]
// of the emitter (a broadcast LDS read) and preR[ti] of the receiver, the ray
// leaving the emitter along uv (fromEm) or the receiver along -uv. The triangles
// left out cannot be hit: the same result as testing every one.
__device__ inline bool wave_anyhit_pre(unsigned cand, const VisConst &vc, float tmin, float tmax,
                                       const float4 *ldsTris, const float4 *preE, const float4 (&preR)[kPreTris],
                                       bool fromEm, F3 ep, F3 rp, F3 uv) {
    if (ballot64(cand != 0u) == 0) return false;
    const F3 o = fromEm ? ep : rp;
    const F3 d = fromEm ? uv : F3{-uv.x, -uv.y, -uv.z};
    bool hit = false;
    // the records are read from LDS at each use (an opaque zero offset keeps the
    // compiler from holding all of them in registers across the emitter loop)
    int off = 0;
    asm volatile("" : "+v"(off));
    ldsTris += off;
#pragma unroll
    for (int ti = 0; ti < kPreTris; ti++) {
        const bool need = ((cand >> ti) & 1u) && !hit;
        if (ballot64(need) == 0) continue;
        TriRec t = tri_unpack(ldsTris[4 * ti], ldsTris[4 * ti + 1], ldsTris[4 * ti + 2], ldsTris[4 * ti + 3]);
        const float4 pe = preE[64 * ti];
        const float4 qt = fromEm ? pe : preR[ti];
        if (need) hit = mt_hit_pre(t, vc.padTri, o, d, tmin, tmax, qt);
        if (ballot64((cand >> (ti + 1)) != 0u && !hit) == 0) break;
    }
    return hit;
}

// Returns this lane's "occluded" flag. active: lane has a ray to trace.
// skipA, skipB: the triangles this lane's end points lie on (-1: none); never tested.
// FEW (at most kBruteTris occluders): test every triangle, no tree, from the
// records staged in LDS (ldsTris).
template <bool FEW>
__device__ inline bool wave_anyhit(const Bvh &bvh, const VisConst &vc, int *stk, bool active, F3 o, F3 d,
                                   float tmin, float tmax, int *errFlag, int skipA, int skipB,
                                   const float4 *ldsTris) {
    if (bvh.ntri <= 0) return false;
    bool hit = false;
    unsigned long long live = ballot64(active);
    if (live == 0) return false;
    if (FEW) {
        // a triangle no lane needs is passed over by the whole wave
        const float4 *tr = ldsTris;
        for (int ti = 0; ti < bvh.ntri; ti++) {
            const float4 q3 = tr[4 * ti + 3];
            const int id = __float_as_int(q3.w);
            const bool need = active && !hit && id != skipA && id != skipB;
            if (ballot64(need) == 0) continue;
            TriRec t = tri_unpack(tr[4 * ti], tr[4 * ti + 1], tr[4 * ti + 2], q3);
            if (need) hit = mt_hit_dev(t, vc.padTri, o, d, tmin, tmax);
            if (ballot64(active && !hit) == 0) break;
        }
        return hit;
    }
    F3 inv{fast_safe_inv(d.x), fast_safe_inv(d.y), fast_safe_inv(d.z)};
    const int lane = __lane_id();
    int sp = 0;
    int node = 0;
    for (;;) {
        // node is an internal node index (wave-uniform)
        const float4 *np = bvh.nodes + 4 * node;
        float4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
        int cl = __float_as_int(n0.w), cr = __float_as_int(n1.w);
        bool want = active && !hit;
        bool hl = want && slab_hit(F3{n0.x, n0.y, n0.z}, F3{n1.x, n1.y, n1.z}, vc.padNode, o, inv, tmin, tmax);
        bool hr = want && slab_hit(F3{n2.x, n2.y, n2.z}, F3{n3.x, n3.y, n3.z}, vc.padNode, o, inv, tmin, tmax);
        unsigned long long bl = ballot64(hl), br = ballot64(hr);
        // leaves are resolved immediately
        if (cl < 0 && bl) {
            int ti = ~cl;
            const float4 q3 = bvh.tris[4 * ti + 3];
            const int id = __float_as_int(q3.w);
            TriRec t = tri_unpack(bvh.tris[4 * ti], bvh.tris[4 * ti + 1], bvh.tris[4 * ti + 2], q3);
            if (hl && id != skipA && id != skipB && mt_hit_dev(t, vc.padTri, o, d, tmin, tmax)) hit = true;
            bl = 0;
        }
        if (cr < 0 && br) {
            int ti = ~cr;
            const float4 q3 = bvh.tris[4 * ti + 3];
            const int id = __float_as_int(q3.w);
            TriRec t = tri_unpack(bvh.tris[4 * ti], bvh.tris[4 * ti + 1], bvh.tris[4 * ti + 2], q3);
            if (hr && !hit && id != skipA && id != skipB && mt_hit_dev(t, vc.padTri, o, d, tmin, tmax)) hit = true;
            br = 0;
        }
        unsigned long long remaining = ballot64(active && !hit);
        if (remaining == 0) break;
        int next = -1;
        if (bl && br) {
            if (sp >= kStack) {  // a plain system-scope store: the flag may be a mapped host word
                if (lane == 0) __hip_atomic_store(errFlag, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                break;
            }
            if (lane == 0) stk[sp] = cr;
            sp++;
            next = cl;
        } else if (bl) {
            next = cl;
        } else if (br) {
            next = cr;
        } else {
            if (sp == 0) break;
            sp--;
            __builtin_amdgcn_wave_barrier();
            next = uniform_int(stk[sp]);
        }
        node = uniform_int(next);
    }
    return hit;
}

// One wave's share: receiver block rb, emitters [ec*emPerWave, (ec+1)*emPerWave).
// The emitters' records are read 64 at a time, one per lane, into the wave's LDS
// slots and read back by the whole wave at one address each.
// FEW: at most kBruteTris occluders, their records staged in LDS at ldsTris.
// sEm: this wave's LDS slots for 64 emitters' points and normals.
template <bool FEW, bool PRE = false>
__device__ __attribute__((always_inline)) inline void vis_wave(const MaskTask &tk, const Bvh &bvh,
                                                               const VisConst &vc, int rb, int ec, int emPerWave,
                                                               int *stk, const float4 *ldsTris, float4 *sEm,
                                                               float4 *sPre = nullptr, unsigned *sBits = nullptr) {
    const int lane = __lane_id();
    const int r = rb * 64 + lane;
    const bool rvalid = r < tk.rc.count;
    float4 rp = rvalid ? tk.rc.p[r] : make_float4(0.f, 0.f, 0.f, 0.f);
    float4 rn = (rvalid && tk.rc.n) ? tk.rc.n[r] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int rfacet = __float_as_int(rn.w);
    // PRE (at most kPreTris occluders): each triangle's origin terms for this
    // lane's receiver as the ray origin, once; the emitters' are staged with them
    float4 preR[kPreTris];
    unsigned bitsR = 0;
    if (PRE) {
#pragma unroll
        for (int ti = 0; ti < kPreTris; ti++) preR[ti] = staged_origin_terms(ldsTris, ti, bvh.ntri, F3{rp.x, rp.y, rp.z});
        bitsR = plane_bits(ldsTris, bvh.ntri, vc.scale, F3{rp.x, rp.y, rp.z},
                           tk.rcTriBase >= 0 ? tk.rcTriBase + rfacet : -1);
    }
    int e0 = ec * emPerWave;
    const int e1 = min(tk.em.count, e0 + emPerWave);
    if (tk.upperOnly) e0 = max(e0, rb * 64);
    for (int e = e0; e < e1; e++) {
        const int k = (e - e0) & 63;
        if (k == 0) {  // the next 64 emitters, one per lane, into the wave's LDS slots
            const int el = min(e + lane, e1 - 1);
            // the wave's reads of the previous 64 precede these writes (its LDS
            // accesses run in order), as these precede the reads below
            const float4 epl = tk.em.p[el];
            const float4 enl = tk.em.n ? tk.em.n[el] : make_float4(0.f, 0.f, 0.f, 0.f);
            sEm[lane] = epl;
            sEm[64 + lane] = enl;
            if (PRE) {
#pragma unroll
                for (int ti = 0; ti < kPreTris; ti++)
                    sPre[64 * ti + lane] = staged_origin_terms(ldsTris, ti, bvh.ntri, F3{epl.x, epl.y, epl.z});
                sBits[lane] = plane_bits(ldsTris, bvh.ntri, vc.scale, F3{epl.x, epl.y, epl.z},
                                         tk.emTriBase >= 0 ? tk.emTriBase + __float_as_int(enl.w) : -1);
            }
        }
        const float4 ep = sEm[k];  // one address for the whole wave: a broadcast
        const float4 en = sEm[64 + k];
        bool pair = rvalid;
        // an emitter on every lane's own facet pairs with none of them: no ray set-up
        if (tk.origin == ORIGIN_LOWER_FACET && ballot64(rvalid && __float_as_int(en.w) != rfacet) == 0ull) {
            if (lane == 0) tk.mask[(size_t)rb * tk.em.count + e] = 0ull;
            continue;
        }
        // Orientation: which end the ray leaves from. The unit direction and the
        // length are formed once, emitter -> receiver; the ray leaving the
        // receiver has exactly the negated direction (its difference vector is
        // the exact negation, so the same length and reciprocal), and the two
        // cosine culls are the same values whichever end it leaves from.
        bool fromEm = true;
        if (tk.origin == ORIGIN_RECEIVER) fromEm = false;
        if (tk.origin == ORIGIN_LOWER_FACET) {
            int ef = __float_as_int(en.w);
            if (ef == rfacet) pair = false;
            fromEm = ef < rfacet;
        }
        F3 uv; float rr;
        ray_setup_dev(F3{ep.x, ep.y, ep.z}, F3{rp.x, rp.y, rp.z}, uv, rr);
        if (tk.useNormalEm && !(f3dot(uv, F3{en.x, en.y, en.z}) >= 1e-6f)) pair = false;
        if (tk.useNormalRc && !(f3dot(F3{-uv.x, -uv.y, -uv.z}, F3{rn.x, rn.y, rn.z}) >= 1e-6f)) pair = false;
        bool occ;
        if (PRE) {
            const unsigned cand = pair ? pair_candidates(sBits[k], bitsR, bvh.ntri) : 0u;
            occ = wave_anyhit_pre(cand, vc, tk.tmin, rr - 1e-3f, ldsTris, sPre + k, preR, fromEm,
                                  F3{ep.x, ep.y, ep.z}, F3{rp.x, rp.y, rp.z}, uv);
        } else {
            const F3 o = fromEm ? F3{ep.x, ep.y, ep.z} : F3{rp.x, rp.y, rp.z};
            if (!fromEm) uv = F3{-uv.x, -uv.y, -uv.z};
            const int triEm = tk.emTriBase >= 0 ? tk.emTriBase + __float_as_int(en.w) : -1;
            const int triRc = tk.rcTriBase >= 0 ? tk.rcTriBase + rfacet : -1;
            occ = wave_anyhit<FEW>(bvh, vc, stk, pair, o, uv, tk.tmin, rr - 1e-3f, tk.errFlag, triEm, triRc, ldsTris);
        }
        unsigned long long word = ballot64(pair && !occ);
        if (lane == 0) tk.mask[(size_t)rb * tk.em.count + e] = word;
    }
}

// Waves of a symmetric set (upperOnly, emitters = receivers) with runs of
// emPerWave dividing 64: only the (receiver block rb, emitter run ec) pairs with
// emitters at or above 64 rb, f = 64 / emPerWave runs fewer per block:
// waves before block rb: P(rb) = rb C - f rb (rb - 1) / 2 (C runs per block).
__host__ __device__ inline long upper_waves_before(long rb, long C, long f) { return rb * C - f * rb * (rb - 1) / 2; }
__host__ __device__ inline bool compact_upper(const MaskTask &tk, int emPerWave) {
    return tk.upperOnly && tk.em.count == tk.rc.count && emPerWave <= 64 && 64 % emPerWave == 0;
}
__device__ inline void upper_wave(long gw, int nrb, int C, int f, int &rb, int &ec) {
    const double a = C + 0.5 * f;
    long r = (long)((a - sqrt(fmax(0.0, a * a - 2.0 * f * (double)gw))) / f);
    r = r < 0 ? 0 : (r > nrb - 1 ? nrb - 1 : r);
    while (r > 0 && upper_waves_before(r, C, f) > gw) r--;
    while (r + 1 < nrb && upper_waves_before(r + 1, C, f) <= gw) r++;
    rb = (int)r;
    ec = (int)(r * f + (gw - upper_waves_before(r, C, f)));
}

// One wave = one receiver block of 64 and a run of emitters. PRE: at most
// kPreTris occluders (tested from LDS with their origin terms kept per origin).
template <int WAVES, bool PRE>
__global__ __launch_bounds__(64 * WAVES) void k_vis_mask(MaskTask tk, Bvh bvh, VisConst vc, int emPerWave) {
    __shared__ int stkAll[WAVES][PRE ? 1 : kStack];
    __shared__ float4 sEmAll[WAVES][128];
    __shared__ float4 sTri[4 * kBruteTris];  // few occluders: their records, staged once per block
    __shared__ float4 sPreAll[WAVES][PRE ? kPreTris * 64 : 1];  // the staged emitters' origin terms
    __shared__ unsigned sBitsAll[WAVES][PRE ? 64 : 1];           // ... and plane-side bits
    const bool brute = bvh.ntri > 0 && bvh.ntri <= kBruteTris;
    if (brute)
        for (int q = threadIdx.x; q < 4 * bvh.ntri; q += 64 * WAVES) sTri[q] = bvh.tris[q];
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const int nrb = (tk.rc.count + 63) >> 6;
    const int nEmChunks = (tk.em.count + emPerWave - 1) / emPerWave;
    const long gw = (long)blockIdx.x * WAVES + wave;
    const bool compact = compact_upper(tk, emPerWave);
    const int f = 64 / emPerWave;
    if (gw >= (compact ? upper_waves_before(nrb, nEmChunks, f) : (long)nrb * nEmChunks)) return;
    int rb = (int)(gw / nEmChunks), ec = (int)(gw % nEmChunks);
    if (compact) upper_wave(gw, nrb, nEmChunks, f, rb, ec);
    if (PRE)
        vis_wave<true, true>(tk, bvh, vc, rb, ec, emPerWave, nullptr, sTri, sEmAll[wave], sPreAll[wave],
                             sBitsAll[wave]);
    else if (brute)
        vis_wave<true>(tk, bvh, vc, rb, ec, emPerWave, stkAll[wave], sTri, sEmAll[wave]);
    else
        vis_wave<false>(tk, bvh, vc, rb, ec, emPerWave, stkAll[wave], nullptr, sEmAll[wave]);
}

// Persistent form over many tasks (scenes): each wave pulls (task, receiver
// block, emitter chunk) items from a shared counter until the list is empty.
constexpr int kVisClaim = 8;  // items a wave claims at once from the work counter
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_vis_mask_multi(const MaskTask *__restrict__ tasks,
                                                               const Bvh *__restrict__ bvhs,
                                                               const VisConst *__restrict__ vcs,
                                                               const int4 *__restrict__ items, int nitems,
                                                               int emPerWave, int *counter) {
    __shared__ int stkAll[WAVES][kStack];
    __shared__ float4 sEmAll[WAVES][128];
    __shared__ float4 sTriW[WAVES][4 * kBruteTris];  // few occluders: this wave's scene's records
    const int wave = threadIdx.x >> 6;
    for (;;) {
        // kVisClaim consecutive items per claim: one atomic on the shared counter
        // per item had queued every wave of the GPU on that one address
        int it0 = 0;
        if (__lane_id() == 0) it0 = atomicAdd(counter, kVisClaim);
        it0 = __builtin_amdgcn_readfirstlane(it0);
        if (it0 >= nitems) break;
        const int it1 = min(nitems, it0 + kVisClaim);
        for (int it = it0; it < it1; it++) {
            const int4 w = items[it];
            const Bvh &b = bvhs[w.x];
            if (b.ntri > 0 && b.ntri <= kBruteTris) {
                // staged by this wave alone: its LDS accesses run in order, so the
                // records it reads next are the ones it wrote (after its own last reads)
                if (__lane_id() < 4 * b.ntri) sTriW[wave][__lane_id()] = b.tris[__lane_id()];
                vis_wave<true>(tasks[w.x], b, vcs[w.x], w.y, w.z, emPerWave, stkAll[wave], sTriW[wave],
                               sEmAll[wave]);
            } else {
                vis_wave<false>(tasks[w.x], b, vcs[w.x], w.y, w.z, emPerWave, stkAll[wave], nullptr, sEmAll[wave]);
            }
        }
    }
}

// Fills the words with emitter e < 64*rb of a symmetric set from the computed
// ones: word (rb, 64*eb + i) bit j = word (eb, 64*rb + j) bit i, for eb < rb.
// One wave per block pair, pair -> (rb, eb) with 0 <= eb < rb row-major over the
// strictly lower triangle; lane i gathers its transposed word from 64 ballots
// and the wave stores the 64 words together.
__device__ inline void transpose_block_pair(uint64_t *mask, int n, long pair, int lane) {
    int rb = (int)((1.0 + sqrt(1.0 + 8.0 * (double)pair)) / 2.0);
    while ((long)rb * (rb - 1) / 2 > pair) rb--;
    while ((long)(rb + 1) * rb / 2 <= pair) rb++;
    const int eb = (int)(pair - (long)rb * (rb - 1) / 2);
    const int r = rb * 64 + lane;
    const uint64_t w = (r < n) ? mask[(size_t)eb * n + r] : 0ull;
    uint64_t mine = 0ull;
#pragma unroll 8
    for (int i = 0; i < 64; i++) {
        const uint64_t col = __ballot((w >> i) & 1ull);
        if (lane == i) mine = col;
    }
    const int e = eb * 64 + lane;
    if (e < n) mask[(size_t)rb * n + e] = mine;
}

__global__ __launch_bounds__(256) void k_mask_transpose(uint64_t *mask, int n) {
    const int nrb = (n + 63) >> 6;
    const long pair = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long npairs = (long)nrb * (nrb - 1) / 2;
    if (pair >= npairs) return;  // uniform over the wave
    transpose_block_pair(mask, n, pair, __lane_id());
}

// ---------------------------------------------------------------------------
// Host launchers.

int lbvh_build(const float4 *dTrisUnsorted, int n, float3 lo, float3 hi, float4 *dTrisSorted, float4 *dNodes,
               hipStream_t st) {
    if (n <= 0) return 0;
    if (n == 1) {
        SR_HIP(hipMemcpyAsync(dTrisSorted, dTrisUnsorted, 4 * sizeof(float4), hipMemcpyDeviceToDevice, st));
        return 0;
    }
    float3 ext = make_float3(hi.x - lo.x, hi.y - lo.y, hi.z - lo.z);
    float3 invExt = make_float3(ext.x > 0 ? 1.0f / ext.x : 0.f, ext.y > 0 ? 1.0f / ext.y : 0.f,
                                ext.z > 0 ? 1.0f / ext.z : 0.f);
    unsigned *keys = nullptr, *keys2 = nullptr;
    int *ids = nullptr, *ids2 = nullptr, *child = nullptr, *pInt = nullptr, *pLeaf = nullptr, *cnt = nullptr;
    int *err = nullptr;
    float4 *boxes = nullptr;
    unsigned *hist = nullptr;
    const int sortTiles = (n + kSortTile - 1) / kSortTile;
    SR_HIP(hipMalloc(&keys, n * sizeof(unsigned)));
    SR_HIP(hipMalloc(&keys2, n * sizeof(unsigned)));
    SR_HIP(hipMalloc(&ids, n * sizeof(int)));
    SR_HIP(hipMalloc(&ids2, n * sizeof(int)));
    SR_HIP(hipMalloc(&child, 2 * (n - 1) * sizeof(int)));
    SR_HIP(hipMalloc(&pInt, (n - 1) * sizeof(int)));
    SR_HIP(hipMalloc(&pLeaf, n * sizeof(int)));
    SR_HIP(hipMalloc(&cnt, (n - 1) * sizeof(int)));
    SR_HIP(hipMalloc(&err, sizeof(int)));
    SR_HIP(hipMalloc(&boxes, 2 * (n - 1) * sizeof(float4)));
    SR_HIP(hipMalloc(&hist, (size_t)256 * sortTiles * sizeof(unsigned)));
    const int B = 256, G = (n + B - 1) / B;
    k_morton<<<G, B, 0, st>>>(dTrisUnsorted, n, lo, invExt, keys, ids);
    SR_HIP(hipGetLastError());
    if (radix_sort_pairs(keys, ids, keys2, ids2, hist, n, 30, st)) return -1;  // 30-bit Morton keys
    k_gather_tris<<<G, B, 0, st>>>(dTrisUnsorted, ids, n, dTrisSorted);
    SR_HIP(hipGetLastError());
    SR_HIP(hipMemsetAsync(pInt, 0xff, (n - 1) * sizeof(int), st));
    SR_HIP(hipMemsetAsync(cnt, 0, (n - 1) * sizeof(int), st));
    SR_HIP(hipMemsetAsync(err, 0, sizeof(int), st));
    const int Gi = (n - 1 + B - 1) / B;
    k_hierarchy<<<Gi, B, 0, st>>>(keys, n, child, pInt, pLeaf);
    SR_HIP(hipGetLastError());
    k_refit<<<G, B, 0, st>>>(dTrisSorted, n, child, pInt, pLeaf, boxes, cnt, err);
    SR_HIP(hipGetLastError());
    k_pack_nodes<<<Gi, B, 0, st>>>(dTrisSorted, n, child, boxes, dNodes);
    SR_HIP(hipGetLastError());
    int herr = 0;
    SR_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
    SR_HIP(hipStreamSynchronize(st));
    void *bufs[] = {keys, keys2, ids, ids2, child, pInt, pLeaf, cnt, err, boxes, hist};
    for (void *b : bufs) (void)hipFree(b);
    if (herr) { set_error("LBVH refit exceeded the depth limit"); return -1; }
    return 0;
}

int vis_mask_fill_lower(uint64_t *mask, int n, hipStream_t st);

// deferLower: leave a symmetric set's words below the diagonal blocks unfilled
// (vis_mask_fill_lower fills them when a caller needs them).
int vis_mask_launch(const MaskTask &tk, const Bvh &bvh, const VisConst &vc, hipStream_t st, bool deferLower) {
    if (tk.rc.count <= 0 || tk.em.count <= 0) return 0;
    constexpr int W = 4;
    const long nrb = (tk.rc.count + 63) / 64;
    // Enough waves to fill 256 CUs several times over, but runs of emitters long
    // enough that the receiver block's data is reused (shorter runs for small sets).
    // A symmetric set launches only the waves with emitters (compact_upper).
    auto waves_for = [&](int epw) {
        const long C = (tk.em.count + epw - 1) / epw;
        return compact_upper(tk, epw) ? upper_waves_before(nrb, C, 64 / epw) : nrb * C;
    };
    int emPerWave = 32;
    while (emPerWave > 1 && waves_for(emPerWave) < 8192) emPerWave /= 2;
    long waves = waves_for(emPerWave);
    while (waves > (1L << 22) && emPerWave < 1024) {
        emPerWave *= 2;
        waves = waves_for(emPerWave);
    }
    long blocks = (waves + W - 1) / W;
    if (blocks > 0x7fffffffL) { set_error("visibility grid too large"); return -1; }
    if (bvh.ntri > 0 && bvh.ntri <= kPreTris) k_vis_mask<W, true><<<(unsigned)blocks, 64 * W, 0, st>>>(tk, bvh, vc, emPerWave);
    else k_vis_mask<W, false><<<(unsigned)blocks, 64 * W, 0, st>>>(tk, bvh, vc, emPerWave);
    SR_HIP(hipGetLastError());
    if (tk.upperOnly && !deferLower) return vis_mask_fill_lower(tk.mask, tk.rc.count, st);
    return 0;
}

// The words below the diagonal blocks of a symmetric set of n, from those above.
int vis_mask_fill_lower(uint64_t *mask, int n, hipStream_t st) {
    const long nrb = (n + 63) / 64;
    if (nrb <= 1) return 0;
    const long pairs = nrb * (nrb - 1) / 2;
    k_mask_transpose<<<(unsigned)((pairs * 64 + 255) / 256), 256, 0, st>>>(mask, n);
    SR_HIP(hipGetLastError());
    return 0;
}

// Multi-task transposition: item (task, lower-triangle block pair index), one wave each.
__global__ __launch_bounds__(256) void k_mask_transpose_multi(const MaskTask *__restrict__ tasks,
                                                              const int2 *__restrict__ items, int nitems) {
    const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (w >= nitems) return;  // uniform over the wave
    const int2 it = items[w];
    transpose_block_pair(tasks[it.x].mask, tasks[it.x].rc.count, it.y, __lane_id());
}

int mask_transpose_multi_launch(const MaskTask *tasks, const int2 *items, int nitems, hipStream_t st) {
    if (nitems <= 0) return 0;
    k_mask_transpose_multi<<<(unsigned)(((long)nitems * 64 + 255) / 256), 256, 0, st>>>(tasks, items, nitems);
    SR_HIP(hipGetLastError());
    return 0;
}

int vis_mask_multi_launch(const MaskTask *tasks, const Bvh *bvhs, const VisConst *vcs, const int4 *items, int nitems,
                          int emPerWave, int *counter, int grid, hipStream_t st) {
    if (nitems <= 0) return 0;
    k_vis_mask_multi<4><<<grid, 256, 0, st>>>(tasks, bvhs, vcs, items, nitems, emPerWave, counter);
    SR_HIP(hipGetLastError());
    return 0;
}

}  // namespace sr
