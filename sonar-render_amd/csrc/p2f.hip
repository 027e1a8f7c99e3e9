// Point->fragment sums of the single-scene model (src/OptiX.cu:35-110).
//
// Compiled with -ffp-contract=off: with few occluders the batched-source kernel
// evaluates each pair's visibility itself, on the visibility kernels' FP32
// arithmetic (vis_fp32.h), instead of reading the sources->fragments words.

#include <hip/hip_runtime.h>
#include <cstdint>

#include "launch.h"
#include "sums.h"
#include "p2f_term.h"
#include "vis_fp32.h"

namespace sr {

template <bool ATT>
__global__ __launch_bounds__(256) void k_p2f(P2FTask t, Phys ph) {
    p2f_elem<ATT>(t, ph, blockIdx.x * blockDim.x + threadIdx.x);
}

// Batched sources (source s writes plane s only): one wave = 64 cells x one
// source, so every (cell, source) update is independent.
// FUSED (at most kPreTris occluder triangles): the pair's bit is the one the
// visibility kernels give the sources->fragments set (ORIGIN_EMITTER: the ray
// leaves the source; the receiver's cosine cull, cos_inc >= 1e-6, on the same
// values p2f_geom forms; no triangle the cell lies on; candidate triangles by
// the plane-side bits, each tested with the FP32 triangle test), computed here.
template <bool ATT, bool FUSED>
__global__ __launch_bounds__(256) void k_p2f_planes(P2FTask t, Phys ph) {
    __shared__ float4 sTri[FUSED ? 4 * kPreTris : 1];
    if (FUSED) {
        for (int q = threadIdx.x; q < 4 * t.vis.ntri; q += 256) sTri[q] = t.vis.tris[q];
        __syncthreads();
    }
    const long w = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int ncb = (t.ncell + 63) >> 6;
    if (w >= (long)ncb * t.ns) return;  // uniform over the wave
    const int s = (int)(w / ncb), cb = (int)(w % ncb);
    const int lane = threadIdx.x & 63;
    const int c = cb * 64 + lane;
    const bool valid = c < t.ncell;
    if (!FUSED) {
        const uint64_t word = t.mask[(size_t)cb * t.ns + s];
        if (!valid || !((word >> lane) & 1ull)) return;
        const float2 v = p2f_term<ATT>(t, ph, t.cp[c], t.cn[c], s);
        float2 &P = t.P[(size_t)s * t.Pstride + c];
        P.x += v.x;
        P.y += v.y;
        return;
    }
    const float4 cp = valid ? t.cp[c] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 cn = valid ? t.cn[c] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 sp = t.sp[s];
    const P2FGeom g = p2f_geom(t, cp, cn, sp);
    bool vis = valid && (!t.useNormal || g.ci >= 1e-6f);
    const F3 sp3{sp.x, sp.y, sp.z};
    const unsigned bE = plane_bits(sTri, t.vis.ntri, t.vis.scale, sp3, -1);  // sources lie on no triangle
    const unsigned bR = plane_bits(sTri, t.vis.ntri, t.vis.scale, F3{cp.x, cp.y, cp.z},
                                   t.vis.rcTriBase >= 0 ? t.vis.rcTriBase + __float_as_int(cn.w) : -1);
    const unsigned cand = vis ? pair_candidates(bE, bR, t.vis.ntri) : 0u;
    if (__ballot(cand != 0u) != 0ull) {
        bool hit = false;
        if (cand != 0u) {
            // the ray leaves the source along the unit direction v / r (ray_setup)
            const F3 d{g.vx * g.inv, g.vy * g.inv, g.vz * g.inv};
#pragma unroll
            for (int ti = 0; ti < kPreTris; ti++) {
                if (ti < t.vis.ntri && ((cand >> ti) & 1u) && !hit) {
                    const float4 *t4 = sTri + 4 * ti;
                    hit = mt_hit_dev(tri_unpack(t4[0], t4[1], t4[2], t4[3]), t.vis.padTri, sp3, d, t.vis.tmin,
                                     g.r - 1e-3f);
                }
            }
        }
        vis = vis && !hit;
    }
    if (!vis) return;
    const float2 v = p2f_value<ATT>(ph, g, t.samp[s]);
    float2 &P = t.P[(size_t)s * t.Pstride + c];
    P.x += v.x;
    P.y += v.y;
}

int p2f_launch(const P2FTask &t, const Phys &ph, hipStream_t st) {
    if (t.ncell <= 0 || t.ns <= 0) return 0;
    const bool fused = t.vis.tris != nullptr;
    if (fused && (!t.batched || t.vis.ntri < 1 || t.vis.ntri > kPreTris)) {
        set_error("point->fragment with its own visibility: batched sources and 1 to 4 occluder triangles");
        return -2;
    }
    if (!fused && !t.mask) { set_error("point->fragment: no visibility words"); return -2; }
    if (t.batched) {
        const long waves = (long)((t.ncell + 63) / 64) * t.ns;
        const unsigned gb = (unsigned)((waves + 3) / 4);
        if (fused) {
            if (ph.alpha != 0.0f) k_p2f_planes<true, true><<<gb, 256, 0, st>>>(t, ph);
            else k_p2f_planes<false, true><<<gb, 256, 0, st>>>(t, ph);
        } else {
            if (ph.alpha != 0.0f) k_p2f_planes<true, false><<<gb, 256, 0, st>>>(t, ph);
            else k_p2f_planes<false, false><<<gb, 256, 0, st>>>(t, ph);
        }
        SR_HIP(hipGetLastError());
        return 0;
    }
    const unsigned g = (unsigned)((t.ncell + 255) / 256);
    if (ph.alpha != 0.0f) k_p2f<true><<<g, 256, 0, st>>>(t, ph);
    else k_p2f<false><<<g, 256, 0, st>>>(t, ph);
    SR_HIP(hipGetLastError());
    return 0;
}

}  // namespace sr
