// FP32 shadow-ray semantics shared by the visibility kernels.
//
// The reference gates every contribution with an OptiX any-hit ray over all
// target triangles, double sided, t in [0, r - 1e-3] (src/OptiX.cu:81-93,
// 163-175, 321-333; geometry from src/OptiX.cpp:264-350). This build defines
// that test as FP32 Moller-Trumbore, evaluated exactly in the order written
// (the translation unit is compiled with -ffp-contract=off):
//   hit iff det != 0, 0 <= u <= 1, v >= 0, u + v <= 1, tmin < t < tmax and the
//   computed hit point lies inside the triangle's box padded by padTri.
// Ray set-up follows src/OptiX/vec_math.h:493-497,549-552:
//   r = sqrtf(dot(v, v)); uv = v * (1.0f / r).
// Node culling uses boxes padded by padNode > padTri, so the any-hit answer is
// the OR of the triangle tests, independent of the tree.
#pragma once
#include <hip/hip_runtime.h>

namespace sr {

struct F3 { float x, y, z; };

__host__ __device__ inline F3 f3sub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline float f3dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline F3 f3cross(F3 a, F3 b) {
    return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Ray from o towards t: unit direction and length exactly as the reference forms them.
__host__ __device__ inline void ray_setup(F3 o, F3 t, F3 &uv, float &r) {
    F3 v = f3sub(t, o);
    r = sqrtf(f3dot(v, v));
    float inv = 1.0f / r;
    uv = F3{v.x * inv, v.y * inv, v.z * inv};
}

// The same ray on the GPU with IEEE sqrtf and 1/r evaluated without the
// special-case paths that cannot apply to these lengths: for r^2 >= 2^-96 the
// square root needs no denormal pre-scaling, and for 2^-60 <= r <= 2^60 the
// division 1/r needs no range scaling (v_div_scale and v_div_fmas then act as
// the identity and a plain FMA) nor a fix-up (finite, normal operands). What
// remains is the correctly rounded evaluation itself: the hardware square root
// corrected by its two one-ulp residual tests, and the reciprocal refined by
// Newton-Raphson steps with FMA residuals. Lanes outside those ranges (r^2
// below 2^-96: coincident points) take the full IEEE operations.
__device__ inline float sqrt_ieee_normal(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __int_as_float(__float_as_int(s) - 1), sp = __int_as_float(__float_as_int(s) + 1);
    const float rm = fmaf(-sm, s, x), rp = fmaf(-sp, s, x);
    const float t = rm <= 0.0f ? sm : s;
    return rp > 0.0f ? sp : t;
}

__device__ inline float rcp_ieee_normal(float r) {
    float y = __builtin_amdgcn_rcpf(r);
    y = fmaf(fmaf(-r, y, 1.0f), y, y);
    float q = y;
    q = fmaf(fmaf(-r, q, 1.0f), y, q);
    return fmaf(fmaf(-r, q, 1.0f), y, q);
}

__device__ inline void ray_setup_dev(F3 o, F3 t, F3 &uv, float &r) {
    F3 v = f3sub(t, o);
    const float r2 = f3dot(v, v);
    const bool normal = r2 >= 0x1p-96f && r2 <= 0x1p100f;
    float rr = sqrt_ieee_normal(r2), inv = rcp_ieee_normal(rr);
    if (__builtin_expect(__ballot(!normal) != 0ull, 0)) {
        if (!normal) {
            rr = sqrtf(r2);
            inv = 1.0f / rr;
        }
    }
    r = rr;
    uv = F3{v.x * inv, v.y * inv, v.z * inv};
}

struct TriRec {
    F3 v0, e1, e2, bmin, bmax;
};

__host__ __device__ inline TriRec tri_unpack(float4 q0, float4 q1, float4 q2, float4 q3) {
    TriRec t;
    t.v0 = F3{q0.x, q0.y, q0.z};
    t.e1 = F3{q0.w, q1.x, q1.y};
    t.e2 = F3{q1.z, q1.w, q2.x};
    t.bmin = F3{q2.y, q2.z, q2.w};
    t.bmax = F3{q3.x, q3.y, q3.z};
    return t;
}

__host__ __device__ inline bool mt_hit(const TriRec &T, float padTri, F3 o, F3 d, float tmin, float tmax) {
    F3 p = f3cross(d, T.e2);
    float det = f3dot(T.e1, p);
    if (det == 0.0f) return false;
    float inv = 1.0f / det;
    F3 s = f3sub(o, T.v0);
    float u = f3dot(s, p) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    F3 q = f3cross(s, T.e1);
    float v = f3dot(d, q) * inv;
    if (!(v >= 0.0f && u + v <= 1.0f)) return false;
    float t = f3dot(T.e2, q) * inv;
    if (!(t > tmin && t < tmax)) return false;
    F3 h{o.x + d.x * t, o.y + d.y * t, o.z + d.z * t};
    return h.x >= T.bmin.x - padTri && h.x <= T.bmax.x + padTri && h.y >= T.bmin.y - padTri &&
           h.y <= T.bmax.y + padTri && h.z >= T.bmin.z - padTri && h.z <= T.bmax.z + padTri;
}

// mt_hit on the GPU: the same operations, with 1/det by rcp_ieee_normal when
// 2^-60 <= |det| <= 2^60 (bit-identical there) and by the full IEEE division
// otherwise.
__device__ inline bool mt_hit_dev(const TriRec &T, float padTri, F3 o, F3 d, float tmin, float tmax) {
    F3 p = f3cross(d, T.e2);
    float det = f3dot(T.e1, p);
    if (det == 0.0f) return false;
    const bool normal = fabsf(det) >= 0x1p-60f && fabsf(det) <= 0x1p60f;
    float inv = rcp_ieee_normal(det);
    if (__builtin_expect(__ballot(!normal) != 0ull, 0)) {
        if (!normal) inv = 1.0f / det;
    }
    F3 s = f3sub(o, T.v0);
    float u = f3dot(s, p) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    F3 q = f3cross(s, T.e1);
    float v = f3dot(d, q) * inv;
    if (!(v >= 0.0f && u + v <= 1.0f)) return false;
    float t = f3dot(T.e2, q) * inv;
    if (!(t > tmin && t < tmax)) return false;
    F3 h{o.x + d.x * t, o.y + d.y * t, o.z + d.z * t};
    return h.x >= T.bmin.x - padTri && h.x <= T.bmax.x + padTri && h.y >= T.bmin.y - padTri &&
           h.y <= T.bmax.y + padTri && h.z >= T.bmin.z - padTri && h.z <= T.bmax.z + padTri;
}

// mt_hit_dev with the terms that depend only on the ray origin and the triangle
// given: q = (o - v0) x e1 and tq = e2 . q, computed by the same operations
// (origin_terms), so every value is the one mt_hit_dev computes. The hit is the
// conjunction of the same tests; t (from tq) is tested first, before u and v,
// so a wave whose rays all end before the triangle's plane skips those.
__device__ inline float4 origin_terms(const TriRec &T, F3 o) {
    const F3 s = f3sub(o, T.v0);
    const F3 q = f3cross(s, T.e1);
    return make_float4(q.x, q.y, q.z, f3dot(T.e2, q));
}

__device__ inline bool mt_hit_pre(const TriRec &T, float padTri, F3 o, F3 d, float tmin, float tmax, float4 qt) {
    F3 p = f3cross(d, T.e2);
    float det = f3dot(T.e1, p);
    if (det == 0.0f) return false;
    const bool normal = fabsf(det) >= 0x1p-60f && fabsf(det) <= 0x1p60f;
    float inv = rcp_ieee_normal(det);
    if (__builtin_expect(__ballot(!normal) != 0ull, 0)) {
        if (!normal) inv = 1.0f / det;
    }
    float t = qt.w * inv;
    if (!(t > tmin && t < tmax)) return false;
    F3 s = f3sub(o, T.v0);
    float u = f3dot(s, p) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    F3 q{qt.x, qt.y, qt.z};
    float v = f3dot(d, q) * inv;
    if (!(v >= 0.0f && u + v <= 1.0f)) return false;
    F3 h{o.x + d.x * t, o.y + d.y * t, o.z + d.z * t};
    return h.x >= T.bmin.x - padTri && h.x <= T.bmax.x + padTri && h.y >= T.bmin.y - padTri &&
           h.y <= T.bmax.y + padTri && h.z >= T.bmin.z - padTri && h.z <= T.bmax.z + padTri;
}

__host__ __device__ inline float safe_inv(float d) {
    float dd = (fabsf(d) < 1e-20f) ? (d < 0.0f ? -1e-20f : 1e-20f) : d;
    return 1.0f / dd;
}

// Reciprocal for the slab test only: any rounding is far inside the box padding.
__device__ inline float fast_safe_inv(float d) {
    float dd = (fabsf(d) < 1e-20f) ? (d < 0.0f ? -1e-20f : 1e-20f) : d;
    return __builtin_amdgcn_rcpf(dd);
}

// Conservative slab test against a box padded by pad (any rounding here only
// widens the accepted set by far less than the padding margin).
__host__ __device__ inline bool slab_hit(F3 bmin, F3 bmax, float pad, F3 o, F3 inv, float tmin, float tmax) {
    float tx0 = (bmin.x - pad - o.x) * inv.x, tx1 = (bmax.x + pad - o.x) * inv.x;
    float ty0 = (bmin.y - pad - o.y) * inv.y, ty1 = (bmax.y + pad - o.y) * inv.y;
    float tz0 = (bmin.z - pad - o.z) * inv.z, tz1 = (bmax.z + pad - o.z) * inv.z;
    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    return tn <= tf && tf >= tmin - pad && tn <= tmax + pad;
}

constexpr int kPreTris = 4;  // few enough occluders for per-point plane-side bits and origin terms

// Plane-side rule for few occluders. With N = e1 x e2 and side(p) = N.p - N.v0
// (FP32, in this order), a ray whose two end points both lie beyond
// m = 1e-4 scale |e1|_1 |e2|_1 on one side of the triangle's plane is never
// accepted by the FP32 triangle test (mt_hit), so the test can be left out. The
// test's t = (e2.q)/det estimates N.(v0 - o)/(N.d): with both ends on one side
// beyond m, t lies outside (0, L) by a distance that needs errors of order m in
// e2.q or m/L in det to close, while each is off by a few ulps of
// scale |e1| |e2| (|o - v0| <= 2 sqrt(3) scale, |d| = 1), a factor of about 30
// below m; the side values themselves are off by less still. The rule is
// checked against the oracle's FP32 test on random, grazing and sliver cases on
// the CPU.
__device__ inline float plane_margin(const TriRec &T, float scale) {
    const float l1 = (fabsf(T.e1.x) + fabsf(T.e1.y)) + fabsf(T.e1.z);
    const float l2 = (fabsf(T.e2.x) + fabsf(T.e2.y)) + fabsf(T.e2.z);
    return ((1e-4f * scale) * l1) * l2;
}

__device__ inline float plane_side(const TriRec &T, F3 p) {
    const F3 n = f3cross(T.e1, T.e2);
    return f3dot(n, p) - f3dot(n, T.v0);
}

// Per end point and triangle ti < ntri (at most kPreTris): bit ti: beyond the
// margin above the plane, bit 4 + ti: below, bit 8 + ti: the triangle is the one
// the point lies on (own: its index, -1 none). A pair's candidate triangles are
// those with neither a shared side bit nor an own bit of either end.
__device__ inline unsigned plane_bits(const float4 *ldsTris, int ntri, float scale, F3 p, int own) {
    unsigned b = 0;
#pragma unroll
    for (int ti = 0; ti < kPreTris; ti++) {
        if (ti >= ntri) break;
        const float4 q3 = ldsTris[4 * ti + 3];
        const TriRec T = tri_unpack(ldsTris[4 * ti], ldsTris[4 * ti + 1], ldsTris[4 * ti + 2], q3);
        const float m = plane_margin(T, scale), sd = plane_side(T, p);
        b |= ((unsigned)(sd > m) << ti) | ((unsigned)(sd < -m) << (ti + 4)) |
             ((unsigned)(__float_as_int(q3.w) == own) << (ti + 8));
    }
    return b;
}

__device__ inline unsigned pair_candidates(unsigned bE, unsigned bR, int ntri) {
    const unsigned same = bE & bR;
    return ~(same | (same >> 4) | ((bE | bR) >> 8)) & ((1u << ntri) - 1u);
}

// Cosine cull decided without the ray set-up where rounding cannot change it.
// The cull keeps a pair when fl(uv . n) >= 1e-6 with uv = v * (1 / sqrtf(v . v))
// (ray_setup_dev, then f3dot; the receiver's cull is the same with -v). Given
// x = v . n and r2 = v . v evaluated in FP32 (any order, fused or not), with
// s = |v| and c = (v . n) / s exact, and |n| <= 1 + 2^-20:
//   |x - v . n| <= 3.01 u s            (u = 2^-24; a three-term dot product)
//   |fl(uv . n) - c| <= 7.6 u          (sqrt, reciprocal and the three products
//                                       of uv: 4.5 u each term, then the dot)
// so for r2 in [2^-60, 2^60] (no underflow in x^2 or K r2, no overflow):
//   x <= 0                   -> c <= 3.01 u, fl(uv . n) <= 10.7 u < 1e-6: culled;
//   x^2 < 1e-14 r2           -> c < 1e-7 + 3.01 u, fl(uv . n) < 7.3e-7: culled;
//   x > 0, x^2 > 9e-12 r2    -> c > 3e-6 (1 - 3u) - 3.01 u, fl(uv . n) > 2.3e-6: kept.
// Returns 1 kept, 0 culled, 2 undecided (cos within about [1e-7, 3e-6], or r2
// out of range): only those need the exact set-up.
__host__ __device__ inline int cull_margin(float x, float r2) {
    if (!(r2 >= 0x1p-60f && r2 <= 0x1p60f)) return 2;
    if (!(x > 0.0f)) return 0;
    const float x2 = x * x;
    if (x2 > 9e-12f * r2) return 1;
    if (x2 < 1e-14f * r2) return 0;
    return 2;
}

}  // namespace sr
