// FP32 shadow-ray visibility for the host oracle (test infrastructure only).
//
// Replaces the reference's optixTrace any-hit (src/OptiX.cu:81-93,163-175,
// 321-333; GAS of all target triangles, double sided, terminate on first hit,
// src/OptiX.cpp:304-350) with a defined FP32 Moller-Trumbore test:
//   * hit iff det != 0, u in [0,1], v >= 0, u+v <= 1, tmin < t < tmax, and the
//     computed hit point lies in the triangle's box padded by padTri;
//   * every step in the order written below, with no FMA contraction;
//   * tmin = 0 for rays leaving a source or field point (as the reference) and
//     tmin = 1e-3 m for rays leaving a fragment (the reference's tmin = 0 from
//     a point on the emitter's own plane has no defined answer; 1e-3 m mirrors
//     the reference's far-end epsilon, OptiX.cu:65,118,199).
// The BVH only culls: a node is entered whenever a box padded by
// padNode > padTri could contain an accepted hit, so BVH and brute force agree
// exactly (checked by the tests).
#pragma once
#include <cmath>
#include <vector>
#include <algorithm>
#include <cstdint>

struct V3 { float x, y, z; };

struct Tri {
    V3 v0, e1, e2;
    V3 bmin, bmax;
};

struct VisParams {
    float scale;
    float padTri;
    float padNode;
    float tminPoint;
    float tminFrag;
};

static inline VisParams make_vis_params(float scale) {
    VisParams p;
    p.scale = scale;
    p.padTri = 1e-5f * scale;
    p.padNode = 4e-5f * scale;
    p.tminPoint = 0.0f;
    p.tminFrag = 1e-3f;
    return p;
}

static inline Tri make_tri(V3 a, V3 b, V3 c) {
    Tri t;
    t.v0 = a;
    t.e1 = V3{b.x - a.x, b.y - a.y, b.z - a.z};
    t.e2 = V3{c.x - a.x, c.y - a.y, c.z - a.z};
    t.bmin = V3{std::min(a.x, std::min(b.x, c.x)), std::min(a.y, std::min(b.y, c.y)), std::min(a.z, std::min(b.z, c.z))};
    t.bmax = V3{std::max(a.x, std::max(b.x, c.x)), std::max(a.y, std::max(b.y, c.y)), std::max(a.z, std::max(b.z, c.z))};
    return t;
}

static inline V3 vcross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static inline float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

static inline bool mt_hit(const Tri &T, float padTri, V3 o, V3 d, float tmin, float tmax) {
    V3 p = vcross(d, T.e2);
    float det = vdot(T.e1, p);
    if (det == 0.0f) return false;
    float inv = 1.0f / det;
    V3 s{o.x - T.v0.x, o.y - T.v0.y, o.z - T.v0.z};
    float u = vdot(s, p) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    V3 q = vcross(s, T.e1);
    float v = vdot(d, q) * inv;
    if (!(v >= 0.0f && u + v <= 1.0f)) return false;
    float t = vdot(T.e2, q) * inv;
    if (!(t > tmin && t < tmax)) return false;
    V3 h{o.x + d.x * t, o.y + d.y * t, o.z + d.z * t};
    return h.x >= T.bmin.x - padTri && h.x <= T.bmax.x + padTri &&
           h.y >= T.bmin.y - padTri && h.y <= T.bmax.y + padTri &&
           h.z >= T.bmin.z - padTri && h.z <= T.bmax.z + padTri;
}

// skipA, skipB: the triangles the ray's two end points lie on (or -1); a ray
// never tests them (it leaves one and ends on the other).
static inline bool anyhit_brute(const Tri *tris, int n, const VisParams &vp, V3 o, V3 d, float tmin, float tmax,
                                int skipA = -1, int skipB = -1) {
    for (int i = 0; i < n; i++)
        if (i != skipA && i != skipB && mt_hit(tris[i], vp.padTri, o, d, tmin, tmax)) return true;
    return false;
}

// ---------------------------------------------------------------------------
// CPU BVH: top-down median split on the widest centroid axis, leaves of <= 4.
struct BvhNode {
    V3 bmin, bmax;
    int left, right;  // children, or -1
    int first, count; // triangle range (leaves)
};
struct Bvh {
    std::vector<BvhNode> nodes;
    std::vector<int> idx;
};

static inline int bvh_build_rec(Bvh &b, const std::vector<Tri> &tris, std::vector<V3> &cent, int first, int count) {
    BvhNode n;
    n.bmin = V3{1e30f, 1e30f, 1e30f};
    n.bmax = V3{-1e30f, -1e30f, -1e30f};
    V3 cmin = n.bmin, cmax = n.bmax;
    for (int i = first; i < first + count; i++) {
        const Tri &t = tris[b.idx[i]];
        n.bmin = V3{std::min(n.bmin.x, t.bmin.x), std::min(n.bmin.y, t.bmin.y), std::min(n.bmin.z, t.bmin.z)};
        n.bmax = V3{std::max(n.bmax.x, t.bmax.x), std::max(n.bmax.y, t.bmax.y), std::max(n.bmax.z, t.bmax.z)};
        V3 c = cent[b.idx[i]];
        cmin = V3{std::min(cmin.x, c.x), std::min(cmin.y, c.y), std::min(cmin.z, c.z)};
        cmax = V3{std::max(cmax.x, c.x), std::max(cmax.y, c.y), std::max(cmax.z, c.z)};
    }
    n.left = n.right = -1;
    n.first = first;
    n.count = count;
    int me = (int)b.nodes.size();
    b.nodes.push_back(n);
    if (count <= 4) return me;
    float ex = cmax.x - cmin.x, ey = cmax.y - cmin.y, ez = cmax.z - cmin.z;
    int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
    int mid = first + count / 2;
    std::nth_element(b.idx.begin() + first, b.idx.begin() + mid, b.idx.begin() + first + count, [&](int a, int c) {
        float ka = axis == 0 ? cent[a].x : (axis == 1 ? cent[a].y : cent[a].z);
        float kc = axis == 0 ? cent[c].x : (axis == 1 ? cent[c].y : cent[c].z);
        return ka < kc || (ka == kc && a < c);
    });
    int l = bvh_build_rec(b, tris, cent, first, mid - first);
    int r = bvh_build_rec(b, tris, cent, mid, first + count - mid);
    b.nodes[me].left = l;
    b.nodes[me].right = r;
    b.nodes[me].count = 0;
    return me;
}

static inline void build_bvh(Bvh &b, const std::vector<Tri> &tris, const VisParams &) {
    b.nodes.clear();
    b.idx.resize(tris.size());
    std::vector<V3> cent(tris.size());
    for (size_t i = 0; i < tris.size(); i++) {
        b.idx[i] = (int)i;
        cent[i] = V3{(tris[i].bmin.x + tris[i].bmax.x) * 0.5f, (tris[i].bmin.y + tris[i].bmax.y) * 0.5f,
                     (tris[i].bmin.z + tris[i].bmax.z) * 0.5f};
    }
    if (!tris.empty()) bvh_build_rec(b, tris, cent, 0, (int)tris.size());
}

static inline float safe_inv(float d) {
    float dd = (std::fabs(d) < 1e-20f) ? (d < 0 ? -1e-20f : 1e-20f) : d;
    return 1.0f / dd;
}

static inline bool box_hit(const BvhNode &n, float pad, V3 o, V3 inv, float tmin, float tmax) {
    float tx0 = (n.bmin.x - pad - o.x) * inv.x, tx1 = (n.bmax.x + pad - o.x) * inv.x;
    float ty0 = (n.bmin.y - pad - o.y) * inv.y, ty1 = (n.bmax.y + pad - o.y) * inv.y;
    float tz0 = (n.bmin.z - pad - o.z) * inv.z, tz1 = (n.bmax.z + pad - o.z) * inv.z;
    float tn = std::max(std::max(std::min(tx0, tx1), std::min(ty0, ty1)), std::min(tz0, tz1));
    float tf = std::min(std::min(std::max(tx0, tx1), std::max(ty0, ty1)), std::max(tz0, tz1));
    return tn <= tf && tf >= tmin - pad && tn <= tmax + pad;
}

static inline bool anyhit_bvh(const Bvh &b, const Tri *tris, const VisParams &vp, V3 o, V3 d, float tmin, float tmax,
                              int skipA = -1, int skipB = -1) {
    if (b.nodes.empty()) return false;
    V3 inv{safe_inv(d.x), safe_inv(d.y), safe_inv(d.z)};
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const BvhNode &n = b.nodes[stack[--sp]];
        if (!box_hit(n, vp.padNode, o, inv, tmin, tmax)) continue;
        if (n.left < 0) {
            for (int i = n.first; i < n.first + n.count; i++) {
                const int ti = b.idx[i];
                if (ti != skipA && ti != skipB && mt_hit(tris[ti], vp.padTri, o, d, tmin, tmax)) return true;
            }
        } else {
            stack[sp++] = n.left;
            stack[sp++] = n.right;
        }
    }
    return false;
}
