// FP64 OpenMP host oracle for sonar-render's discrete Kirchhoff sums.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or
// calls this file: only tests/, __graft_entry__.smoke() and the cpu_baseline
// leg of bench.py use it, as the checker.
//
// It restates, loop for loop, the reference's algorithm:
//   * fragmenter ............ src/Facet.cpp:6-339, src/GeoMath.h:19-106
//   * wavenumber ............ src/Model.cpp:44-54 (omega, k rounded to FP32)
//   * point -> fragment ..... src/OptiX.cu:35-110
//   * fragment -> point ..... src/OptiX.cu:112-190
//   * fragment -> fragment .. src/OptiX.cu:192-406 (SELF / RESP / NO_RESP)
//   * projection sequencing . src/ModelCuda.cpp:242-356
//   * buffer semantics ...... src/FacetCuda.cpp:8-62,114-137, src/FacetCuda.cu:87-101
//   * occluder set .......... src/OptiX.cpp:264-296 (all target triangles, double sided)
// Arithmetic is FP32 where the reference is FP32 (geometry, areas, cosines,
// the 1/r^2 term and the A*cos products) and FP64 where the reference uses
// thrust::complex<double> (the Green's function and the accumulations).
// thrust's exp is replaced by glibc's FP64 exp/cos/sin.
//
// Visibility: optixTrace is replaced by an any-hit over all target triangles
// using the FP32 Moller-Trumbore test defined in vis_oracle.h. A CPU BVH with
// padded boxes gives the same answer as brute force (tested); the brute-force
// path is kept as a switch.
//
// Build: g++ -O2 -fopenmp -ffp-contract=off (no FMA contraction: the FP32
// steps must round exactly as written).

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

#include "vis_oracle.h"

using cd = std::complex<double>;

namespace orc {

enum { OBJ_TARGET = 0, OBJ_SOURCE = 1, OBJ_FIELD = 2 };

struct Facet {
    V3 v1, v2, v3;
    V3 normal, xAxis, yAxis, base;
    float baseLength, height, lenNeg, lenPos, area;
    int nX, nY, nNeg;
    std::vector<float> frag;  // nX*nY, index j*nX+i
};

struct Object {
    int type;
    float delta;
    std::vector<Facet> facets;
    std::vector<size_t> cellOff;  // per facet offset into the flat cell arrays
    size_t ncell = 0;
    std::vector<cd> P, Pin, Pout;  // field objects alias all three to P
    int triBase = -1;              // targets: index of the first facet among the occluder triangles
};

// ----------------------------------------------------------------------------
// GeoMath restatement (src/GeoMath.h:19-106), FP32, no contraction.
static inline float length_between(V3 a, V3 b) {
    return sqrtf((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) + (a.z - b.z) * (a.z - b.z));
}
static inline V3 vec_from(V3 origin, V3 dest) { return V3{dest.x - origin.x, dest.y - origin.y, dest.z - origin.z}; }
static inline V3 unit_div(V3 v) {
    float l = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    return V3{v.x / l, v.y / l, v.z / l};
}
static inline float dotp(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 crossp(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// GeoMath::MakeVectorThroughPointPerpToVector (GeoMath.h:58-70)
static inline V3 perp_through(V3 vectPnt, V3 vect, V3 perpPnt, V3 &pntOnLine) {
    V3 vpp = vec_from(vectPnt, perpPnt);
    V3 u = unit_div(vect);
    float proj = dotp(u, vpp);
    pntOnLine = V3{vectPnt.x + u.x * proj, vectPnt.y + u.y * proj, vectPnt.z + u.z * proj};
    return vec_from(pntOnLine, perpPnt);
}

// Facet::Facet + GenerateFacetLimits (Facet.cpp:6-109)
static void make_facet(Facet &f, V3 v1, V3 v2, V3 v3) {
    f.v1 = v1; f.v2 = v2; f.v3 = v3;
    V3 a = vec_from(v2, v1), b = vec_from(v3, v1);
    f.normal = unit_div(crossp(a, b));
    float L[3] = {length_between(v1, v2), length_between(v1, v3), length_between(v2, v3)};
    int mi = 0; float mv = L[0];
    for (int i = 1; i < 3; i++) if (L[i] > mv) { mv = L[i]; mi = i; }  // GeoMath.h:24-38
    V3 X, Y;
    switch (mi) {
    case 0:
        X = vec_from(v1, v2); Y = perp_through(v1, X, v3, f.base);
        f.baseLength = L[0]; f.height = sqrtf(Y.x * Y.x + Y.y * Y.y + Y.z * Y.z);
        f.lenNeg = length_between(v1, f.base); f.lenPos = length_between(f.base, v2);
        break;
    case 1:
        X = vec_from(v1, v3); Y = perp_through(v1, X, v2, f.base);
        f.baseLength = L[1]; f.height = sqrtf(Y.x * Y.x + Y.y * Y.y + Y.z * Y.z);
        f.lenNeg = length_between(v1, f.base); f.lenPos = length_between(f.base, v3);
        break;
    default:
        X = vec_from(v2, v3); Y = perp_through(v2, X, v1, f.base);
        f.baseLength = L[2]; f.height = sqrtf(Y.x * Y.x + Y.y * Y.y + Y.z * Y.z);
        f.lenNeg = length_between(v2, f.base); f.lenPos = length_between(f.base, v3);
        break;
    }
    f.xAxis = unit_div(X);
    f.yAxis = unit_div(Y);
    f.area = (f.baseLength * f.height) / 2.0f;
}

// Facet::MakeFragmentData (Facet.cpp:111-339): exact clipped cell areas.
static void fragment_facet(Facet &f, float delta) {
    const float H = f.height;
    const int nNeg = (int)std::ceil(std::fabs(f.lenNeg) / delta);
    const int nPos = (int)std::ceil(std::fabs(f.lenPos) / delta);
    const int nX = nNeg + nPos;
    const int nY = (int)std::ceil(H / delta);
    f.nX = nX; f.nY = nY; f.nNeg = nNeg;
    f.frag.assign((size_t)nX * nY, 0.0f);
    std::vector<float> vX(nX), vY(nY);
    for (int i = 0; i < nX; i++)
        vX[i] = (i < nNeg) ? ((-nNeg + i) * delta) + delta / 2 : ((i - nNeg) * delta) + delta / 2;
    for (int i = 0; i < nY; i++) vY[i] = (i * delta) + delta / 2;
    std::vector<char> yxL(nY + 1), yxR(nY + 1);
    float *A = f.frag.data();
    const float m1 = H / (-f.lenPos);
    const float m2 = H / (f.lenNeg);
    for (int i = 0; i < nX; i++) {
        const float x = vX[i];
        const float xL = x - (delta / 2);
        const float xR = x + (delta / 2);
        if (x > 0) {
            const float yL = m1 * xL + H;
            const float yR = m1 * xR + H;
            for (int j = 0; j < nY + 1; j++) { yxL[j] = delta * j <= yL; yxR[j] = delta * j <= yR; }
            for (int j = 0; j < nY; j++) {
                const float yB = delta * j, yT = delta * (j + 1);
                const bool lb = yxL[j], rb = yxR[j], lt = yxL[j + 1], rt = yxR[j + 1];
                if (!lb && !rb && !lt && !rt) break;
                if (lb && rb && lt && rt) { A[j * nX + i] = delta * delta; continue; }
                if (lb && rb && lt && !rt) {
                    float xi = (yT - H) / m1;
                    float sub = (yT - yR) * (xR - xi) / 2;
                    A[j * nX + i] = delta * delta - sub;
                    continue;
                }
                if (lb && rb && !lt && !rt) { A[j * nX + i] = ((yL - yR) * delta / 2) + (yR - yB) * delta; continue; }
                if (lb && !rb && lt && !rt) {
                    float xt = ((yT - H) / m1) - xL;
                    float xb = ((yB - H) / m1) - xL;
                    A[j * nX + i] = (xt * delta) + (xb - xt) * delta / 2;
                    continue;
                }
                if (lb && !rb && !lt && !rt) { A[j * nX + i] = (yL - yB) * (((yB - H) / m1) - xL) / 2; continue; }
            }
        } else {
            const float yL = m2 * xL + H;
            const float yR = m2 * xR + H;
            for (int j = 0; j < nY + 1; j++) { yxL[j] = delta * j <= yL; yxR[j] = delta * j <= yR; }
            for (int j = 0; j < nY; j++) {
                const bool lb = yxL[j], rb = yxR[j], lt = yxL[j + 1], rt = yxR[j + 1];
                if (!lb && !rb && !lt && !rt) break;
                if (lb && rb && lt && rt) { A[j * nX + i] = delta * delta; continue; }
                if (lb && rb && !lt && rt) {
                    A[j * nX + i] = delta * delta - (vY[j] + delta / 2 - yL) * (-(xL - ((vY[j] + (delta / 2) - H) / m2)) / 2);
                    continue;
                }
                if (lb && rb && !lt && !rt) {
                    A[j * nX + i] = (yR - yL) * delta / 2 + ((yL - (vY[j] - delta / 2)) * delta);
                    continue;
                }
                if (!lb && rb && !lt && rt) {
                    A[j * nX + i] = -((vY[j] - (delta / 2) - H) / m2 - (vY[j] + (delta / 2) - H) / m2) * delta / 2 -
                                    (((vY[j] + (delta / 2) - H) / m2) - xR) * delta;
                    continue;
                }
                if (!lb && rb && !lt && !rt) {
                    A[j * nX + i] = (yR - (vY[j] - delta / 2)) * (-((((vY[j] - delta / 2) - H) / m2) - xR)) / 2;
                    continue;
                }
            }
        }
    }
}

// Fragment positions. Emitter form: x offset delta*(i-nNeg) (OptiX.cu:129-135,
// 210-216, 273-279). Point->fragment receiver form adds delta/2 (OptiX.cu:51-57).
static inline V3 cell_pos(const Facet &f, float delta, int i, int j, bool recvShift) {
    float xo = delta * (float)(i - f.nNeg);
    if (recvShift) xo = xo + delta / 2;
    float yo = delta * (float)j + delta / 2;
    V3 xa{f.xAxis.x * xo, f.xAxis.y * xo, f.xAxis.z * xo};
    V3 ya{f.yAxis.x * yo, f.yAxis.y * yo, f.yAxis.z * yo};
    V3 p{f.base.x + xa.x, f.base.y + xa.y, f.base.z + xa.z};
    return V3{p.x + ya.x, p.y + ya.y, p.z + ya.z};
}

struct Scene {
    float cp = 1480.f, freq = 1000.f, atten = 0.f, density = 0.f;
    float kf = 0.f;       // FP32 k (Model.cpp:50-52)
    cd k{0, 0};           // (double)k + i*attenuation
    std::vector<Object> targets, fields;
    // source objects (type 1): fragmented like targets, every cell at the
    // coherent amplitude srcObjAmp, projected by the fragment->fragment sum
    std::vector<Object> sources;
    std::vector<cd> srcObjAmp;
    std::vector<V3> srcPos, fldPos;
    std::vector<cd> srcP, fldP;
    // Occluders: all target triangles in load order (OptiX.cpp:273-292).
    std::vector<Tri> tris;
    Bvh bvh;
    VisParams vp;
    bool rendered = false;
    bool bruteForce = false;
    // near-field quadrature (0 = the reference's point rule everywhere)
    int nearSub = 0;
    float nearRadius = 0.f;
};

// Sub-fragment quadrature of one fragment->fragment term (test infrastructure's
// definition of the near-field judge): the delta x delta square centred on the
// emitter point (axes xa, ya of the emitter's facet) is split into nsub x nsub
// cells; each cell centre is an emitter with area A/nsub^2, the cosine culls
// and the r < delta/nsub regularisation of the point rule (OptiX.cu:340-354),
// in FP64. nb == nullptr: a field destination (receiver cosine 1). Returns the
// replacement for A_e cos_s cos_i 0.9 G(r).
static cd near_quadrature_pair(cd k, int n, float dA, float Ae, V3 pe, V3 xa, V3 ya, V3 na, const V3 *nb, V3 pr) {
    const double d = dA, h = d / n;
    const cd ik = cd(0, 1) * k;
    cd acc(0, 0);
    for (int sy = 0; sy < n; sy++)
        for (int sx = 0; sx < n; sx++) {
            const double ox = ((sx + 0.5) * h - 0.5 * d), oy = ((sy + 0.5) * h - 0.5 * d);
            const double px = pe.x + xa.x * ox + ya.x * oy;
            const double py = pe.y + xa.y * ox + ya.y * oy;
            const double pz = pe.z + xa.z * ox + ya.z * oy;
            const double vx = pr.x - px, vy = pr.y - py, vz = pr.z - pz;
            const double r = std::sqrt(vx * vx + vy * vy + vz * vz);
            const double cs = (vx * na.x + vy * na.y + vz * na.z) / r;
            const double ci = nb ? -(vx * nb->x + vy * nb->y + vz * nb->z) / r : 1.0;
            if (!(cs >= 1e-6) || !(ci >= 1e-6)) continue;
            cd G;
            if (r < h) G = ik / (4.0 * M_PI);
            else G = -std::exp(cd(0, 1) * k * r) / (4 * M_PI) * (ik / r) + 1.0 / (r * r);
            acc += cs * ci * 0.9 * G;
        }
    return acc * ((double)Ae / (double)(n * n));
}

static cd near_quadrature(const Scene &s, const Facet &FA, float dA, float Ae, V3 pe, const Facet *FB, V3 pr,
                          bool fieldDst) {
    return near_quadrature_pair(s.k, s.nearSub, dA, Ae, pe, FA.xAxis, FA.yAxis, FA.normal,
                                fieldDst ? nullptr : &FB->normal, pr);
}

static cd expik(const cd &k, float r) {
    // thrust::exp(i1 * k * r) with r promoted to double (OptiX.cu:102,182,350)
    cd z = cd(0, 1) * k * (double)r;
    double m = std::exp(z.real());
    return cd(m * std::cos(z.imag()), m * std::sin(z.imag()));
}

// skipA, skipB: the occluder triangles the ray's end points lie on (-1: a
// source, field point or field object), which the ray never tests.
static inline bool occluded(const Scene &s, V3 o, V3 d, float tmin, float tmax, int skipA, int skipB) {
    if (s.bruteForce) return anyhit_brute(s.tris.data(), (int)s.tris.size(), s.vp, o, d, tmin, tmax, skipA, skipB);
    return anyhit_bvh(s.bvh, s.tris.data(), s.vp, o, d, tmin, tmax, skipA, skipB);
}
static inline int tri_of(const Object &o, int facet) { return o.triBase >= 0 ? o.triBase + facet : -1; }

// Ray set-up exactly as OptiX.cu: r = sqrtf(dot(v,v)); uv = v * (1.0f / r).
static inline void ray_dir(V3 v, float &r, V3 &uv) {
    r = sqrtf(dotp(v, v));
    float inv = 1.0f / r;
    uv = V3{v.x * inv, v.y * inv, v.z * inv};
}

// projectSourcePointsToFacet (OptiX.cu:35-110). Writes every cell (no area test).
static void project_src_to_object(Scene &s, Object &B) {
    const float eps = 1e-3f;
    const int nsrc = (int)s.srcPos.size();
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < (int)B.facets.size(); b++) {
        const Facet &F = B.facets[b];
        for (int x = 0; x < F.nX; x++)
            for (int y = 0; y < F.nY; y++) {
                const size_t cj = B.cellOff[b] + (size_t)y * F.nX + x;
                V3 pj = cell_pos(F, B.delta, x, y, true);
                for (int i = 0; i < nsrc; i++) {
                    V3 pi = s.srcPos[i];
                    V3 v{pj.x - pi.x, pj.y - pi.y, pj.z - pi.z};
                    float r; V3 uv; ray_dir(v, r, uv);
                    float cinc = dotp(V3{-uv.x, -uv.y, -uv.z}, F.normal);
                    if (B.type == OBJ_FIELD) cinc = 1;
                    if (cinc < 1e-6f) continue;
                    if (occluded(s, pi, uv, s.vp.tminPoint, r - eps, -1, tri_of(B, b))) continue;
                    cd Pj = (double)cinc * (s.srcP[i] * expik(s.k, r)) / (double)r;
                    B.P[cj] += Pj;
                }
            }
    }
}

// projectFacetToFieldPoints (OptiX.cu:112-190). Emitter value is the total P.
static void project_object_to_points(Scene &s, Object &A) {
    const float eps = 1e-3f;
    const int nq = (int)s.fldPos.size();
    const double fourpi = 4 * M_PI;
    const cd ik = cd(0, 1) * s.k;
    int nthr = omp_get_max_threads();
    std::vector<std::vector<cd>> part(nthr, std::vector<cd>(nq, cd(0, 0)));
#pragma omp parallel
    {
        std::vector<cd> &acc = part[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 1)
        for (int a = 0; a < (int)A.facets.size(); a++) {
            const Facet &F = A.facets[a];
            for (int x = 0; x < F.nX; x++)
                for (int y = 0; y < F.nY; y++) {
                    const size_t ci = A.cellOff[a] + (size_t)y * F.nX + x;
                    V3 pi = cell_pos(F, A.delta, x, y, false);
                    const float Ai = F.frag[(size_t)y * F.nX + x];
                    if (Ai <= 1e-6f) continue;
                    const cd Pi = A.P[ci];
                    for (int j = 0; j < nq; j++) {
                        V3 pj = s.fldPos[j];
                        V3 v{pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
                        float r; V3 uv; ray_dir(v, r, uv);
                        float cs = dotp(V3{-uv.x, -uv.y, -uv.z}, F.normal);
                        if (cs < 1e-6f) continue;
                        if (occluded(s, pj, uv, s.vp.tminPoint, r - eps, -1, tri_of(A, a))) continue;
                        float invr2 = 1 / (r * r);  // FP32 in the reference
                        cd G = -expik(s.k, r) / fourpi * (ik / (double)r) + (double)invr2;
                        float Acs = Ai * cs;
                        acc[j] += (double)Acs * Pi * G;
                    }
                }
        }
    }
    for (int t = 0; t < nthr; t++)
        for (int j = 0; j < nq; j++) s.fldP[j] += part[t][j];
}

// projectFacetToFacets (OptiX.cu:192-406). mode: 0 SELF, 1 RESP, 2 NO_RESP.
// Accumulates into dst.Pout (and src.Pout for reciprocity).
static void project_facets(Scene &s, Object &A, Object &B, int mode) {
    const bool recip = (mode != 2), self = (mode == 0);
    const float eps = 1e-3f;
    const float dI = A.delta;
    const double fourpi = 4 * M_PI;
    const cd ik = cd(0, 1) * s.k;
    const std::vector<cd> &srcVal = (mode == 2) ? A.P : A.Pin;
    const std::vector<cd> &dstVal = (mode == 2) ? B.P : B.Pin;
    int nthr = omp_get_max_threads();
    std::vector<std::vector<cd>> accB(nthr), accA(nthr);
#pragma omp parallel
    {
        int tid = omp_get_thread_num();
        accB[tid].assign(B.ncell, cd(0, 0));
        if (recip) accA[tid].assign(A.ncell, cd(0, 0));
        std::vector<cd> &oB = accB[tid];
        std::vector<cd> &oA = accA[tid];
#pragma omp for schedule(dynamic, 1)
        for (int a = 0; a < (int)A.facets.size(); a++) {
            const Facet &FA = A.facets[a];
            for (int xi = 0; xi < FA.nX; xi++)
                for (int yi = 0; yi < FA.nY; yi++) {
                    const size_t li = (size_t)yi * FA.nX + xi;
                    const size_t ci = A.cellOff[a] + li;
                    V3 pi = cell_pos(FA, A.delta, xi, yi, false);
                    const float Ai = FA.frag[li];
                    if (Ai <= 1e-6f) continue;
                    const cd Pri = srcVal[ci];
                    if (Pri.real() == 0 && Pri.imag() == 0) continue;  // OptiX.cu:241-245
                    const int b0 = self ? a + 1 : 0;
                    for (int b = b0; b < (int)B.facets.size(); b++) {
                        if (self && b == a) continue;
                        const Facet &FB = B.facets[b];
                        for (int xj = 0; xj < FB.nX; xj++)
                            for (int yj = 0; yj < FB.nY; yj++) {
                                const size_t lj = (size_t)yj * FB.nX + xj;
                                const size_t cj = B.cellOff[b] + lj;
                                V3 pj = cell_pos(FB, B.delta, xj, yj, false);
                                V3 v{pj.x - pi.x, pj.y - pi.y, pj.z - pi.z};
                                float r; V3 uv; ray_dir(v, r, uv);
                                const float Bj = FB.frag[lj];
                                if (Bj <= 1e-6f) continue;
                                float cs = dotp(uv, FA.normal);
                                if (cs < 1e-6f) continue;
                                float cinc = dotp(V3{-uv.x, -uv.y, -uv.z}, FB.normal);
                                if (B.type == OBJ_FIELD) cinc = 1;
                                if (cinc < 1e-6f) continue;
                                if (occluded(s, pi, uv, s.vp.tminFrag, r - eps, tri_of(A, a), tri_of(B, b))) continue;
                                if (s.nearSub > 0 && r < s.nearRadius) {
                                    // near-field judge: quadrature over each emitter's own cell
                                    const bool fld = (B.type == OBJ_FIELD);
                                    oB[cj] += near_quadrature(s, FA, A.delta, Ai, pi, &FB, pj, fld) * Pri;
                                    if (recip)
                                        oA[ci] += near_quadrature(s, FB, B.delta, Bj, pj, &FA, pi, false) * dstVal[cj];
                                    continue;
                                }
                                cd G;
                                if (r < dI) {
                                    G = ik / (4.0 * M_PI);
                                } else {
                                    float invr2 = 1 / (r * r);
                                    G = -expik(s.k, r) / fourpi * (ik / (double)r) + (double)invr2;
                                }
                                G = 0.9 * G;
                                float w = Ai * cs * cinc;
                                oB[cj] += (double)w * Pri * G;
                                if (recip) {
                                    float rw = Bj * cs * cinc;
                                    oA[ci] += (double)rw * dstVal[cj] * G;
                                }
                            }
                    }
                }
        }
    }
    for (int t = 0; t < nthr; t++) {
        for (size_t c = 0; c < B.ncell; c++) B.Pout[c] += accB[t][c];
        if (recip)
            for (size_t c = 0; c < A.ncell; c++) A.Pout[c] += accA[t][c];
    }
}

static void alias_sync_field(Object &o) {
    // Field objects alias P = P_in = P_out (FacetCuda.cpp:35-37): fold P_out into P.
    if (o.type == OBJ_FIELD) {
        for (size_t c = 0; c < o.ncell; c++) { o.P[c] += o.Pout[c]; o.Pout[c] = 0; }
    }
}

}  // namespace orc

using namespace orc;

extern "C" {

void *orc_create() { return new Scene(); }
void orc_destroy(void *h) { delete (Scene *)h; }
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int orc_get_threads() { return omp_get_max_threads(); }
void orc_set_bruteforce(void *h, int on) { ((Scene *)h)->bruteForce = on != 0; }
// One pair of the near-field judge, outside any scene: emitter point pe with
// its facet's axes xa, ya and normal na, cell edge delta and area Ae; receiver
// point pr with normal nb (null: a field destination); k = kr + i atten.
// out2 = the complex replacement of A_e cos_s cos_i 0.9 G(r).
void orc_near_pair(double kr, double atten, int nsub, float delta, float Ae, const float *pe, const float *xa,
                   const float *ya, const float *na, const float *nb, const float *pr, double *out2) {
    auto v = [](const float *q) { return V3{q[0], q[1], q[2]}; };
    V3 nbv = nb ? v(nb) : V3{0, 0, 0};
    cd r = near_quadrature_pair(cd(kr, atten), nsub, delta, Ae, v(pe), v(xa), v(ya), v(na), nb ? &nbv : nullptr,
                                v(pr));
    out2[0] = r.real();
    out2[1] = r.imag();
}

// The oracle's cosine cull of count pairs outside any scene: kept[i] = 1 when
// the cosine of v_i (the difference vector) with unit normal n_i, formed as the
// sums form it (ray_dir, dotp: OptiX.cu:326-333), is not below 1e-6.
void orc_cosine_kept(const float *v, const float *n, long count, unsigned char *kept) {
    for (long i = 0; i < count; i++) {
        float r; V3 uv;
        ray_dir(V3{v[3 * i], v[3 * i + 1], v[3 * i + 2]}, r, uv);
        kept[i] = dotp(uv, V3{n[3 * i], n[3 * i + 1], n[3 * i + 2]}) < 1e-6f ? 0 : 1;
    }
}

void orc_set_near_quadrature(void *h, int nsub, float radius) {
    ((Scene *)h)->nearSub = nsub;
    ((Scene *)h)->nearRadius = radius;
}

// Model::set_inital_conditions (Model.cpp:44-54): omega and k stored as float.
void orc_set_initial_conditions(void *h, float cp, float f, float att, float dens) {
    Scene &s = *(Scene *)h;
    s.cp = cp; s.freq = f; s.atten = att; s.density = dens;
    float omega = (float)(2 * M_PI * f);
    s.kf = omega / cp;
    s.k = cd((double)s.kf, (double)att);
}

int orc_load_geometry(void *h, const float *v, int nf, int type, float res) {
    Scene &s = *(Scene *)h;
    if (type < 0 || type > 2) return -1;
    Object o; o.type = type; o.delta = res;
    o.facets.resize(nf);
    for (int i = 0; i < nf; i++) {
        const float *t = v + (size_t)9 * i;
        make_facet(o.facets[i], V3{t[0], t[1], t[2]}, V3{t[3], t[4], t[5]}, V3{t[6], t[7], t[8]});
    }
    if (type == OBJ_TARGET) s.targets.push_back(std::move(o));
    else if (type == OBJ_FIELD) s.fields.push_back(std::move(o));
    else {
        // Source objects: stored but never used by the reference (Model.cpp:24-42);
        // here the piston-like emitter of README.md:2 (coherent pressure across the object).
        s.sources.push_back(std::move(o));
        s.srcObjAmp.push_back(cd(1, 0));
    }
    return 0;
}
void orc_set_source_object_amplitude(void *h, int i, double re, double im) {
    Scene &s = *(Scene *)h;
    if (i >= 0 && i < (int)s.srcObjAmp.size()) s.srcObjAmp[i] = cd(re, im);
}

void orc_load_source_points(void *h, const float *p, int n) {
    Scene &s = *(Scene *)h;
    for (int i = 0; i < n; i++) { s.srcPos.push_back(V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}); s.srcP.push_back(cd(1, 0)); }
}
void orc_load_field_points(void *h, const float *p, int n) {
    Scene &s = *(Scene *)h;
    for (int i = 0; i < n; i++) { s.fldPos.push_back(V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}); s.fldP.push_back(cd(0, 0)); }
}
// Source amplitude override (the reference fixes 1+0i, Model.cpp:18).
void orc_set_source_amplitude(void *h, int i, double re, double im) {
    Scene &s = *(Scene *)h;
    if (i >= 0 && i < (int)s.srcP.size()) s.srcP[i] = cd(re, im);
}

static void prep_object(Object &o) {
    o.cellOff.resize(o.facets.size());
    size_t off = 0;
    for (size_t f = 0; f < o.facets.size(); f++) {
        fragment_facet(o.facets[f], o.delta);
        o.cellOff[f] = off;
        off += (size_t)o.facets[f].nX * o.facets[f].nY;
    }
    o.ncell = off;
    o.P.assign(off, cd(0, 0)); o.Pin.assign(off, cd(0, 0)); o.Pout.assign(off, cd(0, 0));
}

// render_cuda: fragment every object, build the occluder set (python_interface.cpp:61-65).
int orc_render(void *h) {
    Scene &s = *(Scene *)h;
    if (s.rendered) return -2;
    for (auto &o : s.targets) prep_object(o);
    for (auto &o : s.fields) prep_object(o);
    for (auto &o : s.sources) prep_object(o);
    s.tris.clear();
    float scale = 1.0f;
    auto upd = [&](V3 p) { scale = std::max(scale, std::max(std::fabs(p.x), std::max(std::fabs(p.y), std::fabs(p.z)))); };
    for (auto &o : s.targets) {
        o.triBase = (int)s.tris.size();
        for (auto &f : o.facets) { s.tris.push_back(make_tri(f.v1, f.v2, f.v3)); upd(f.v1); upd(f.v2); upd(f.v3); }
    }
    for (auto &o : s.fields)
        for (auto &f : o.facets) { upd(f.v1); upd(f.v2); upd(f.v3); }
    for (auto &o : s.sources)  // never occluders (only targets are, OptiX.cpp:273-292)
        for (auto &f : o.facets) { upd(f.v1); upd(f.v2); upd(f.v3); }
    for (auto &p : s.srcPos) upd(p);
    for (auto &p : s.fldPos) upd(p);
    s.vp = make_vis_params(scale);
    build_bvh(s.bvh, s.tris, s.vp);
    s.rendered = true;
    return 0;
}

// A source object's cells (all at its amplitude) into the total pressure P of
// object B: the fragment->fragment sum without reciprocity (FACET_NO_RESP,
// OptiX.cu:192-406), receivers at B's fragment points, B's cosine cull unless B
// is a field object.
static void project_source_object(Scene &s, Object &S, cd amp, Object &B) {
    std::fill(S.P.begin(), S.P.end(), amp);
    std::vector<cd> keep = B.Pout;
    std::fill(B.Pout.begin(), B.Pout.end(), cd(0, 0));
    project_facets(s, S, B, 2);
    for (size_t c = 0; c < B.ncell; c++) B.P[c] += B.Pout[c];
    B.Pout.swap(keep);
}

int orc_project_src_to_objects(void *h, int fieldObjects) {
    Scene &s = *(Scene *)h;
    if (!s.rendered) return -3;
    for (auto &o : s.targets) project_src_to_object(s, o);
    if (fieldObjects)
        for (auto &o : s.fields) project_src_to_object(s, o);
    for (size_t i = 0; i < s.sources.size(); i++) {
        for (auto &o : s.targets) project_source_object(s, s.sources[i], s.srcObjAmp[i], o);
        if (fieldObjects)
            for (auto &o : s.fields) project_source_object(s, s.sources[i], s.srcObjAmp[i], o);
    }
    return 0;
}

// ModelCuda::ProjectTargetToTargetObjects (ModelCuda.cpp:323-356)
int orc_project_target_to_target(void *h, int nref) {
    Scene &s = *(Scene *)h;
    if (!s.rendered) return -3;
    for (auto &o : s.targets) o.Pout = o.P;  // PrimeReflections
    for (int it = 0; it < nref; it++)
        for (auto &o : s.targets) {
            o.Pin = o.Pout;                                // SwapOutputToInputPressure
            std::fill(o.Pout.begin(), o.Pout.end(), cd(0, 0));
            project_facets(s, o, o, 0);                     // FACET_SELF
            for (size_t c = 0; c < o.ncell; c++) o.P[c] += o.Pout[c];  // AccumulatePressure
        }
    return 0;
}

// FACET_RESP between two distinct target objects (OptiX.cu:427-430). Never called by
// the reference's orchestration; provided so the kernel mode has a checker.
int orc_project_resp(void *h, int ia, int ib) {
    Scene &s = *(Scene *)h;
    if (!s.rendered) return -3;
    if (ia < 0 || ib < 0 || ia >= (int)s.targets.size() || ib >= (int)s.targets.size() || ia == ib) return -4;
    project_facets(s, s.targets[ia], s.targets[ib], 1);
    return 0;
}
// Copy P_out into P_in of a target (the RESP checker's input set-up).
int orc_target_pout_to_pin(void *h, int it) {
    Scene &s = *(Scene *)h;
    if (it < 0 || it >= (int)s.targets.size()) return -4;
    Object &o = s.targets[it];
    o.Pin = o.P;
    std::fill(o.Pout.begin(), o.Pout.end(), cd(0, 0));
    return 0;
}

int orc_project_target_to_field_objects(void *h) {
    Scene &s = *(Scene *)h;
    if (!s.rendered) return -3;
    for (auto &t : s.targets)
        for (auto &f : s.fields) {
            project_facets(s, t, f, 2);  // FACET_NO_RESP into the aliased field buffer
            alias_sync_field(f);
        }
    return 0;
}

int orc_project_target_to_field_points(void *h) {
    Scene &s = *(Scene *)h;
    if (!s.rendered) return -3;
    for (auto &t : s.targets) project_object_to_points(s, t);
    return 0;
}

int orc_num_field_points(void *h) { return (int)((Scene *)h)->fldPos.size(); }
void orc_get_field_points(void *h, double *out) {
    Scene &s = *(Scene *)h;
    for (size_t i = 0; i < s.fldP.size(); i++) { out[2 * i] = s.fldP[i].real(); out[2 * i + 1] = s.fldP[i].imag(); }
}

static Object *obj(Scene &s, int type, int idx) {
    std::vector<Object> &v = (type == OBJ_FIELD) ? s.fields : (type == OBJ_SOURCE ? s.sources : s.targets);
    if (idx < 0 || idx >= (int)v.size()) return nullptr;
    return &v[idx];
}
long orc_object_ncell(void *h, int type, int idx) {
    Object *o = obj(*(Scene *)h, type, idx);
    return o ? (long)o->ncell : -1;
}
int orc_object_nfacet(void *h, int type, int idx) {
    Object *o = obj(*(Scene *)h, type, idx);
    return o ? (int)o->facets.size() : -1;
}
// Per-facet grid: nX, nY, nNeg.
void orc_object_grids(void *h, int type, int idx, int *out) {
    Object *o = obj(*(Scene *)h, type, idx);
    for (size_t f = 0; f < o->facets.size(); f++) {
        out[3 * f] = o->facets[f].nX; out[3 * f + 1] = o->facets[f].nY; out[3 * f + 2] = o->facets[f].nNeg;
    }
}
// Per-cell data: area, emitter position (3), receiver position (3), facet id.
void orc_object_cells(void *h, int type, int idx, float *area, float *pe, float *pr, int *fid) {
    Object *o = obj(*(Scene *)h, type, idx);
    for (size_t f = 0; f < o->facets.size(); f++) {
        const Facet &F = o->facets[f];
        for (int y = 0; y < F.nY; y++)
            for (int x = 0; x < F.nX; x++) {
                size_t c = o->cellOff[f] + (size_t)y * F.nX + x;
                area[c] = F.frag[(size_t)y * F.nX + x];
                V3 a = cell_pos(F, o->delta, x, y, false), b = cell_pos(F, o->delta, x, y, true);
                pe[3 * c] = a.x; pe[3 * c + 1] = a.y; pe[3 * c + 2] = a.z;
                pr[3 * c] = b.x; pr[3 * c + 1] = b.y; pr[3 * c + 2] = b.z;
                fid[c] = (int)f;
            }
    }
}
// Facet frames: normal(3) xAxis(3) yAxis(3) base(3) area lenNeg lenPos height.
void orc_object_frames(void *h, int type, int idx, float *out) {
    Object *o = obj(*(Scene *)h, type, idx);
    for (size_t f = 0; f < o->facets.size(); f++) {
        const Facet &F = o->facets[f];
        float *q = out + 16 * f;
        V3 vs[4] = {F.normal, F.xAxis, F.yAxis, F.base};
        for (int k = 0; k < 4; k++) { q[3 * k] = vs[k].x; q[3 * k + 1] = vs[k].y; q[3 * k + 2] = vs[k].z; }
        q[12] = F.area; q[13] = F.lenNeg; q[14] = F.lenPos; q[15] = F.height;
    }
}
void orc_get_object_pressure(void *h, int type, int idx, int which, double *out) {
    Object *o = obj(*(Scene *)h, type, idx);
    const std::vector<cd> &v = which == 0 ? o->P : (which == 1 ? o->Pin : o->Pout);
    for (size_t c = 0; c < o->ncell; c++) { out[2 * c] = v[c].real(); out[2 * c + 1] = v[c].imag(); }
}
void orc_set_object_pressure(void *h, int type, int idx, int which, const double *in) {
    Object *o = obj(*(Scene *)h, type, idx);
    std::vector<cd> &v = which == 0 ? o->P : (which == 1 ? o->Pin : o->Pout);
    for (size_t c = 0; c < o->ncell; c++) v[c] = cd(in[2 * c], in[2 * c + 1]);
}

// Direct any-hit queries on the scene's occluders (for visibility parity tests).
// skipA, skipB: per-ray end-point triangles (may be null: none).
void orc_anyhit(void *h, const float *o, const float *d, const float *tmin, const float *tmax, long n,
                unsigned char *hit, int brute, const int *skipA, const int *skipB) {
    Scene &s = *(Scene *)h;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        V3 O{o[3 * i], o[3 * i + 1], o[3 * i + 2]}, D{d[3 * i], d[3 * i + 1], d[3 * i + 2]};
        const int a = skipA ? skipA[i] : -1, b = skipB ? skipB[i] : -1;
        hit[i] = brute ? anyhit_brute(s.tris.data(), (int)s.tris.size(), s.vp, O, D, tmin[i], tmax[i], a, b)
                       : anyhit_bvh(s.bvh, s.tris.data(), s.vp, O, D, tmin[i], tmax[i], a, b);
    }
}
void orc_vis_params(void *h, float *out) {
    Scene &s = *(Scene *)h;
    out[0] = s.vp.scale; out[1] = s.vp.padTri; out[2] = s.vp.padNode; out[3] = s.vp.tminPoint; out[4] = s.vp.tminFrag;
}
float orc_kf(void *h) { return ((Scene *)h)->kf; }

}  // extern "C"
