// Host FP32 fragmenter. Compiled without FMA contraction so every rounding
// happens where the reference's host compiler puts it (src/Facet.cpp is built
// by the host C++ compiler with default flags).
#include "fragmenter.h"

#include <cmath>

namespace sr {

namespace {

inline Vec3 sub(Vec3 dest, Vec3 origin) { return Vec3{dest.x - origin.x, dest.y - origin.y, dest.z - origin.z}; }
inline float dist(Vec3 a, Vec3 b) {
    return sqrtf((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) + (a.z - b.z) * (a.z - b.z));
}
inline float norm(Vec3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
// GeoMath::MakeUnitVectorSc divides by the length (three divisions).
inline Vec3 unit(Vec3 v) {
    float l = norm(v);
    return Vec3{v.x / l, v.y / l, v.z / l};
}

// Foot of `apex` on the line through `p0` along `dir`; returns foot -> apex.
inline Vec3 drop_perpendicular(Vec3 p0, Vec3 dir, Vec3 apex, Vec3 &foot) {
    Vec3 toApex = sub(apex, p0);
    Vec3 u = unit(dir);
    float along = u.x * toApex.x + u.y * toApex.y + u.z * toApex.z;
    foot = Vec3{p0.x + u.x * along, p0.y + u.y * along, p0.z + u.z * along};
    return sub(apex, foot);
}

// Corner-inside flags of one cell.
enum : unsigned { LB = 1u, RB = 2u, LT = 4u, RT = 8u };

}  // namespace

FacetFrame facet_frame(Vec3 v1, Vec3 v2, Vec3 v3) {
    FacetFrame f;
    f.v1 = v1; f.v2 = v2; f.v3 = v3;
    Vec3 a = sub(v1, v2), b = sub(v1, v3);
    f.normal = unit(Vec3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x});

    // Longest edge, first maximum wins; candidates (v1v2, v1v3, v2v3).
    const float len[3] = {dist(v1, v2), dist(v1, v3), dist(v2, v3)};
    int longest = 0;
    if (len[1] > len[longest]) longest = 1;
    if (len[2] > len[longest]) longest = 2;
    // Edge start, edge end and the opposite vertex for each choice.
    const Vec3 start[3] = {v1, v1, v2};
    const Vec3 end[3] = {v2, v3, v3};
    const Vec3 apex[3] = {v3, v2, v1};
    Vec3 edge = sub(end[longest], start[longest]);
    Vec3 up = drop_perpendicular(start[longest], edge, apex[longest], f.base);
    f.baseLength = len[longest];
    f.height = norm(up);
    f.lenNeg = dist(start[longest], f.base);
    f.lenPos = dist(f.base, end[longest]);
    f.xAxis = unit(edge);
    f.yAxis = unit(up);
    f.area = (f.baseLength * f.height) / 2.0f;
    f.longest = longest;
    return f;
}

FacetGrid facet_grid(const FacetFrame &f, float delta) {
    FacetGrid g;
    const float H = f.height;
    g.nNeg = (int)std::ceil(std::fabs(f.lenNeg) / delta);
    const int nPos = (int)std::ceil(std::fabs(f.lenPos) / delta);
    g.nX = g.nNeg + nPos;
    g.nY = (int)std::ceil(H / delta);
    const int nX = g.nX, nY = g.nY, nNeg = g.nNeg;
    g.area.assign((size_t)nX * nY, 0.0f);

    const float half = delta / 2;
    const float slopeR = H / (-f.lenPos);  // edge from the apex down to the positive end
    const float slopeL = H / (f.lenNeg);   // edge from the negative end up to the apex
    const float full = delta * delta;

    for (int i = 0; i < nX; i++) {
        const float xm = (i < nNeg) ? ((-nNeg + i) * delta) + half : ((i - nNeg) * delta) + half;
        const float xL = xm - half, xR = xm + half;
        const bool right = xm > 0;
        const float m = right ? slopeR : slopeL;
        const float yL = m * xL + H, yR = m * xR + H;
        // Row boundaries below the edge line on the left / right of the column.
        auto below = [&](int j, float y) { return delta * j <= y; };
        for (int j = 0; j < nY; j++) {
            unsigned c = (below(j, yL) ? LB : 0u) | (below(j, yR) ? RB : 0u) | (below(j + 1, yL) ? LT : 0u) |
                         (below(j + 1, yR) ? RT : 0u);
            if (c == 0) break;
            float a = 0.0f;
            if (c == (LB | RB | LT | RT)) {
                a = full;
            } else if (right) {
                const float yB = delta * j, yT = delta * (j + 1);
                switch (c) {
                case LB | RB | LT: {  // top-right corner cut off
                    float xTop = (yT - H) / m;
                    float cut = (yT - yR) * (xR - xTop) / 2;
                    a = full - cut;
                    break;
                }
                case LB | RB:  // both top corners out
                    a = ((yL - yR) * delta / 2) + (yR - yB) * delta;
                    break;
                case LB | LT: {  // line crosses top and bottom
                    float xt = ((yT - H) / m) - xL;
                    float xb = ((yB - H) / m) - xL;
                    a = (xt * delta) + (xb - xt) * delta / 2;
                    break;
                }
                case LB:  // only the bottom-left corner in
                    a = (yL - yB) * (((yB - H) / m) - xL) / 2;
                    break;
                default:
                    break;  // pattern not covered by the reference: area stays 0
                }
            } else {
                const float yc = (j * delta) + half;  // row mid point
                switch (c) {
                case LB | RB | RT:  // top-left corner cut off
                    a = full - (yc + half - yL) * (-(xL - ((yc + half - H) / m)) / 2);
                    break;
                case LB | RB:
                    a = (yR - yL) * delta / 2 + ((yL - (yc - half)) * delta);
                    break;
                case RB | RT:
                    a = -((yc - half - H) / m - (yc + half - H) / m) * delta / 2 - (((yc + half - H) / m) - xR) * delta;
                    break;
                case RB:
                    a = (yR - (yc - half)) * (-((((yc - half) - H) / m) - xR)) / 2;
                    break;
                default:
                    break;
                }
            }
            g.area[(size_t)j * nX + i] = a;
        }
    }
    return g;
}

Vec3 cell_point(const FacetFrame &f, const FacetGrid &g, float delta, int i, int j, bool recvShift) {
    float xo = delta * (float)(i - g.nNeg);
    if (recvShift) xo = xo + delta / 2;
    const float yo = delta * (float)j + delta / 2;
    Vec3 p{f.base.x + f.xAxis.x * xo, f.base.y + f.xAxis.y * xo, f.base.z + f.xAxis.z * xo};
    return Vec3{p.x + f.yAxis.x * yo, p.y + f.yAxis.y * yo, p.z + f.yAxis.z * yo};
}

}  // namespace sr
