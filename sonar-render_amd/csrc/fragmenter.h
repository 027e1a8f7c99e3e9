// Host FP32 fragmenter: splits a triangle into delta-sized planar cells with
// exact clipped areas, and places the cell points the sums use.
#pragma once
#include <vector>

namespace sr {

struct Vec3 { float x, y, z; };

struct FacetFrame {
    Vec3 v1, v2, v3;
    Vec3 normal, xAxis, yAxis, base;
    float baseLength, height, lenNeg, lenPos, area;
    int longest;  // which edge is the base: 0 v1v2, 1 v1v3, 2 v2v3
};

struct FacetGrid {
    int nX, nY, nNeg;
    std::vector<float> area;  // nX*nY, cell (i,j) at j*nX + i
};

// Facet frame (src/Facet.cpp:6-109, src/GeoMath.h:19-106).
FacetFrame facet_frame(Vec3 v1, Vec3 v2, Vec3 v3);

// Cell grid and clipped areas (src/Facet.cpp:111-339).
FacetGrid facet_grid(const FacetFrame &f, float delta);

// Cell point. recvShift=false: x offset delta*(i-nNeg) (emitters, and
// receivers of fragment->fragment; src/OptiX.cu:129-135,210-216,273-279).
// recvShift=true adds delta/2 (point->fragment receivers, src/OptiX.cu:51-57).
Vec3 cell_point(const FacetFrame &f, const FacetGrid &g, float delta, int i, int j, bool recvShift);

}  // namespace sr
