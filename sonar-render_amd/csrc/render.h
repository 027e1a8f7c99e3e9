// Host render of the scene's surfaces (the reference's RenderOpenGL without
// OpenGL): textured facets, a perspective camera and a depth test, rasterised
// on the CPU and written as a PNG.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "model.h"

namespace sr {

struct RenderSettings {
    float maxDb = -20.f, minDb = -90.f;  // ModelCuda.hpp:45-48 defaults
    bool phaseTarget = false, phaseField = false;
};

// One object's facets and their per-cell colours (RGBA, cells in facet order,
// cell (i, j) of facet f at cellOff[f] + j*nX + i).
struct RenderObject {
    const std::vector<FacetFrame> *frames;
    const std::vector<FacetGrid> *grids;
    const std::vector<long> *cellOff;
    const float *rgba;
};

// RGB image, rows top to bottom (as the reference's PNG after its vertical flip),
// of the objects drawn in order. Host only.
int render_objects(const std::vector<RenderObject> &objs, int width, int height, const float view[9],
                   std::vector<uint8_t> &rgb);
// The scene's target objects then field objects, coloured from their total
// pressures P of one incidence plane (read back from the GPU).
int render_rgb(Model &m, int plane, int width, int height, const float view[9], const RenderSettings &rs,
               std::vector<uint8_t> &rgb);
// 8-bit RGB PNG (zlib-compressed IDAT).
int write_png(const std::string &path, int width, int height, const uint8_t *rgb, std::string &err);

}  // namespace sr
