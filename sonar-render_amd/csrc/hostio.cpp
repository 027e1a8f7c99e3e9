// Host-side scene I/O and the surface colour map.
//
// * STL: binary and ASCII readers, binary writer (the reference's scenes load
//   meshes through numpy-stl, run_scripts/modules/geo.py:150-156).
// * Colour map of the surface render (src/FacetCuda.cu:10-72): pressure ->
//   brightness from dB between min_dB and max_dB, optional phase -> hue on an
//   HSV wheel with saturation 0.8. The render itself (EGL/GL, PNG) stays
//   outside this library; this produces the per-cell RGBA it would upload.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "hostio.h"

namespace sr {

namespace {

bool looks_ascii(const std::vector<unsigned char> &buf) {
    if (buf.size() < 84) return true;
    uint32_t n;
    std::memcpy(&n, buf.data() + 80, 4);
    if ((size_t)84 + (size_t)n * 50 == buf.size()) return false;  // exact binary size
    return std::strncmp((const char *)buf.data(), "solid", 5) == 0;
}

}  // namespace

int stl_read(const std::string &path, std::vector<float> &tris, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { err = "cannot open " + path; return -2; }
    std::vector<unsigned char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    tris.clear();
    if (!looks_ascii(buf)) {
        uint32_t n;
        std::memcpy(&n, buf.data() + 80, 4);
        tris.resize((size_t)n * 9);
        for (uint32_t i = 0; i < n; i++) std::memcpy(&tris[(size_t)9 * i], buf.data() + 84 + (size_t)50 * i + 12, 36);
        return (int)n;
    }
    std::istringstream in(std::string(buf.begin(), buf.end()));
    std::string tok;
    int inFacet = 0;
    while (in >> tok) {
        if (tok == "vertex") {
            float x, y, z;
            if (!(in >> x >> y >> z)) { err = "bad vertex line in " + path; return -2; }
            tris.push_back(x); tris.push_back(y); tris.push_back(z);
            inFacet++;
        } else if (tok == "endfacet") {
            if (inFacet != 3) { err = "facet without three vertices in " + path; return -2; }
            inFacet = 0;
        }
    }
    if (tris.size() % 9) { err = "truncated ASCII STL " + path; return -2; }
    return (int)(tris.size() / 9);
}

int stl_write(const std::string &path, const float *tris, int n, std::string &err) {
    std::ofstream f(path, std::ios::binary);
    if (!f) { err = "cannot create " + path; return -2; }
    char header[80] = {0};
    std::snprintf(header, sizeof(header), "sonar-render-amd binary STL");
    f.write(header, 80);
    uint32_t cnt = (uint32_t)n;
    f.write((const char *)&cnt, 4);
    for (int i = 0; i < n; i++) {
        const float *t = tris + (size_t)9 * i;
        // normal as the reference computes it: unit((v1 - v2) x (v1 - v3))
        float a[3] = {t[0] - t[3], t[1] - t[4], t[2] - t[5]}, b[3] = {t[0] - t[6], t[1] - t[7], t[2] - t[8]};
        float nx = a[1] * b[2] - a[2] * b[1], ny = a[2] * b[0] - a[0] * b[2], nz = a[0] * b[1] - a[1] * b[0];
        float l = std::sqrt(nx * nx + ny * ny + nz * nz);
        float nrm[3] = {l > 0 ? nx / l : 0.f, l > 0 ? ny / l : 0.f, l > 0 ? nz / l : 0.f};
        f.write((const char *)nrm, 12);
        f.write((const char *)t, 36);
        uint16_t attr = 0;
        f.write((const char *)&attr, 2);
    }
    if (!f) { err = "write failed: " + path; return -1; }
    return 0;
}

// hsv2rgb of src/FacetCuda.cu:10-41 (evaluated in double as written there).
static void hsv_to_rgb(float h, float s, float v, float *out) {
    float c = v * s;
    float x = (float)(c * (1.0 - std::fabs(std::fmod(h * 6.0, 2.0) - 1.0)));
    float m = v - c;
    float r, g, b;
    if (h < 1.0 / 6.0) { r = c; g = x; b = 0; }
    else if (h < 2.0 / 6.0) { r = x; g = c; b = 0; }
    else if (h < 3.0 / 6.0) { r = 0; g = c; b = x; }
    else if (h < 4.0 / 6.0) { r = 0; g = x; b = c; }
    else if (h < 5.0 / 6.0) { r = x; g = 0; b = c; }
    else { r = c; g = 0; b = x; }
    out[0] = r + m; out[1] = g + m; out[2] = b + m; out[3] = 1.0f;
}

void colour_map(const double *p, long n, float minDb, float maxDb, bool phase, float *rgba) {
    const float sat = 0.8f;
    for (long i = 0; i < n; i++) {
        const float re = (float)p[2 * i], im = (float)p[2 * i + 1];
        const float mag0 = std::hypot(re, im);
        const float dB = 20.0f * std::log10(mag0 + 1e-8f);
        float mag = (dB - minDb) / (maxDb - minDb);
        mag = std::fmin(std::fmax(mag, 0.0f), 1.0f);
        float *o = rgba + 4 * i;
        if (phase) {
            const float ph = (float)((std::atan2(im, re) + M_PI) / (2.0f * M_PI));
            hsv_to_rgb(ph, 0.8f, mag, o);
        } else {
            o[0] = o[1] = o[2] = sat * mag;
            o[3] = 1.0f;
        }
    }
}

}  // namespace sr
