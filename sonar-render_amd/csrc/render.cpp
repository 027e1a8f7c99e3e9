// Host render of the scene's surfaces: what the reference draws with EGL and
// OpenGL (src/ModelGl.cpp:14-341, src/ObjectGl.cpp:3-110, src/FacetGl.cpp),
// rasterised on the CPU so no GL context or GPU interop is needed.
//
// Same picture as the reference:
//   * camera: glm::perspective(45 deg, width/height, 0.1, 10000) and
//     glm::lookAt(eye, centre, up) from viewSettings[9] (ModelGl.cpp:94-104);
//   * beige background (0.96, 0.96, 0.86), depth test GL_LESS, no face culling;
//   * every facet of the target objects, then of the field objects, drawn as one
//     triangle textured with its nX x nY cell colours (texel (i, j) = cell
//     j*nX + i, the colour map of MakeSurface, FacetCuda.cu:43-72), bilinear
//     filtering with clamp-to-edge, texture coordinates per vertex as
//     Facet.cpp:20-65 assigns them (base edge from u = 0 to 1 at v = 0, the apex
//     at u = lenNeg / baseLength, v = 1);
//   * RGB bytes, rows top to bottom (the reference flips glReadPixels' rows).
// Geometry is clipped against the near and far planes in clip space; texture
// coordinates are interpolated perspective-correctly, depth linearly in
// window space, pixels sampled at their centres.
#include "render.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "hostio.h"
#include "launch.h"
#include "model_ops.h"

namespace sr {

namespace {

struct D3 { double x, y, z; };
inline D3 sub3(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross3(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline D3 unit3(D3 a) {
    const double l = std::sqrt(dot3(a, a));
    return D3{a.x / l, a.y / l, a.z / l};
}

struct M4 { double m[4][4]; };

M4 mul(const M4 &a, const M4 &b) {
    M4 c{};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++) c.m[i][j] += a.m[i][k] * b.m[k][j];
    return c;
}

// glm::lookAt (right-handed): rows s, u, -f.
M4 look_at(D3 eye, D3 centre, D3 up) {
    const D3 f = unit3(sub3(centre, eye)), s = unit3(cross3(f, up)), u = cross3(s, f);
    M4 v{};
    v.m[0][0] = s.x; v.m[0][1] = s.y; v.m[0][2] = s.z; v.m[0][3] = -dot3(s, eye);
    v.m[1][0] = u.x; v.m[1][1] = u.y; v.m[1][2] = u.z; v.m[1][3] = -dot3(u, eye);
    v.m[2][0] = -f.x; v.m[2][1] = -f.y; v.m[2][2] = -f.z; v.m[2][3] = dot3(f, eye);
    v.m[3][3] = 1.0;
    return v;
}

// glm::perspective (right-handed, depth -1..1).
M4 perspective(double fovy, double aspect, double zn, double zf) {
    const double t = std::tan(fovy / 2.0);
    M4 p{};
    p.m[0][0] = 1.0 / (aspect * t);
    p.m[1][1] = 1.0 / t;
    p.m[2][2] = -(zf + zn) / (zf - zn);
    p.m[2][3] = -2.0 * zf * zn / (zf - zn);
    p.m[3][2] = -1.0;
    return p;
}

struct CV { double x, y, z, w, u, v; };  // clip-space vertex with texture coordinates

CV lerp(const CV &a, const CV &b, double t) {
    return CV{a.x + t * (b.x - a.x), a.y + t * (b.y - a.y), a.z + t * (b.z - a.z),
              a.w + t * (b.w - a.w), a.u + t * (b.u - a.u), a.v + t * (b.v - a.v)};
}

// Sutherland-Hodgman against one clip plane d(v) >= 0.
template <class F>
void clip_poly(std::vector<CV> &poly, F dist) {
    std::vector<CV> out;
    const size_t n = poly.size();
    for (size_t i = 0; i < n; i++) {
        const CV &a = poly[i], &b = poly[(i + 1) % n];
        const double da = dist(a), db = dist(b);
        if (da >= 0) out.push_back(a);
        if ((da >= 0) != (db >= 0)) out.push_back(lerp(a, b, da / (da - db)));
    }
    poly.swap(out);
}

struct Texture {
    int nx, ny;
    const float *rgba;  // nx * ny * 4
    void sample(double u, double v, float out[3]) const {
        const double x = u * nx - 0.5, y = v * ny - 0.5;
        const double fx = std::floor(x), fy = std::floor(y);
        const double a = x - fx, b = y - fy;
        int i0 = (int)fx, j0 = (int)fy, i1 = i0 + 1, j1 = j0 + 1;
        i0 = std::clamp(i0, 0, nx - 1); i1 = std::clamp(i1, 0, nx - 1);
        j0 = std::clamp(j0, 0, ny - 1); j1 = std::clamp(j1, 0, ny - 1);
        for (int c = 0; c < 3; c++) {
            const double t00 = rgba[4 * ((size_t)j0 * nx + i0) + c], t10 = rgba[4 * ((size_t)j0 * nx + i1) + c];
            const double t01 = rgba[4 * ((size_t)j1 * nx + i0) + c], t11 = rgba[4 * ((size_t)j1 * nx + i1) + c];
            out[c] = (float)((1 - b) * ((1 - a) * t00 + a * t10) + b * ((1 - a) * t01 + a * t11));
        }
    }
};

inline uint8_t to_u8(float c) { return (uint8_t)std::lround(std::clamp(c, 0.0f, 1.0f) * 255.0f); }

struct Raster {
    int W, H;
    std::vector<float> depth;
    std::vector<uint8_t> rgb;  // window rows bottom to top (GL order)

    void triangle(const CV &a, const CV &b, const CV &c, const Texture &tex) {
        std::vector<CV> poly{a, b, c};
        clip_poly(poly, [](const CV &q) { return q.z + q.w; });  // near
        clip_poly(poly, [](const CV &q) { return q.w - q.z; });  // far
        if (poly.size() < 3) return;
        for (size_t k = 1; k + 1 < poly.size(); k++) fill(poly[0], poly[k], poly[k + 1], tex);
    }

    void fill(const CV &a, const CV &b, const CV &c, const Texture &tex) {
        struct SV { double x, y, z, iw, uw, vw; };
        auto win = [&](const CV &q) {
            const double iw = 1.0 / q.w;
            return SV{(q.x * iw + 1.0) * 0.5 * W, (q.y * iw + 1.0) * 0.5 * H, (q.z * iw + 1.0) * 0.5, iw, q.u * iw,
                      q.v * iw};
        };
        const SV p0 = win(a), p1 = win(b), p2 = win(c);
        double area = (p1.x - p0.x) * (p2.y - p0.y) - (p2.x - p0.x) * (p1.y - p0.y);
        if (!(std::fabs(area) > 0)) return;
        const int x0 = std::max(0, (int)std::floor(std::min({p0.x, p1.x, p2.x})));
        const int x1 = std::min(W - 1, (int)std::ceil(std::max({p0.x, p1.x, p2.x})));
        const int y0 = std::max(0, (int)std::floor(std::min({p0.y, p1.y, p2.y})));
        const int y1 = std::min(H - 1, (int)std::ceil(std::max({p0.y, p1.y, p2.y})));
        const double inv = 1.0 / area;
        for (int y = y0; y <= y1; y++)
            for (int x = x0; x <= x1; x++) {
                const double px = x + 0.5, py = y + 0.5;
                // barycentric weights (signed areas of the opposite sub-triangles)
                const double w0 = ((p1.x - px) * (p2.y - py) - (p2.x - px) * (p1.y - py)) * inv;
                const double w1 = ((p2.x - px) * (p0.y - py) - (p0.x - px) * (p2.y - py)) * inv;
                const double w2 = 1.0 - w0 - w1;
                if (w0 < 0 || w1 < 0 || w2 < 0) continue;
                const double z = w0 * p0.z + w1 * p1.z + w2 * p2.z;
                float &d = depth[(size_t)y * W + x];
                if (!(z < d) || z < 0.0 || z > 1.0) continue;  // GL_LESS, depth range [0, 1]
                d = (float)z;
                const double iw = w0 * p0.iw + w1 * p1.iw + w2 * p2.iw;
                const double u = (w0 * p0.uw + w1 * p1.uw + w2 * p2.uw) / iw;
                const double v = (w0 * p0.vw + w1 * p1.vw + w2 * p2.vw) / iw;
                float col[3];
                tex.sample(u, v, col);
                uint8_t *o = &rgb[3 * ((size_t)y * W + x)];
                o[0] = to_u8(col[0]);
                o[1] = to_u8(col[1]);
                o[2] = to_u8(col[2]);
            }
    }
};

// Texture coordinates of v1, v2, v3 (Facet.cpp:20-65): the base edge runs from
// u = 0 to u = 1 at v = 0, the apex sits at u = lenNeg / (lenNeg + lenPos), v = 1.
void tex_coords(const FacetFrame &F, double uv[6]) {
    const double a = (double)(F.lenNeg / (F.lenPos + F.lenNeg));
    static const double base[3][6] = {{0, 0, 1, 0, -1, 1}, {0, 0, -1, 1, 1, 0}, {-1, 1, 0, 0, 1, 0}};
    for (int k = 0; k < 6; k++) uv[k] = base[F.longest][k] < 0 ? a : base[F.longest][k];
}

void put_u32(std::vector<uint8_t> &b, uint32_t v) {
    b.push_back((uint8_t)(v >> 24)); b.push_back((uint8_t)(v >> 16));
    b.push_back((uint8_t)(v >> 8)); b.push_back((uint8_t)v);
}

void put_chunk(std::vector<uint8_t> &out, const char *type, const uint8_t *data, size_t n) {
    put_u32(out, (uint32_t)n);
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data, data + n);
    put_u32(out, (uint32_t)crc32(0L, out.data() + start, (uInt)(n + 4)));
}

}  // namespace

int render_objects(const std::vector<RenderObject> &objs, int width, int height, const float view[9],
                   std::vector<uint8_t> &rgb) {
    if (width <= 0 || height <= 0 || (long)width * height > (1L << 28)) {
        set_error("render: bad image size");
        return -2;
    }
    const D3 eye{view[0], view[1], view[2]}, centre{view[3], view[4], view[5]}, up{view[6], view[7], view[8]};
    const D3 fwd = sub3(centre, eye);
    if (!(dot3(fwd, fwd) > 0) || !(dot3(cross3(fwd, up), cross3(fwd, up)) > 0)) {
        set_error("render: the camera needs eye != centre and an up vector not along the view direction");
        return -2;
    }
    const M4 pv = mul(perspective(45.0 * M_PI / 180.0, (double)width / height, 0.1, 10000.0), look_at(eye, centre, up));
    Raster R{width, height, std::vector<float>((size_t)width * height, 1.0f),
             std::vector<uint8_t>((size_t)width * height * 3)};
    const uint8_t bg[3] = {to_u8(0.96f), to_u8(0.96f), to_u8(0.86f)};
    for (size_t i = 0; i < (size_t)width * height; i++) std::memcpy(&R.rgb[3 * i], bg, 3);
    for (const RenderObject &o : objs)
        for (size_t f = 0; f < o.frames->size(); f++) {
            const FacetFrame &F = (*o.frames)[f];
            const FacetGrid &G = (*o.grids)[f];
            if (G.nX <= 0 || G.nY <= 0) continue;
            const Texture tex{G.nX, G.nY, o.rgba + 4 * (size_t)(*o.cellOff)[f]};
            double uv[6];
            tex_coords(F, uv);
            CV cv[3];
            const Vec3 vs[3] = {F.v1, F.v2, F.v3};
            for (int k = 0; k < 3; k++) {
                const double X[4] = {vs[k].x, vs[k].y, vs[k].z, 1.0};
                double c[4] = {0, 0, 0, 0};
                for (int i = 0; i < 4; i++)
                    for (int j = 0; j < 4; j++) c[i] += pv.m[i][j] * X[j];
                cv[k] = CV{c[0], c[1], c[2], c[3], uv[2 * k], uv[2 * k + 1]};
            }
            R.triangle(cv[0], cv[1], cv[2], tex);
        }
    rgb.resize((size_t)width * height * 3);
    for (int y = 0; y < height; y++)  // top row first
        std::memcpy(&rgb[(size_t)y * width * 3], &R.rgb[(size_t)(height - 1 - y) * width * 3], (size_t)width * 3);
    return 0;
}

int render_rgb(Model &m, int plane, int width, int height, const float view[9], const RenderSettings &rs,
               std::vector<uint8_t> &rgb) {
    if (!m.rendered) { set_error("render: call render_cuda first"); return -3; }
    std::vector<std::vector<float>> cols;
    std::vector<RenderObject> objs;
    cols.reserve(m.targets.size() + m.fields.size());
    for (int pass = 0; pass < 2; pass++) {  // targets, then field objects (ModelGl.cpp:290-298)
        std::vector<Object> &v = pass == 0 ? m.targets : m.fields;
        const int type = pass == 0 ? OBJ_TARGET : OBJ_FIELD;
        for (int oi = 0; oi < (int)v.size(); oi++) {
            Object &o = v[oi];
            if (!o.ncell) continue;
            std::vector<double> p(2 * (size_t)o.ncell);
            if (int rc = model_get_object_pressure(m, type, oi, 0, plane, p.data(), o.ncell)) return rc;
            cols.emplace_back(4 * (size_t)o.ncell);
            colour_map(p.data(), o.ncell, rs.minDb, rs.maxDb, pass == 0 ? rs.phaseTarget : rs.phaseField,
                       cols.back().data());
            objs.push_back(RenderObject{&o.frames, &o.grids, &o.cellOff, cols.back().data()});
        }
    }
    return render_objects(objs, width, height, view, rgb);
}

int write_png(const std::string &path, int width, int height, const uint8_t *rgb, std::string &err) {
    std::vector<uint8_t> raw((size_t)height * (1 + (size_t)width * 3));
    for (int y = 0; y < height; y++) {
        uint8_t *row = &raw[(size_t)y * (1 + (size_t)width * 3)];
        row[0] = 0;  // filter: none
        std::memcpy(row + 1, rgb + (size_t)y * width * 3, (size_t)width * 3);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) {
        err = "png: compression failed";
        return -1;
    }
    std::vector<uint8_t> out{0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    put_u32(ihdr, (uint32_t)width);
    put_u32(ihdr, (uint32_t)height);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit RGB, deflate, no filter, no interlace
    put_chunk(out, "IHDR", ihdr.data(), ihdr.size());
    put_chunk(out, "IDAT", z.data(), zlen);
    put_chunk(out, "IEND", nullptr, 0);
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) { err = "png: cannot open " + path; return -2; }
    const bool ok = std::fwrite(out.data(), 1, out.size(), fp) == out.size();
    if (std::fclose(fp) != 0 || !ok) { err = "png: write failed: " + path; return -1; }
    return 0;
}

}  // namespace sr
