// The C-ABI: the reference's 13 entry points (src/python_interface.cpp) on one
// process-global scene, plus status-returning sr_* extensions. No exception
// crosses this boundary.
#include <chrono>
#include <cstdio>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "../../include/sonar_render.h"
#include "launch.h"
#include "model_ops.h"
#include "hostio.h"
#include "nearfield.h"
#include "batch.h"
#include "render.h"
#include "noalloc.h"
#include "measure.h"

namespace {

sr::Model gModel;  // python_interface.cpp:11: the default context, handle 0
// created contexts by handle, and each thread's current one (null: the default)
std::mutex gCtxMu;
std::map<int, std::unique_ptr<sr::Model>> gCtx;
int gNextCtx = 1;
thread_local sr::Model *tCur = nullptr;
thread_local int tCurId = 0;
sr::Model &cur() { return tCur ? *tCur : gModel; }
// Status of the calling thread's last call: each thread reads its own, so
// threads working on different contexts never see (or free) each other's message.
thread_local int gStatus = 0;
thread_local std::string gMsg;
thread_local float gLastMs = 0.f;

template <class F>
int guarded(const char *name, F &&f) {
    int rc;
    try {
        sr::clear_error();
        rc = f();
    } catch (const std::exception &e) {
        sr::set_error(std::string(name) + ": " + e.what());
        rc = -1;
    } catch (...) {
        sr::set_error(std::string(name) + ": unknown exception");
        rc = -1;
    }
    gStatus = rc < 0 ? rc : 0;
    gMsg = rc < 0 ? sr::last_error() : std::string();
    return rc;
}

template <class F>
int timed(const char *name, F &&f) {
    auto t0 = std::chrono::steady_clock::now();
    int rc = guarded(name, f);
    gLastMs = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

void report(const char *name, int rc) {
    if (rc < 0) std::fprintf(stderr, "sonar_render: %s failed (%d): %s\n", name, rc, gMsg.c_str());
}

}  // namespace

extern "C" {

const char *sr_last_error(void) { return gMsg.c_str(); }
int sr_last_status(void) { return gStatus; }
float sr_last_call_ms(void) { return gLastMs; }

int sr_set_initial_conditions(float cp, float f, float a, float d) {
    return guarded("set_initial_conditions", [&] { return sr::model_set_initial_conditions(cur(), cp, f, a, d); });
}
int sr_load_geometry(const float *v, int nf, int type, float res) {
    return guarded("load_geometry", [&] { return sr::model_load_geometry(cur(), v, nf, type, res); });
}
int sr_load_field_points(const float *p, int n) {
    return guarded("load_field_points", [&] { return sr::model_load_points(cur(), p, n, false); });
}
int sr_load_source_points(const float *p, int n) {
    return guarded("load_source_points", [&] { return sr::model_load_points(cur(), p, n, true); });
}
int sr_set_batched_sources(int on) {
    return guarded("set_batched_sources", [&] { return sr::model_set_batched_sources(cur(), on); });
}
int sr_render(void) { return timed("render_cuda", [&] { return sr::model_render(cur()); }); }
int sr_project_src_points_to_objects(int fo) {
    return timed("ProjectSrcPointsToObjects", [&] { return sr::model_project_src(cur(), fo != 0); });
}
int sr_project_target_to_field_objects(void) {
    return timed("ProjectTargetToFieldObjects", [&] { return sr::model_project_t2field_objects(cur()); });
}
int sr_project_target_to_field_points(void) {
    return timed("ProjectTargetToFieldPoints", [&] { return sr::model_project_t2field_points(cur()); });
}
int sr_project_target_to_field_points_noalloc(void) {
    return timed("ProjectTargetToFieldPoints (allocation-free)", [&] {
        float2 *scr = nullptr;
        size_t elems = 0;
        if (sr::noalloc_scratch(&scr, &elems)) return -1;
        return sr::model_project_t2field_points_noalloc(cur(), scr, elems);
    });
}
long sr_device_allocations(void) { return sr::model_device_allocations(); }
int sr_project_target_to_target_objects(int n) {
    return timed("ProjectTargetToTargetObjects", [&] { return sr::model_project_t2t(cur(), n); });
}
int sr_prepare_resp(int t) { return guarded("prepare_resp", [&] { return sr::model_prepare_resp(cur(), t); }); }
int sr_project_resp(int a, int b) {
    return timed("project_resp", [&] { return sr::model_project_resp(cur(), a, b); });
}
int sr_get_field_point_pressures(sr_dcomplex *out, int n) {
    return guarded("GetFieldPointPressures", [&] {
        int rc = sr::model_get_field_points(cur(), (double *)out, n, 0);
        return rc < 0 ? rc : 0;
    });
}
int sr_get_field_point_pressures_plane(int plane, sr_dcomplex *out, int n) {
    return guarded("get_field_point_pressures_plane", [&] {
        int rc = sr::model_get_field_points(cur(), (double *)out, n, plane);
        return rc < 0 ? rc : 0;
    });
}
int sr_get_field_point_pressures_c64(float *out, int n) {
    return guarded("get_field_point_pressures_c64", [&] {
        if (n <= 0) return 0;
        double *tmp = new double[2 * (size_t)n];
        int rc = sr::model_get_field_points(cur(), tmp, n, 0);
        for (int i = 0; rc > 0 && i < rc; i++) { out[2 * i] = (float)tmp[2 * i]; out[2 * i + 1] = (float)tmp[2 * i + 1]; }
        delete[] tmp;
        return rc < 0 ? rc : 0;
    });
}
int sr_set_source_amplitude(int s, float re, float im) {
    return guarded("set_source_amplitude", [&] { return sr::model_set_source_amplitude(cur(), s, re, im); });
}
int sr_num_planes(void) { return cur().nplanes; }
int sr_set_source_object_amplitude(int index, float re, float im) {
    return guarded("set_source_object_amplitude",
                   [&] { return sr::model_set_source_object_amplitude(cur(), index, re, im); });
}
int sr_context_create(void) {
    return guarded("context_create", [&] {
        std::lock_guard<std::mutex> g(gCtxMu);
        const int h = gNextCtx++;
        gCtx[h] = std::make_unique<sr::Model>();
        return h;
    });
}
int sr_context_select(int handle) {
    return guarded("context_select", [&] {
        if (handle == 0) {
            tCur = nullptr;
            tCurId = 0;
            return 0;
        }
        std::lock_guard<std::mutex> g(gCtxMu);
        auto it = gCtx.find(handle);
        if (it == gCtx.end()) { sr::set_error("no such context"); return -2; }
        tCur = it->second.get();
        tCurId = handle;
        return 0;
    });
}
int sr_context_current(void) { return tCurId; }
int sr_context_destroy(int handle) {
    return guarded("context_destroy", [&] {
        if (handle == 0) { sr::set_error("the default context cannot be destroyed"); return -2; }
        if (handle == tCurId) { sr::set_error("the calling thread's current context cannot be destroyed"); return -2; }
        std::unique_ptr<sr::Model> m;
        {
            std::lock_guard<std::mutex> g(gCtxMu);
            auto it = gCtx.find(handle);
            if (it == gCtx.end()) { sr::set_error("no such context"); return -2; }
            m = std::move(it->second);
            gCtx.erase(it);
        }
        sr::model_release(*m);
        return 0;
    });
}

int sr_device_synchronize(void) {
    return guarded("device_synchronize", [&] {
        SR_HIP(hipDeviceSynchronize());
        return 0;
    });
}
int sr_teardown(void) {
    return guarded("TearDownCuda", [&] { sr::model_teardown(cur()); return 0; });
}

int sr_object_count(int type) {
    if (type == sr::OBJ_TARGET) return (int)cur().targets.size();
    if (type == sr::OBJ_FIELD) return (int)cur().fields.size();
    if (type == sr::OBJ_SOURCE) return (int)cur().sources.size();
    return -2;
}
static sr::Object *obj(int type, int idx) {
    if (type != sr::OBJ_FIELD && type != sr::OBJ_TARGET && type != sr::OBJ_SOURCE) return nullptr;
    auto &v = (type == sr::OBJ_FIELD) ? cur().fields : (type == sr::OBJ_SOURCE ? cur().sources : cur().targets);
    if (idx < 0 || idx >= (int)v.size()) return nullptr;
    return &v[idx];
}
long sr_object_cell_count(int type, int idx) {
    sr::Object *o = obj(type, idx);
    if (!o || !cur().rendered) return -2;
    return o->ncell;
}
int sr_object_active_count(int type, int idx) {
    sr::Object *o = obj(type, idx);
    if (!o || !cur().rendered) return -2;
    return o->nact;
}
int sr_object_active_cells(int type, int idx, int *out, int n) {
    sr::Object *o = obj(type, idx);
    if (!o || !cur().rendered || n < o->nact) return -2;
    for (int k = 0; k < o->nact; k++) out[k] = o->actCellHost[k];
    return 0;
}
int sr_get_object_pressures(int type, int idx, int which, int plane, sr_dcomplex *out, long n) {
    return guarded("get_object_pressures",
                   [&] { return sr::model_get_object_pressure(cur(), type, idx, which, plane, (double *)out, n); });
}

long sr_mask_words(int kind, int ta, int ia, int tb, int ib) {
    if (!cur().rendered) return -3;
    const long ns = (long)cur().srcHost.size() / 3, nq = (long)cur().fldHost.size() / 3;
    sr::Object *A = obj(ta, ia), *B = obj(tb, ib);
    long ne, nr;
    switch (kind) {
    case 0: if (!B) return -2; ne = ns; nr = B->ncell; break;
    case 1: if (!A) return -2; ne = A->nact; nr = nq; break;
    case 2: if (!A) return -2; ne = nr = A->nact; break;
    case 3: case 4: if (!A || !B) return -2; ne = A->nact; nr = B->nact; break;
    case 5: if (!A || !B) return -2; ne = B->nact; nr = A->nact; break;
    case 6: A = obj(sr::OBJ_SOURCE, ia); if (!A || !B) return -2; ne = A->nact; nr = B->nact; break;
    default: return -2;
    }
    return ((nr + 63) / 64) * ne;
}
int sr_get_mask(int kind, int ta, int ia, int tb, int ib, unsigned long long *out, long nwords) {
    return guarded("get_mask", [&] {
        long need = sr_mask_words(kind, ta, ia, tb, ib);
        if (need < 0) { sr::set_error("bad mask request"); return (int)need; }
        if (nwords < need) { sr::set_error("mask buffer too short"); return -2; }
        return sr::model_get_mask(cur(), kind, ta, ia, tb, ib, (uint64_t *)out, need);
    });
}
int sr_vis_params(float *out) {
    if (!cur().rendered) return -3;
    out[0] = cur().vc.scale; out[1] = cur().vc.padTri; out[2] = cur().vc.padNode;
    out[3] = cur().vc.tminPoint; out[4] = cur().vc.tminFrag;
    return 0;
}

long sr_fragment_facet(const float *v, float delta, int *grid3, float *area, float *pe, float *pr, long cap) {
    if (!v || !grid3 || !(delta > 0.f)) return -2;
    sr::FacetFrame F = sr::facet_frame(sr::Vec3{v[0], v[1], v[2]}, sr::Vec3{v[3], v[4], v[5]}, sr::Vec3{v[6], v[7], v[8]});
    sr::FacetGrid G = sr::facet_grid(F, delta);
    grid3[0] = G.nX; grid3[1] = G.nY; grid3[2] = G.nNeg;
    const long n = (long)G.nX * G.nY;
    if (cap >= n && area && pe && pr) {
        for (int j = 0; j < G.nY; j++)
            for (int i = 0; i < G.nX; i++) {
                const long c = (long)j * G.nX + i;
                area[c] = G.area[c];
                sr::Vec3 a = sr::cell_point(F, G, delta, i, j, false), b = sr::cell_point(F, G, delta, i, j, true);
                pe[3 * c] = a.x; pe[3 * c + 1] = a.y; pe[3 * c + 2] = a.z;
                pr[3 * c] = b.x; pr[3 * c + 1] = b.y; pr[3 * c + 2] = b.z;
            }
    }
    return n;
}

int sr_reset_pressures(void) {
    return guarded("reset_pressures", [&] { return sr::model_reset_pressures(cur()); });
}
int sr_clear_vis_cache(void) {
    return guarded("clear_vis_cache", [&] { return sr::model_clear_vis_cache(cur()); });
}
int sr_set_transfer_mode(int mode) {
    return guarded("set_transfer_mode", [&] { return sr::model_set_transfer_mode(cur(), mode); });
}
int sr_transfer_mode(void) { return cur().transferMode; }
int sr_set_transfer_splits(int n) {
    return guarded("set_transfer_splits", [&] { return sr::model_set_transfer_splits(cur(), n); });
}
int sr_transfer_splits(void) { return cur().transferSplits; }
int sr_set_transfer_symmetric(int on) {
    return guarded("set_transfer_symmetric", [&] { return sr::model_set_transfer_symmetric(cur(), on); });
}
int sr_transfer_symmetric(void) { return cur().transferSymmetric; }
int sr_set_inline_visibility(int on) {
    return guarded("set_inline_visibility", [&] { return sr::model_set_inline_visibility(cur(), on); });
}
int sr_inline_visibility(void) { return cur().inlineVis; }
int sr_set_reduce_in_product(int on) {
    return guarded("set_reduce_in_product", [&] { return sr::model_set_reduce_in_product(cur(), on); });
}
int sr_reduce_in_product(void) { return cur().reduceInGemm; }
int sr_set_gather_mode(int mode) {
    return guarded("set_gather_mode", [&] { return sr::model_set_gather_mode(cur(), mode); });
}
int sr_gather_mode(void) { return cur().gatherMode; }
int sr_set_near_quadrature_radius(float deltas) {
    return guarded("set_near_quadrature_radius", [&] { return sr::model_set_near_quadrature_radius(cur(), deltas); });
}
int sr_set_near_mode(int mode, float radiusWavelengths) {
    return guarded("set_near_mode", [&] { return sr::model_set_near_mode(cur(), mode, radiusWavelengths); });
}
int sr_nf_init(unsigned seed) { return guarded("nf_init", [&] { return sr::nf_init(seed); }); }
int sr_nf_train(int nsamples, int steps, int batch, float lr, unsigned seed, float *report2) {
    return guarded("nf_train", [&] { return sr::nf_train(nsamples, steps, batch, lr, seed, report2); });
}
int sr_nf_get_params(float *out, int n) { return guarded("nf_get_params", [&] { return sr::nf_get_params(out, n); }); }
int sr_nf_set_params(const float *in, int n) {
    return guarded("nf_set_params", [&] { return sr::nf_set_params(in, n); });
}
int sr_nf_samples(unsigned seed, int n, float *X, float *Y) {
    return guarded("nf_samples", [&] { return sr::nf_samples_host(seed, n, X, Y); });
}
int sr_nf_forward(const float *X, int n, float *out) {
    return guarded("nf_forward", [&] { return sr::nf_eval(X, nullptr, n, out, nullptr, nullptr); });
}
int sr_nf_gradient(const float *params, const float *X, const float *Y, int n, float *grad) {
    return guarded("nf_gradient", [&] { return sr::nf_eval(X, Y, n, nullptr, grad, params); });
}
int sr_profile(int on) {
    return guarded("profile", [&] { sr::model_profile(cur(), on); return 0; });
}
int sr_kernel_stats(int kind, double *ms, long *launches, double *pairs, double *flops, double *bytes) {
    return guarded("kernel_stats", [&] {
        if (kind < 0 || kind >= sr::PK_COUNT) { sr::set_error("unknown kernel family"); return -2; }
        return sr::model_kernel_stats(cur(), kind, ms, launches, pairs, flops, bytes);
    });
}

int sr_time_transfer_gemm(int reps, double *msPerLaunch) {
    return guarded("time_transfer_gemm", [&] { return sr::model_time_transfer_gemm(cur(), reps, msPerLaunch); });
}
int sr_measure_stream_read(long bytes, int reps, double *gbs) {
    return guarded("measure_stream_read", [&] { return sr::measure_stream_read(bytes, reps, gbs); });
}
int sr_measure_stream_write(long bytes, int reps, int nontemporal, double *gbs) {
    return guarded("measure_stream_write", [&] { return sr::measure_stream_write(bytes, reps, nontemporal, gbs); });
}
int sr_measure_valu_issue(int reps, double *ginstr) {
    return guarded("measure_valu_issue", [&] { return sr::measure_valu_issue(reps, ginstr); });
}

long sr_read_stl(const char *path, float *out, long cap) {
    long rc = -1;
    guarded("read_stl", [&] {
        std::vector<float> t;
        std::string err;
        int n = sr::stl_read(path ? path : "", t, err);
        if (n < 0) { sr::set_error(err); rc = n; return n; }
        if (out && cap >= n) std::copy(t.begin(), t.end(), out);
        rc = n;
        return 0;
    });
    return rc;
}
int sr_write_stl(const char *path, const float *v, int n) {
    return guarded("write_stl", [&] {
        std::string err;
        int rc = sr::stl_write(path ? path : "", v, n, err);
        if (rc) sr::set_error(err);
        return rc;
    });
}
int sr_load_stl(const char *path, int type, float res) {
    return guarded("load_stl", [&] {
        std::vector<float> t;
        std::string err;
        int n = sr::stl_read(path ? path : "", t, err);
        if (n < 0) { sr::set_error(err); return n; }
        return sr::model_load_geometry(cur(), t.data(), n, type, res);
    });
}
int sr_colour_map(const double *p, long n, float minDb, float maxDb, int phase, float *rgba) {
    if (n < 0 || (n > 0 && (!p || !rgba))) return -2;
    sr::colour_map(p, n, minDb, maxDb, phase != 0, rgba);
    return 0;
}
int sr_surface_rgba(int type, int idx, int plane, float *rgba, long n) {
    return guarded("surface_rgba", [&] {
        std::vector<double> p((size_t)2 * std::max<long>(n, 0));
        int rc = sr::model_get_object_pressure(cur(), type, idx, 0, plane, p.data(), n);
        if (rc) return rc;
        const sr::Model &m = cur();
        const bool ph = (type == sr::OBJ_FIELD) ? m.renderPhaseField : m.renderPhaseTarget;
        sr::colour_map(p.data(), n, m.renderMinDb, m.renderMaxDb, ph, rgba);
        return 0;
    });
}

static sr::RenderSettings render_settings(const sr::Model &m) {
    sr::RenderSettings rs;
    rs.maxDb = m.renderMaxDb;
    rs.minDb = m.renderMinDb;
    rs.phaseTarget = m.renderPhaseTarget;
    rs.phaseField = m.renderPhaseField;
    return rs;
}
int sr_sound_visualisation_init(float maxDb, float minDb, int phaseTarget, int phaseField) {
    return guarded("SoundVisualisationInit", [&] {
        if (!(maxDb > minDb)) { sr::set_error("SoundVisualisationInit: maxDb must exceed minDb"); return -2; }
        sr::Model &m = cur();
        m.renderMaxDb = maxDb;
        m.renderMinDb = minDb;
        m.renderPhaseTarget = phaseTarget != 0;
        m.renderPhaseField = phaseField != 0;
        return 0;
    });
}
int sr_render_rgb(int plane, int width, int height, const float *view9, unsigned char *rgb) {
    return guarded("render_rgb", [&] {
        if (!view9 || !rgb) { sr::set_error("render_rgb: null argument"); return -2; }
        std::vector<uint8_t> img;
        int rc = sr::render_rgb(cur(), plane, width, height, view9, render_settings(cur()), img);
        if (rc) return rc;
        std::copy(img.begin(), img.end(), rgb);
        return 0;
    });
}
int sr_render_png(const char *filename, int width, int height, const float *view9, int plane) {
    return guarded("RenderOpenGL", [&] {
        if (!filename || !view9) { sr::set_error("RenderOpenGL: null file name or view settings"); return -2; }
        std::vector<uint8_t> img;
        int rc = sr::render_rgb(cur(), plane, width, height, view9, render_settings(cur()), img);
        if (rc) return rc;
        std::string err;
        rc = sr::write_png(filename, width, height, img.data(), err);
        if (rc) sr::set_error(err);
        return rc;
    });
}
int sr_render_facets_rgb(const float *v9, int nf, float delta, const double *p2, long ncell, int phase, float minDb,
                         float maxDb, int width, int height, const float *view9, unsigned char *rgb) {
    return guarded("render_facets_rgb", [&] {
        if (nf < 0 || (nf > 0 && !v9) || !(delta > 0.f) || !view9 || !rgb) {
            sr::set_error("render_facets_rgb: bad argument");
            return -2;
        }
        std::vector<sr::FacetFrame> frames;
        std::vector<sr::FacetGrid> grids;
        std::vector<long> off;
        long tot = 0;
        for (int f = 0; f < nf; f++) {
            const float *t = v9 + (size_t)9 * f;
            frames.push_back(sr::facet_frame(sr::Vec3{t[0], t[1], t[2]}, sr::Vec3{t[3], t[4], t[5]},
                                             sr::Vec3{t[6], t[7], t[8]}));
            grids.push_back(sr::facet_grid(frames.back(), delta));
            off.push_back(tot);
            tot += (long)grids.back().nX * grids.back().nY;
        }
        if (ncell != tot || (tot > 0 && !p2)) {
            sr::set_error("render_facets_rgb: the pressure array must hold every cell of every facet");
            return -2;
        }
        std::vector<float> col(4 * (size_t)tot);
        sr::colour_map(p2, tot, minDb, maxDb, phase != 0, col.data());
        std::vector<sr::RenderObject> objs{sr::RenderObject{&frames, &grids, &off, col.data()}};
        std::vector<uint8_t> img;
        int rc = sr::render_objects(objs, width, height, view9, img);
        if (rc) return rc;
        std::copy(img.begin(), img.end(), rgb);
        return 0;
    });
}
int sr_write_png(const char *path, int width, int height, const unsigned char *rgb) {
    return guarded("write_png", [&] {
        if (!path || !rgb || width <= 0 || height <= 0) { sr::set_error("write_png: bad argument"); return -2; }
        std::string err;
        int rc = sr::write_png(path, width, height, rgb, err);
        if (rc) sr::set_error(err);
        return rc;
    });
}

int sr_batch_clear(void) { return guarded("batch_clear", [&] { return sr::batch_clear(); }); }
int sr_batch_add_scene(const float *tris, int ntri, float delta, float cp, float f, float atten, const float *src,
                       int nsrc, const float *pts, int npts) {
    int idx = -1;
    int rc = guarded("batch_add_scene", [&] {
        idx = sr::batch_add_scene(tris, ntri, delta, cp, f, atten, src, nsrc, pts, npts);
        return idx < 0 ? idx : 0;
    });
    return rc < 0 ? rc : idx;
}
int sr_batch_set_bounces(int n) { return guarded("batch_set_bounces", [&] { return sr::batch_set_bounces(n); }); }
int sr_batch_set_near(int mode, float radiusWavelengths, float quadDeltas) {
    return guarded("batch_set_near", [&] { return sr::batch_set_near(mode, radiusWavelengths, quadDeltas); });
}
int sr_batch_prepare(void) { return guarded("batch_prepare", [&] { return sr::batch_prepare(); }); }
int sr_batch_run(int useGraph, float *ms) { return guarded("batch_run", [&] { return sr::batch_run(useGraph, ms); }); }
int sr_batch_autotune(float *report, int n) {
    int k = 0;
    int rc = guarded("batch_autotune", [&] {
        k = sr::batch_autotune(report, n);
        return k < 0 ? k : 0;
    });
    return rc < 0 ? rc : k;
}
int sr_batch_field_points(int scene, sr_dcomplex *out, int n) {
    return guarded("batch_field_points", [&] {
        int rc = sr::batch_field_points(scene, (double *)out, n);
        return rc < 0 ? rc : 0;
    });
}
double sr_batch_pairs(void) { return sr::batch_pairs(); }
int sr_batch_get_tune(int *out4) { return sr::batch_get_tune(out4); }
int sr_batch_set_tune(const int *in4) { return guarded("batch_set_tune", [&] { return sr::batch_set_tune(in4); }); }

// ---- the reference's void entry points ----
void set_initial_conditions(float cp, float f, float a, float d) { report("set_initial_conditions", sr_set_initial_conditions(cp, f, a, d)); }
void load_geometry(float *v, int nf, int type, float res) { report("load_geometry", sr_load_geometry(v, nf, type, res)); }
void load_field_points(float *p, int n) { report("load_field_points", sr_load_field_points(p, n)); }
void load_source_points(float *p, int n) { report("load_source_points", sr_load_source_points(p, n)); }
void render_cuda(void) { report("render_cuda", sr_render()); }
void ProjectSrcPointsToObjects(bool fo) { report("ProjectSrcPointsToObjects", sr_project_src_points_to_objects(fo)); }
void ProjectTargetToFieldObjects(void) { report("ProjectTargetToFieldObjects", sr_project_target_to_field_objects()); }
void ProjectTargetToFieldPoints(void) { report("ProjectTargetToFieldPoints", sr_project_target_to_field_points()); }
void ProjectTargetToTargetObjects(int n) { report("ProjectTargetToTargetObjects", sr_project_target_to_target_objects(n)); }
void GetFieldPointPressures(sr_dcomplex *out, int n) { report("GetFieldPointPressures", sr_get_field_point_pressures(out, n)); }
void RenderOpenGL(int width, int height, char *filename, float viewSettings[9]) {
    report("RenderOpenGL", sr_render_png(filename, width, height, viewSettings, 0));
}
void SoundVisualisationInit(float maxDb, float minDb, bool phaseTarget, bool phaseField) {
    report("SoundVisualisationInit", sr_sound_visualisation_init(maxDb, minDb, phaseTarget, phaseField));
}
void TearDownCuda(void) { report("TearDownCuda", sr_teardown()); }

}  // extern "C"
