/*
 * sonar_render.h: C ABI of the MI355X discrete Kirchhoff engine.
 *
 * Part 1 is the reference's own boundary, symbol for symbol, with the same
 * argument types and meanings as src/python_interface.cpp, so a ctypes caller
 * written for the reference (run_scripts/modules/api.py) binds unchanged.
 * These entry points return void like the reference; instead of letting a
 * C++ exception cross the boundary they record the error (sr_last_error) and
 * print it to stderr.
 *
 * Part 2 adds status-returning twins and the batched-incidence, readback and
 * visibility-mask entry points. Every sr_* function returns 0 on success or a
 * negative status: -1 device/runtime failure, -2 bad argument, -3 wrong call
 * order.
 *
 * Threading: the reference keeps one process-global scene
 * (python_interface.cpp:11). Here that scene is context 0, and further
 * independent scenes are contexts (sr_context_*): every entry point acts on
 * the calling thread's current context, and one context must not be used by
 * two threads at once. The status and message of a failed call
 * (sr_last_status, sr_last_error) belong to the calling thread.
 * Projections return once their kernels are queued on the context's stream;
 * the calls that read results (GetFieldPointPressures, sr_get_*) wait for
 * that work and report any error it raised.
 */
#ifndef SONAR_RENDER_H
#define SONAR_RENDER_H

#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* complex128 as the reference lays it out: src/dcomplex.h:12-15 */
typedef struct { double r, i; } sr_dcomplex;

/* ---- Part 1: the reference's entry points ---------------------------- */

/* python_interface.cpp:21-24 (Model.cpp:44-54). omega and k are rounded to
 * FP32; attenuation becomes Im k unchanged; density is unused. */
void set_initial_conditions(float cp, float frequency, float attenuation, float density);

/* python_interface.cpp:26-41. v = num_facets x 9 floats (v1, v2, v3);
 * objectType 0 target, 1 source, 2 field; resolution is the fragment edge
 * delta. A source object (stored but never used by the reference,
 * Model.cpp:33-35) is the piston-like emitter of README.md:2: every fragment
 * at one coherent amplitude (1+0i, sr_set_source_object_amplitude),
 * projected by ProjectSrcPointsToObjects onto the targets (and field objects)
 * with the fragment->fragment sum without reciprocity. It never occludes, and
 * needs unbatched source points. */
void load_geometry(float *v, int num_facets, int objectType, float resolution);

/* python_interface.cpp:43-50: n x 3 floats, appended. */
void load_field_points(float *xyz, int num_field_points);

/* python_interface.cpp:52-59: n x 3 floats, amplitude 1+0i (Model.cpp:18). */
void load_source_points(float *xyz, int num_source_points);

/* python_interface.cpp:61-65: fragment every object, build the LBVH over the
 * target triangles, upload. Computes no pressures. */
void render_cuda(void);

/* python_interface.cpp:67-70: incident field onto targets (and onto field
 * objects if projectFieldObjects). */
void ProjectSrcPointsToObjects(bool projectFieldObjects);

/* python_interface.cpp:72-75: targets -> field objects (FACET_NO_RESP). */
void ProjectTargetToFieldObjects(void);

/* python_interface.cpp:77-80: targets -> field points (FIELD_POINTS). */
void ProjectTargetToFieldPoints(void);

/* python_interface.cpp:82-85: number_reflections FACET_SELF bounces per
 * target object. */
void ProjectTargetToTargetObjects(int number_reflections);

/* python_interface.cpp:87-90: copies min(NumPoints, #field points) complex128
 * values (the reference ignores NumPoints, ModelCuda.cpp:206). */
void GetFieldPointPressures(sr_dcomplex *field_points_pressure, int NumPoints);

/* python_interface.cpp:92-96 (ModelGl.cpp:14-341): writes a PNG of the target
 * objects then the field objects, each facet textured with its cells' colours,
 * seen through a 45-degree perspective camera; viewSettings = eye (3), look-at
 * point (3), up vector (3). Rendered on the host (no OpenGL): sr_render_png. */
void RenderOpenGL(int width, int height, char *filename, float viewSettings[9]);

/* python_interface.cpp:103-107: colour range (dB) and phase flags of the
 * render and of sr_surface_rgba, per context (defaults -20 / -90 dB and no
 * phase, ModelCuda.hpp:45-48). */
void SoundVisualisationInit(float maxDbValue, float minDbValue, bool renderPhaseTarget, bool renderPhaseField);

/* python_interface.cpp:98-101: free the scene; the library can be loaded
 * with a new scene afterwards. */
void TearDownCuda(void);

/* ---- Part 2: extensions (no counterpart in the reference) ------------- */

/* Message of the last failed call ("" if none) and its status. */
const char *sr_last_error(void);
int sr_last_status(void);

/* Status-returning twins of Part 1. */
int sr_set_initial_conditions(float cp, float frequency, float attenuation, float density);
int sr_load_geometry(const float *v, int num_facets, int objectType, float resolution);
int sr_load_field_points(const float *xyz, int n);
int sr_load_source_points(const float *xyz, int n);
int sr_render(void);
int sr_project_src_points_to_objects(int projectFieldObjects);
int sr_project_target_to_field_objects(void);
int sr_project_target_to_field_points(void);
int sr_project_target_to_target_objects(int number_reflections);
int sr_get_field_point_pressures(sr_dcomplex *out, int n);
/* complex64 read-back: 2 floats per point. */
int sr_get_field_point_pressures_c64(float *out, int n);
int sr_teardown(void);

/* Batched incidences: before render, make every source point its own
 * incidence ("plane") instead of summing them coherently. Every projection
 * then runs over all planes in one pass; plane p is what a separate
 * reference run with only source p would produce. */
int sr_set_batched_sources(int on);
int sr_num_planes(void);
int sr_get_field_point_pressures_plane(int plane, sr_dcomplex *out, int n);
int sr_set_source_amplitude(int source, float re, float im);
/* Coherent amplitude of source object `index` (default 1+0i). */
int sr_set_source_object_amplitude(int index, float re, float im);

/* FACET_RESP between two target objects (OptiX.cu:427-430; the reference
 * never dispatches it). sr_prepare_resp sets P_in <- P and P_out <- 0. */
int sr_prepare_resp(int target);
int sr_project_resp(int targetA, int targetB);

/* Scene inspection. objectType 0 target, 2 field. which: 0 P, 1 P_in, 2 P_out.
 * Pressures are per cell in facet order, cell (i,j) of a facet at j*nX+i. */
int sr_object_count(int objectType);
long sr_object_cell_count(int objectType, int index);
int sr_object_active_count(int objectType, int index);
int sr_object_active_cells(int objectType, int index, int *out, int n);
int sr_get_object_pressures(int objectType, int index, int which, int plane, sr_dcomplex *out, long n);

/* Visibility words. kind: 0 sources->cells of (tb,ib), 1 active cells of
 * (ta,ia)->field points, 2 SELF of (ta,ia), 3 NO_RESP (ta,ia)->(tb,ib),
 * 4 RESP forward, 5 RESP reciprocal, 6 source object ia -> object (tb,ib).
 * Word [rblock][emitter], bit = receiver rblock*64+bit contributes. */
long sr_mask_words(int kind, int ta, int ia, int tb, int ib);
int sr_get_mask(int kind, int ta, int ia, int tb, int ib, unsigned long long *out, long nwords);

/* Visibility constants in use: scale, padTri, padNode, tminPoint, tminFrag. */
int sr_vis_params(float *out5);

/* Host-only fragmenter probe (no GPU): fragment one triangle with edge delta.
 * Writes nX, nY, nNeg to grid3; if cap >= nX*nY also the cell areas, the
 * emitter points (3 floats per cell) and the point->fragment receiver points.
 * Returns the cell count, or a negative status. */
long sr_fragment_facet(const float *v9, float delta, int *grid3, float *area, float *pe, float *pr, long cap);

/* Zero every pressure buffer (P, P_in, P_out, field points) of the loaded
 * scene, and drop the cached visibility words (rebuilt on next use). */
int sr_reset_pressures(void);
int sr_clear_vis_cache(void);

/* Bounce arithmetic. 0 (default): plain FP32 gathers, the pair kernel
 * evaluated on the fly for every bounce. 1: the pair kernel is stored once
 * per geometry and frequency as a complex FP16 matrix and every bounce is a
 * FP16 MFMA GEMM with FP32 accumulation over all incidence planes. */
int sr_set_transfer_mode(int mode);
int sr_transfer_mode(void);
/* Split-K of the transfer form's bounce product: each 128-row group's
 * non-zero K chunks are dealt out evenly over n splits (n workgroups per group,
 * partial sums reduced after). 0 (default): chosen from the geometry and the
 * GPU's CU count when the transfer matrix is prepared. Takes effect at the next
 * preparation (a changed value rebuilds the matrix); n up to 4096, capped at
 * the group's chunk count. */
int sr_set_transfer_splits(int n);
int sr_transfer_splits(void);
/* How the transfer matrix S is built (no counterpart in the reference, whose
 * pair kernel has no stored matrix). 1: where the rows and emitters are padded
 * alike and no near-field pairs are listed, only the 64 x 64 blocks on and
 * above the diagonal are evaluated and each is also stored transposed (S is
 * symmetric bit for bit); elsewhere, and with 0, every block is evaluated.
 * 2 (default): as 1, and in a scene of at most 4 occluder triangles that
 * build evaluates the object's own fragment-pair visibility itself (the same
 * FP32 test as the visibility words, which are then not built for it). All
 * give the same matrix; a changed value rebuilds it. */
int sr_set_transfer_symmetric(int on);
int sr_transfer_symmetric(void);

/* Point->fragment sums of the single-scene model with one incidence plane per
 * source (no counterpart in the reference, whose sums read a visibility
 * buffer filled beforehand: src/OptiX.cu:35-110). 1 (default): in a scene of
 * at most 4 occluder triangles the sum evaluates each (source, fragment)
 * pair's visibility itself, the same FP32 test as the sources->fragments
 * visibility words, which are then not built; 0: always from those words.
 * Both give the same result bit for bit. Reset to 1 by TearDownCuda. */
int sr_set_inline_visibility(int on);
int sr_inline_visibility(void);

/* Split-K reduction of the transfer product (no counterpart in the reference,
 * whose bounce sums each receiver's terms in one thread: src/OptiX.cu:297-354).
 * 1: the last workgroup of each 128-row group to finish reduces the group's
 * partial sums, applies the zero-input gate and writes P_out and P; 0
 * (default): the same reduction as a pass of its own. Both sum the splits in
 * the same order: the same pressures bit for bit. The handover between
 * workgroups needs device-scope release and acquire fences, which on MI355X
 * write back and invalidate an XCD's L2: slower than the pass. Reset to 0 by
 * TearDownCuda. */
int sr_set_reduce_in_product(int on);
int sr_reduce_in_product(void);

/* Plain fragment->fragment sums (transfer mode 0). 0 (default): with 16 or
 * more incidence planes the pair kernel is evaluated in FP32 on the vector
 * units and the complex products accumulate on the FP32 matrix cores, 64
 * planes per pass; 1: vector units only (16 planes per pass). Both are FP32
 * throughout. Reset to 0 by TearDownCuda. */
int sr_set_gather_mode(int mode);
int sr_gather_mode(void);

/* Near-field terms of the transfer matrix (mode 1 transfer only; pairs of
 * fragments closer than radiusWavelengths): 0 the reference's point rule,
 * 1 16x16 sub-fragment quadrature of the same kernel over the emitter's
 * cell, 2 the learned correction (a 10-64-64-2 MLP, BF16 inference). */
int sr_set_near_mode(int mode, float radiusWavelengths);
/* Learned mode: pairs closer than `deltas` fragment sizes (1 to 16, default 3)
 * use the quadrature instead of the network. The network is trained on pairs
 * at 3 or more fragment sizes; a smaller radius applies it outside that
 * range. Reset to 3 by TearDownCuda. */
int sr_set_near_quadrature_radius(float deltas);

/* The learned correction's network: trained on the GPU with the library's own
 * forward/backward/Adam kernels on pairs labelled by the quadrature.
 * report2: relative RMS error on training and held-out pairs. Parameters are
 * 4994 floats (W1[64][10] b1 W2[64][64] b2 W3[2][64] b3). sr_nf_samples
 * returns labelled pairs (10 features, 2 targets each); sr_nf_forward the FP32
 * network output; sr_nf_gradient the loss gradient of a given parameter set
 * on given pairs (L = sum |out - y|^2 / n). */
int sr_nf_init(unsigned seed);
int sr_nf_train(int nsamples, int steps, int batch, float lr, unsigned seed, float *report2);
int sr_nf_get_params(float *out, int n);
int sr_nf_set_params(const float *in, int n);
int sr_nf_samples(unsigned seed, int n, float *X, float *Y);
int sr_nf_forward(const float *X, int n, float *out);
int sr_nf_gradient(const float *params, const float *X, const float *Y, int n, float *grad);

/* Per-launch profiling with HIP events on the library's stream. sr_profile(1)
 * starts (and clears) the record. Kernel families: 0 point->fragment,
 * 1 visibility words, 2 fragment->fragment, 3 fragment->field object,
 * 4 fragment->field point, 5 split-K reductions, 6 transfer-matrix build,
 * 7 transfer GEMM, 8 incidence-matrix build, 9 exact recomputation of the
 * rows a zero input gates (matrix-core forms), 10 near-field terms of the
 * transfer matrix (quadrature / learned network; pairs = near pairs). Reports summed kernel time,
 * launches, ordered pairs (x planes), flops (vector gathers: 37 FP32
 * flops per pair plus 8 per pair and plane; matrix-core fragment->fragment:
 * the 8 per pair and plane that run on the matrix cores; transfer GEMM: 2 M N K as executed;
 * transfer-matrix build: the pair evaluations executed, padding included) and
 * algorithmic HBM bytes (transfer GEMM: S, B and the partial sums). */
int sr_profile(int on);
int sr_kernel_stats(int kind, double *ms, long *launches, double *pairs, double *flops, double *bytes);
/* Measurement hook with no counterpart in the reference: replays the last
 * transfer-form bounce product (k_transfer_gemm over the same S and B) `reps`
 * times back to back and returns the mean time per launch, free of the
 * per-launch event overhead of sr_kernel_stats. Overwrites only scratch. */
int sr_time_transfer_gemm(int reps, double *msPerLaunch);
/* Measurement hook with no counterpart in the reference: the attainable HBM
 * read bandwidth of this GPU, a plain streaming read of `bytes` (>= 1 MiB,
 * its own buffer, allocated and freed here) timed over `reps` passes after two
 * warm-up passes; gbs = bytes / mean pass time in GB/s. */
int sr_measure_stream_read(long bytes, int reps, double *gbs);
/* The same for writes (no counterpart in the reference): a streaming write of
 * `bytes` with 16-byte stores, non-temporal (nontemporal = 1, past the caches)
 * or ordinary (0). */
int sr_measure_stream_write(long bytes, int reps, int nontemporal, double *gbs);
/* The attainable vector-instruction issue rate of this GPU (no counterpart in
 * the reference): every SIMD full of waves running independent integer vector
 * instructions, timed over `reps` launches; ginstr in G wave-instructions/s. */
int sr_measure_valu_issue(int reps, double *ginstr);

/* STL meshes (binary or ASCII in, binary out): 9 floats per triangle. */
long sr_read_stl(const char *path, float *out9, long cap);
int sr_write_stl(const char *path, const float *v9, int n);
int sr_load_stl(const char *path, int objectType, float resolution);

/* Surface colour map of src/FacetCuda.cu:43-72: brightness from
 * 20 log10(|p| + 1e-8) between min_dB and max_dB, grey with saturation 0.8 or
 * (phase) hue (arg p + pi) / 2 pi on an HSV wheel. rgba: 4 floats per value. */
int sr_colour_map(const double *p2, long n, float minDb, float maxDb, int phase, float *rgba);
int sr_surface_rgba(int objectType, int index, int plane, float *rgba, long n);

/* Host render (RenderOpenGL's picture). sr_render_png writes incidence plane
 * `plane` of the current scene to a PNG file; sr_render_rgb returns the same
 * image as width x height x 3 bytes, rows top to bottom. sr_render_facets_rgb
 * renders one object given as triangles, fragment edge delta and complex
 * per-cell pressures (p2: 2 doubles per cell, cells of facet f after those of
 * facets 0..f-1, cell (i,j) at j*nX+i) with the given colour settings, on the
 * host only. sr_write_png writes an RGB image. */
int sr_sound_visualisation_init(float maxDb, float minDb, int phaseTarget, int phaseField);
int sr_render_png(const char *filename, int width, int height, const float *view9, int plane);
int sr_render_rgb(int plane, int width, int height, const float *view9, unsigned char *rgb);
int sr_render_facets_rgb(const float *v9, int nf, float delta, const double *p2, long ncell, int phase, float minDb,
                         float maxDb, int width, int height, const float *view9, unsigned char *rgb);
int sr_write_png(const char *path, int width, int height, const unsigned char *rgb);

/* Batched independent scenes: each its own mesh (target), fragment edge,
 * sound speed, frequency, attenuation, sources (summed) and field points.
 * sr_batch_prepare fragments and uploads them; sr_batch_run projects all of
 * them (sources -> fragments, n bounces, fragments -> field points) in one
 * sequence of persistent multi-scene launches, replayed as one hipGraph when
 * useGraph is 1. sr_batch_autotune times candidate tilings (persistent grid,
 * emitters per visibility item, emitters per gather item) and keeps the
 * fastest; report gets 5 floats per candidate (ms, vis grid, gather grid,
 * emitters per vis item, emitters per gather item). Tune: the same 4 ints. */
int sr_batch_clear(void);
int sr_batch_add_scene(const float *tris9, int ntri, float delta, float cp, float frequency, float attenuation,
                       const float *src, int nsrc, const float *pts, int npts);
int sr_batch_set_bounces(int n);
/* Near-field terms of the batched bounces (before sr_batch_prepare): mode 0 the
 * reference's point rule (default), 1 the 16x16 sub-fragment quadrature, 2 the
 * learned correction (needs sr_nf_train; pairs closer than quadDeltas fragment
 * sizes use the quadrature), for visible pairs closer than radiusWavelengths. */
int sr_batch_set_near(int mode, float radiusWavelengths, float quadDeltas);
int sr_batch_prepare(void);
int sr_batch_run(int useGraph, float *ms);
int sr_batch_autotune(float *report, int nmax);
int sr_batch_field_points(int scene, sr_dcomplex *out, int n);
double sr_batch_pairs(void);
int sr_batch_get_tune(int *out4);
int sr_batch_set_tune(const int *in4);

/* ProjectTargetToFieldPoints that allocates nothing. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops). Its split-K partial sums go to a fixed
 * 64 MiB device array per GPU (two halves of 4 Mi complex FP32 values) instead of
 * the library's scratch; the field-point visibility words it needs are allocated
 * by render_cuda. It uses fewer emitter splits when the partials would not fit,
 * and fails with -2 when not even one split fits (more than 65,536 field points
 * in the 64-plane form). The array is shared by every call on that GPU and keeps
 * the last call's contents: calls that overlap on one GPU (other threads'
 * contexts) race and must not run concurrently. Results equal
 * ProjectTargetToFieldPoints' whenever its partial sums fit the array. */
int sr_project_target_to_field_points_noalloc(void);
/* Device allocations the scene model has made so far in this process. */
long sr_device_allocations(void);

/* hipDeviceSynchronize on the library's runtime. */
int sr_device_synchronize(void);

/* Contexts: independent scenes (geometry, sources, field points, settings, device
 * buffers, streams), one current per calling thread; every entry point above acts
 * on the calling thread's current context. Handle 0 is the default context (the
 * reference's single global Model, python_interface.cpp:11) and stays current
 * until a thread selects another. Shared by all contexts: the learned near-field
 * network (sr_nf_*) and the batched scenes (sr_batch_*). Replaces the
 * reference's one-model process (python_interface.cpp:11), per the build
 * contract of one context per handle.
 * sr_context_create: a new empty context, returns its handle (> 0) or a negative status.
 * sr_context_select: makes handle current for the calling thread (0: the default).
 * sr_context_current: the calling thread's current handle.
 * sr_context_destroy: tears a created context down and frees it; not 0, not the
 *   calling thread's current one (another thread must not be using it). */
int sr_context_create(void);
int sr_context_select(int handle);
int sr_context_current(void);
int sr_context_destroy(int handle);

/* Wall time of the last projection entry point, milliseconds. */
float sr_last_call_ms(void);

#ifdef __cplusplus
}
#endif

#endif /* SONAR_RENDER_H */
