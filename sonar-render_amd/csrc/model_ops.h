// Operations on the scene model; each returns 0 on success or a negative
// status and records a message retrievable with last_error().
#pragma once
#include <cstdint>
#include <string>

#include "model.h"

namespace sr {

std::string last_error();
void clear_error();

int model_set_initial_conditions(Model &m, float cp, float f, float att, float density);
int model_load_geometry(Model &m, const float *v, int nf, int type, float res);
int model_load_points(Model &m, const float *p, int n, bool source);
int model_set_batched_sources(Model &m, int on);
int model_render(Model &m);
int model_set_source_amplitude(Model &m, int i, float re, float im);
int model_project_src(Model &m, bool fieldObjects);
int model_set_source_object_amplitude(Model &m, int i, float re, float im);
int model_project_t2t(Model &m, int nref);
int model_project_t2field_objects(Model &m);
int model_project_t2field_points(Model &m);
int model_project_t2field_points_noalloc(Model &m, float2 *scratch, size_t elems);
// Device allocations the model has made so far (process-wide count).
long model_device_allocations();
int model_project_resp(Model &m, int ia, int ib);
int model_prepare_resp(Model &m, int it);
int model_get_field_points(Model &m, double *out, long n, int plane);
int model_get_object_pressure(Model &m, int type, int idx, int which, int plane, double *out, long n);
int model_get_mask(Model &m, int kind, int ta, int ia, int tb, int ib, uint64_t *out, long nwords);
int model_reset_pressures(Model &m);
int model_clear_vis_cache(Model &m);
int model_set_transfer_mode(Model &m, int mode);
int model_set_transfer_splits(Model &m, int n);
int model_set_transfer_symmetric(Model &m, int on);
int model_set_inline_visibility(Model &m, int on);
int model_set_reduce_in_product(Model &m, int on);
int model_set_gather_mode(Model &m, int mode);
int model_set_near_quadrature_radius(Model &m, float deltas);
int model_set_near_mode(Model &m, int mode, float radiusWl);
void model_profile(Model &m, int on);
void model_profile_reset(Model &m);
int model_kernel_stats(Model &m, int kind, double *ms, long *launches, double *pairs, double *flops, double *bytes);
int model_time_transfer_gemm(Model &m, int reps, double *msPerLaunch);
void model_teardown(Model &m);
// Teardown plus the context's own streams, events and pinned words (a context that is discarded).
void model_release(Model &m);

// Visibility-word task of one pair set (kinds as sr_get_mask), without launching it.
int mask_task(Model &m, int kind, int ta, int ia, int tb, int ib, MaskTask &tk, size_t &words);

}  // namespace sr
