// Many independent scenes (own mesh, frequency, source and field points) in
// one pass: persistent multi-scene kernels, optionally replayed as one hipGraph.
#pragma once
#include <cstdint>

namespace sr {

struct BatchTune {
    int visGrid = 1024;       // persistent workgroups of the visibility pass
    int gatherGrid = 1024;    // persistent workgroups of the gathers
    int visEmPerWave = 16;    // emitters per visibility work item
    int minPerSplit = 128;    // emitters per gather work item (split-K granularity)
};

int batch_clear();
int batch_add_scene(const float *tris, int ntri, float delta, float cp, float f, float atten, const float *src,
                    int nsrc, const float *pts, int npts);
int batch_set_bounces(int n);
int batch_set_near(int mode, float radiusWavelengths, float quadDeltas);
int batch_prepare();
int batch_run(int useGraph, float *ms);
int batch_autotune(float *report, int nreport);
int batch_field_points(int scene, double *out, int n);
double batch_pairs();
int batch_get_tune(int *out4);
int batch_set_tune(const int *in4);

}  // namespace sr
