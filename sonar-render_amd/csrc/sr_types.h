// Plain-data descriptors shared by the host model and the HIP kernels.
// No torch types, no classes with behaviour: these cross the internal
// host -> kernel-launcher boundary by value.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sr {

enum ObjType { OBJ_TARGET = 0, OBJ_SOURCE = 1, OBJ_FIELD = 2 };

// Which end of a pair a shadow ray leaves from.
enum RayOrigin {
    ORIGIN_EMITTER = 0,    // point->fragment (source) and fragment->fragment across objects
    ORIGIN_RECEIVER = 1,   // fragment->point: the ray leaves the field point (OptiX.cu:163-175)
    ORIGIN_LOWER_FACET = 2 // fragment->fragment within one object: from the lower facet index
};

// Visibility constants; identical on host (oracle) and device.
struct VisConst {
    float scale;     // max(1, max |coordinate| over the scene)
    float padTri;    // 1e-5 * scale: accepted hit point must lie in the padded triangle box
    float padNode;   // 4e-5 * scale: conservative BVH node padding
    float tminPoint; // 0: rays leaving a source or field point
    float tminFrag;  // 1e-3 m: rays leaving a fragment
};

// Triangle record for the shadow test: 4 x float4 = 64 B.
//   q0 = (v0.x, v0.y, v0.z, e1.x)   q1 = (e1.y, e1.z, e2.x, e2.y)
//   q2 = (e2.z, bmin.x, bmin.y, bmin.z)  q3 = (bmax.x, bmax.y, bmax.z, triangle index bits)
// BVH node (internal): 4 x float4 = 64 B.
//   n0 = (Lmin.xyz, L child bits)  n1 = (Lmax.xyz, R child bits)
//   n2 = (Rmin.xyz, 0)             n3 = (Rmax.xyz, 0)
// child >= 0: internal node index; child < 0: leaf, sorted triangle ~child.
struct Bvh {
    const float4 *tris;   // sorted triangles (ntri x 4 float4)
    const float4 *nodes;  // ntri-1 internal nodes (x 4 float4)
    int ntri;
};

// One side of a pair set. Points carry (x,y,z,w) with w = area for fragments.
// n carries (nx,ny,nz, facet id as int bits); null when the side has no normal.
struct PointSet {
    const float4 *p;
    const float4 *n;
    int count;
};

struct MaskTask {
    PointSet em;          // emitters (columns)
    PointSet rc;          // receivers (rows, 64 per word)
    int origin;           // RayOrigin
    int useNormalEm;      // cull with the emitter-side normal
    int useNormalRc;      // cull with the receiver-side normal (0 for field destinations)
    int sameObject;       // SELF: pairs on one facet never interact
    int upperOnly;        // symmetric set: compute emitters >= 64*rb only (the rest by transposition)
    float tmin;
    uint64_t *mask;       // [ceil(rc.count/64)][em.count]
    int *errFlag;
    // occluder triangle of each point = base + its facet id (normal w bits);
    // -1: that side is not an occluder (sources, field points, field objects).
    // A ray never tests the triangles its end points lie on.
    int emTriBase, rcTriBase;
};

}  // namespace sr
