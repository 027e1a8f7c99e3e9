// Projection sequencing: the reference's render_cuda / Project* call sequence
// (src/ModelCuda.cpp:242-356, src/Model.cpp:56-93) over SoA fragment tables.
#include "model.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <atomic>
#include <mutex>

#include "launch.h"
#include "model_ops.h"
#include "transfer.h"
#include "nearfield.h"

namespace sr {

// Kernel launchers from vis.hip
int lbvh_build(const float4 *dTrisUnsorted, int n, float3 lo, float3 hi, float4 *dTrisSorted, float4 *dNodes,
               hipStream_t st);
int vis_mask_launch(const MaskTask &tk, const Bvh &bvh, const VisConst &vc, hipStream_t st,
                    bool deferLower = false);
int vis_mask_fill_lower(uint64_t *mask, int n, hipStream_t st);

namespace {

// The error text of the calling thread's current call (one per thread, like
// the status the C-ABI reports from it).
thread_local std::string gErr;

// Device allocations made by the model (sr_device_allocations reports the count).
std::atomic<long> gDevAllocs{0};

template <class T>
int dev_alloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) return 0;
    gDevAllocs++;
    SR_HIP(hipMalloc((void **)p, n * sizeof(T)));
    return 0;
}
template <class T>
void dev_free(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

enum MaskKind {
    MK_P2F = 0, MK_F2P = 1, MK_SELF = 2, MK_NORESP = 3, MK_RESP_FWD = 4, MK_RESP_REC = 5,
    MK_SRCOBJ = 6  // source object (ta = 1, ia) -> target or field object (tb, ib)
};

Object *object_at(Model &m, int type, int idx) {
    std::vector<Object> &v = (type == OBJ_FIELD) ? m.fields : (type == OBJ_SOURCE ? m.sources : m.targets);
    if (idx < 0 || idx >= (int)v.size()) return nullptr;
    return &v[idx];
}

Phys make_phys(const Model &m, float deltaReg, float damp) {
    Phys p;
    p.kr = m.kf;
    const double krev = (double)m.kf / (2.0 * M_PI);
    p.krevHi = (float)krev;
    p.krevLo = (float)(krev - (double)p.krevHi);
    p.alpha = m.atten;
    p.alphaLog2e = (float)((double)m.atten * 1.4426950408889634);
    p.deltaReg = deltaReg;
    p.damp = damp;
    return p;
}

int upload_object(Model &m, Object &o) {
    const int nf = (int)o.frames.size();
    o.grids.resize(nf);
    o.cellOff.resize(nf);
    long off = 0;
    for (int f = 0; f < nf; f++) {
        o.grids[f] = facet_grid(o.frames[f], o.delta);
        o.cellOff[f] = off;
        off += (long)o.grids[f].nX * o.grids[f].nY;
    }
    o.ncell = off;
    std::vector<float4> cp(off), cn(off), ap, an;
    o.actCellHost.clear();
    for (int f = 0; f < nf; f++) {
        const FacetFrame &F = o.frames[f];
        const FacetGrid &G = o.grids[f];
        const float4 nrm = make_float4(F.normal.x, F.normal.y, F.normal.z, 0.f);
        float fid;
        std::memcpy(&fid, &f, sizeof(int));
        for (int j = 0; j < G.nY; j++)
            for (int i = 0; i < G.nX; i++) {
                const long c = o.cellOff[f] + (long)j * G.nX + i;
                const float A = G.area[(size_t)j * G.nX + i];
                Vec3 r = cell_point(F, G, o.delta, i, j, true);
                cp[c] = make_float4(r.x, r.y, r.z, A);
                cn[c] = make_float4(nrm.x, nrm.y, nrm.z, fid);
                if (A > 1e-6f) {  // OptiX.cu:138-142,219-223,286-290
                    Vec3 e = cell_point(F, G, o.delta, i, j, false);
                    ap.push_back(make_float4(e.x, e.y, e.z, A));
                    an.push_back(make_float4(nrm.x, nrm.y, nrm.z, fid));
                    o.actCellHost.push_back((int)c);
                }
            }
    }
    o.nact = (int)ap.size();
    o.actFacetHost.resize(o.nact);
    for (int k = 0; k < o.nact; k++) std::memcpy(&o.actFacetHost[k], &an[k].w, sizeof(int));
    {
        std::vector<double> perFacet(nf, 0.0);
        for (auto &q : an) { int f; std::memcpy(&f, &q.w, sizeof(int)); perFacet[f] += 1.0; }
        o.sameFacetPairs = 0;
        for (double k : perFacet) o.sameFacetPairs += k * k;
    }
    const size_t np = (size_t)m.nplanes;
    if (dev_alloc(&o.cellP, off) || dev_alloc(&o.cellN, off) || dev_alloc(&o.actP, o.nact) ||
        dev_alloc(&o.actN, o.nact) || dev_alloc(&o.actCell, o.nact) || dev_alloc(&o.P, np * off))
        return -1;
    if (off) {
        SR_HIP(hipMemcpy(o.cellP, cp.data(), off * sizeof(float4), hipMemcpyHostToDevice));
        SR_HIP(hipMemcpy(o.cellN, cn.data(), off * sizeof(float4), hipMemcpyHostToDevice));
        SR_HIP(hipMemset(o.P, 0, np * off * sizeof(float2)));
    }
    {
        std::vector<float4> fr(2 * std::max(1, nf));
        for (int f = 0; f < nf; f++) {
            fr[2 * f] = make_float4(o.frames[f].xAxis.x, o.frames[f].xAxis.y, o.frames[f].xAxis.z, 0.f);
            fr[2 * f + 1] = make_float4(o.frames[f].yAxis.x, o.frames[f].yAxis.y, o.frames[f].yAxis.z, 0.f);
        }
        if (dev_alloc(&o.frameTab, fr.size())) return -1;
        SR_HIP(hipMemcpy(o.frameTab, fr.data(), fr.size() * sizeof(float4), hipMemcpyHostToDevice));
    }
    if (o.nact) {
        SR_HIP(hipMemcpy(o.actP, ap.data(), o.nact * sizeof(float4), hipMemcpyHostToDevice));
        SR_HIP(hipMemcpy(o.actN, an.data(), o.nact * sizeof(float4), hipMemcpyHostToDevice));
        SR_HIP(hipMemcpy(o.actCell, o.actCellHost.data(), o.nact * sizeof(int), hipMemcpyHostToDevice));
    }
    if (o.type == OBJ_TARGET && o.nact) {
        if (dev_alloc(&o.Pin, np * o.nact) || dev_alloc(&o.Pout, np * o.nact)) return -1;
        SR_HIP(hipMemset(o.Pin, 0, np * o.nact * sizeof(float2)));
        SR_HIP(hipMemset(o.Pout, 0, np * o.nact * sizeof(float2)));
    }
    return 0;
}

void free_transfer(Object &o) {
    auto &t = o.tc;
    dev_free(t.A); dev_free(t.B); dev_free(t.wgSkip); dev_free(t.colScale); dev_free(t.colInv);
    dev_free(t.state); dev_free(t.rowFs);
    dev_free(t.rowAct); dev_free(t.actRow); dev_free(t.kB); dev_free(t.cls); dev_free(t.symItems);
    dev_free(t.grpCount);
    dev_free(t.nearQ); dev_free(t.nearM); dev_free(t.nearCount);
    t = Object::Transfer{};
}

void free_object(Object &o) {
    dev_free(o.cellP); dev_free(o.cellN); dev_free(o.actP); dev_free(o.actN); dev_free(o.actCell);
    dev_free(o.P); dev_free(o.Pin); dev_free(o.Pout); dev_free(o.frameTab);
    free_transfer(o);
}

// Every cell of a source object at its amplitude (host copy; set-up time only).
int fill_pressure(Object &o, float2 amp) {
    if (!o.ncell) return 0;
    std::vector<float2> h(o.ncell, amp);
    SR_HIP(hipMemcpy(o.P, h.data(), o.ncell * sizeof(float2), hipMemcpyHostToDevice));
    return 0;
}

// The partial-sum buffers of the current call: the library's, or in an
// allocation-free call the fixed ones its entry point provides.
float2 *scr_l(Model &m) { return m.noAlloc ? m.fixL : m.scratchL; }
float2 *scr_u(Model &m) { return m.noAlloc ? m.fixU : m.scratchU; }

int ensure_scratch(Model &m, size_t elems) {
    if (m.noAlloc) {
        if (elems <= m.fixElems) return 0;
        set_error("partial sums exceed the allocation-free scratch (" + std::to_string(m.fixElems) +
                  " complex values per half)");
        return -2;
    }
    if (elems <= m.scratchElems) return 0;
    dev_free(m.scratchL);
    dev_free(m.scratchU);
    m.scratchElems = 0;
    if (dev_alloc(&m.scratchL, elems) || dev_alloc(&m.scratchU, elems)) return -1;
    m.scratchElems = elems;
    return 0;
}

Bvh bvh_of(const Model &m) { return Bvh{m.trisSorted, m.nodes, m.ntri}; }

// Brackets one launch with HIP events on the model's stream (or st) when profiling.
struct ProfScope {
    Model &m;
    int kind;
    double pairs, flops, bytes;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    int workSlot = -1;
    double flopsPerStep = 0.0;
    // A counter the launch adds its executed work to (profiling only, null otherwise);
    // the record's flops then come from it at flopsPerStep each.
    unsigned long long *work(double perStep) {
        if (!a || !m.workCnt || m.workUsed >= m.workCap) return nullptr;
        workSlot = m.workUsed++;
        flopsPerStep = perStep;
        return m.workCnt + workSlot;
    }
    ProfScope(Model &mm, int k, double p, double f, double by = 0.0, hipStream_t st = nullptr)
        : m(mm), kind(k), pairs(p), flops(f), bytes(by), s(st ? st : mm.stream) {
        if (!m.profiling || m.noAlloc) return;  // allocation-free calls record no events
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { a = b = nullptr; return; }
        (void)hipEventRecord(a, s);
    }
    ~ProfScope() {
        if (!a) return;
        (void)hipEventRecord(b, s);
        m.recs.push_back(Model::Rec{kind, a, b, pairs, flops, bytes, workSlot, flopsPerStep});
    }
};

int ensure_scratch_c(Model &m, size_t elems) {
    if (elems <= m.scratchCElems) return 0;
    dev_free(m.scratchC);
    m.scratchCElems = 0;
    if (dev_alloc(&m.scratchC, elems)) return -1;
    m.scratchCElems = elems;
    return 0;
}

// Marks the main stream's current point; side work queued after side_follow
// starts there, concurrently with what the main stream runs next.
int side_mark(Model &m) {
    SR_HIP(hipEventRecord(m.sideIn, m.stream));
    return 0;
}
int side_follow(Model &m) {
    SR_HIP(hipStreamWaitEvent(m.side, m.sideIn, 0));
    return 0;
}
// The main stream continues after everything queued on the side stream so far.
int side_join(Model &m) {
    SR_HIP(hipEventRecord(m.sideOut, m.side));
    SR_HIP(hipStreamWaitEvent(m.stream, m.sideOut, 0));
    return 0;
}

// A device count copied into pinned host word `slot`; the host waits for that
// copy alone (word_wait), not for the work queued after it.
int word_readback(Model &m, const int *dev, int slot) {
    SR_HIP(hipMemcpyAsync(m.hostWords + slot, dev, sizeof(int), hipMemcpyDeviceToHost, m.stream));
    SR_HIP(hipEventRecord(m.wordEv[slot], m.stream));
    return 0;
}
int word_wait(Model &m, int slot, int &out) {
    SR_HIP(hipEventSynchronize(m.wordEv[slot]));
    out = m.hostWords[slot];
    return 0;
}

}  // namespace

int settle(Model &m);

void set_error(const std::string &msg) { gErr = msg; }
std::string last_error() { return gErr; }
void clear_error() { set_error(""); }

// ---------------------------------------------------------------------------

int model_set_initial_conditions(Model &m, float cp, float f, float att, float density) {
    if (!(cp > 0.f)) { set_error("set_initial_conditions: sound speed must be > 0"); return -2; }
    m.cp = cp; m.freq = f; m.atten = att; m.density = density;
    const float omega = (float)(2 * M_PI * f);  // Model.cpp:50 (float member)
    m.kf = omega / cp;                           // Model.cpp:51
    for (auto &o : m.targets) o.tc.valid = false;  // S depends on k
    return 0;
}

int model_load_geometry(Model &m, const float *v, int nf, int type, float res) {
    if (type < OBJ_TARGET || type > OBJ_FIELD) { set_error("load_geometry: invalid object type"); return -2; }
    if (m.rendered) { set_error("load_geometry: scene already rendered; call TearDownCuda first"); return -3; }
    if (nf < 0 || (nf > 0 && !v)) { set_error("load_geometry: bad facet array"); return -2; }
    if (!(res > 0.f)) { set_error("load_geometry: resolution must be > 0"); return -2; }
    Object o;
    o.type = type;
    o.delta = res;
    o.frames.reserve(nf);
    for (int i = 0; i < nf; i++) {
        const float *t = v + (size_t)9 * i;
        o.frames.push_back(facet_frame(Vec3{t[0], t[1], t[2]}, Vec3{t[3], t[4], t[5]}, Vec3{t[6], t[7], t[8]}));
    }
    if (type == OBJ_SOURCE) {
        m.sources.push_back(std::move(o));
        m.srcObjAmp.push_back(make_float2(1.f, 0.f));
        return 0;
    }
    (type == OBJ_TARGET ? m.targets : m.fields).push_back(std::move(o));
    return 0;
}

int model_load_points(Model &m, const float *p, int n, bool source) {
    if (m.rendered) { set_error("points loaded after render_cuda (already allocated on the GPU)"); return -3; }
    if (n < 0 || (n > 0 && !p)) { set_error("load points: bad array"); return -2; }
    std::vector<float> &dst = source ? m.srcHost : m.fldHost;
    dst.insert(dst.end(), p, p + (size_t)3 * n);
    return 0;
}

int model_set_batched_sources(Model &m, int on) {
    if (m.rendered) { set_error("batched sources must be chosen before render_cuda"); return -3; }
    m.batchedSources = on != 0;
    return 0;
}

int model_render(Model &m) {
    if (m.rendered) { set_error("render_cuda called twice without TearDownCuda"); return -3; }
    if (!m.stream) SR_HIP(hipStreamCreateWithFlags(&m.stream, hipStreamNonBlocking));
    if (!m.side) {
        SR_HIP(hipStreamCreateWithFlags(&m.side, hipStreamNonBlocking));
        SR_HIP(hipEventCreateWithFlags(&m.sideIn, hipEventDisableTiming));
        SR_HIP(hipEventCreateWithFlags(&m.sideOut, hipEventDisableTiming));
    }
    if (!m.hostWords) {  // kept, like the stream, for the life of the process
        SR_HIP(hipHostMalloc((void **)&m.hostWords, Model::kWords * sizeof(int), hipHostMallocDefault));
        for (auto &e : m.wordEv) SR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        SR_HIP(hipHostMalloc((void **)&m.errHost, sizeof(int), hipHostMallocCoherent | hipHostMallocMapped));
        SR_HIP(hipHostGetDevicePointer((void **)&m.errFlag, m.errHost, 0));
    }
    *(volatile int *)m.errHost = 0;
    m.errPending = false;
    const int ns = (int)(m.srcHost.size() / 3), nq = (int)(m.fldHost.size() / 3);
    if (m.batchedSources && !m.sources.empty()) {
        set_error("source objects radiate into every incidence: use them with unbatched source points");
        return -3;
    }
    m.nplanes = m.batchedSources ? std::max(1, ns) : 1;
    for (auto &o : m.targets) if (upload_object(m, o)) return -1;
    for (auto &o : m.fields) if (upload_object(m, o)) return -1;
    for (size_t i = 0; i < m.sources.size(); i++) {
        if (upload_object(m, m.sources[i])) return -1;
        if (fill_pressure(m.sources[i], m.srcObjAmp[i])) return -1;
    }

    // Occluders: every target triangle in load order (OptiX.cpp:273-292).
    std::vector<float4> tris;
    float scale = 1.0f;
    float3 lo = make_float3(1e30f, 1e30f, 1e30f), hi = make_float3(-1e30f, -1e30f, -1e30f);
    auto upd = [&](float x, float y, float z) {
        scale = std::max(scale, std::max(std::fabs(x), std::max(std::fabs(y), std::fabs(z))));
    };
    for (auto &o : m.targets) {
        o.triBase = (int)(tris.size() / 4);
        for (auto &F : o.frames) {
            Vec3 a = F.v1, b = F.v2, c = F.v3;
            Vec3 e1{b.x - a.x, b.y - a.y, b.z - a.z}, e2{c.x - a.x, c.y - a.y, c.z - a.z};
            Vec3 mn{std::min(a.x, std::min(b.x, c.x)), std::min(a.y, std::min(b.y, c.y)), std::min(a.z, std::min(b.z, c.z))};
            Vec3 mx{std::max(a.x, std::max(b.x, c.x)), std::max(a.y, std::max(b.y, c.y)), std::max(a.z, std::max(b.z, c.z))};
            tris.push_back(make_float4(a.x, a.y, a.z, e1.x));
            tris.push_back(make_float4(e1.y, e1.z, e2.x, e2.y));
            tris.push_back(make_float4(e2.z, mn.x, mn.y, mn.z));
            // w: the triangle's index in load order (the facet a ray end point may lie on)
            const int id = (int)(tris.size() / 4);
            float idBits;
            std::memcpy(&idBits, &id, sizeof(float));
            tris.push_back(make_float4(mx.x, mx.y, mx.z, idBits));
            lo = make_float3(std::min(lo.x, mn.x), std::min(lo.y, mn.y), std::min(lo.z, mn.z));
            hi = make_float3(std::max(hi.x, mx.x), std::max(hi.y, mx.y), std::max(hi.z, mx.z));
            upd(a.x, a.y, a.z); upd(b.x, b.y, b.z); upd(c.x, c.y, c.z);
        }
    }
    for (auto *v : {&m.fields, &m.sources})  // never occluders: only targets are (OptiX.cpp:273-292)
        for (auto &o : *v)
            for (auto &F : o.frames) { upd(F.v1.x, F.v1.y, F.v1.z); upd(F.v2.x, F.v2.y, F.v2.z); upd(F.v3.x, F.v3.y, F.v3.z); }
    for (size_t i = 0; i < m.srcHost.size(); i += 3) upd(m.srcHost[i], m.srcHost[i + 1], m.srcHost[i + 2]);
    for (size_t i = 0; i < m.fldHost.size(); i += 3) upd(m.fldHost[i], m.fldHost[i + 1], m.fldHost[i + 2]);
    m.vc.scale = scale;
    m.vc.padTri = 1e-5f * scale;
    m.vc.padNode = 4e-5f * scale;
    m.vc.tminPoint = 0.0f;
    m.vc.tminFrag = 1e-3f;

    m.ntri = (int)(tris.size() / 4);
    if (m.ntri > 0) {
        float4 *unsorted = nullptr;
        if (dev_alloc(&unsorted, tris.size()) || dev_alloc(&m.trisSorted, tris.size()) ||
            dev_alloc(&m.nodes, (size_t)4 * std::max(1, m.ntri - 1)))
            return -1;
        SR_HIP(hipMemcpy(unsorted, tris.data(), tris.size() * sizeof(float4), hipMemcpyHostToDevice));
        int rc = lbvh_build(unsorted, m.ntri, lo, hi, m.trisSorted, m.nodes, m.stream);
        SR_HIP(hipStreamSynchronize(m.stream));
        dev_free(unsorted);
        if (rc) return rc;
    }

    // Points (ModelCuda.cpp:39-94); sources carry amplitude 1+0i (Model.cpp:18).
    if (ns) {
        std::vector<float4> sp(ns);
        std::vector<float2> sa(ns, make_float2(1.f, 0.f));
        for (int i = 0; i < ns; i++) sp[i] = make_float4(m.srcHost[3 * i], m.srcHost[3 * i + 1], m.srcHost[3 * i + 2], 0.f);
        if (dev_alloc(&m.srcP, ns) || dev_alloc(&m.srcAmp, ns)) return -1;
        SR_HIP(hipMemcpy(m.srcP, sp.data(), ns * sizeof(float4), hipMemcpyHostToDevice));
        SR_HIP(hipMemcpy(m.srcAmp, sa.data(), ns * sizeof(float2), hipMemcpyHostToDevice));
    }
    if (nq) {
        std::vector<float4> fp(nq);
        for (int i = 0; i < nq; i++) fp[i] = make_float4(m.fldHost[3 * i], m.fldHost[3 * i + 1], m.fldHost[3 * i + 2], 0.f);
        if (dev_alloc(&m.fldP, nq) || dev_alloc(&m.fldVal, (size_t)nq * m.nplanes)) return -1;
        SR_HIP(hipMemcpy(m.fldP, fp.data(), nq * sizeof(float4), hipMemcpyHostToDevice));
        SR_HIP(hipMemset(m.fldVal, 0, (size_t)nq * m.nplanes * sizeof(double2)));
    }
    // The field-point visibility words of every target are allocated here, so that
    // field-point projections (the allocation-free one included) allocate nothing.
    if (nq)
        for (int i = 0; i < (int)m.targets.size(); i++) {
            if (!m.targets[i].nact) continue;
            MaskTask tk{};
            size_t words = 0;
            if (mask_task(m, MK_F2P, OBJ_TARGET, i, 0, 0, tk, words)) return -1;
            Model::MaskBuf &mb = m.masks[std::make_tuple((int)MK_F2P, (int)OBJ_TARGET, i, 0, 0)];
            if (dev_alloc(&mb.p, words)) return -1;
            mb.words = words;
            mb.valid = false;
        }
    SR_HIP(hipDeviceSynchronize());
    m.rendered = true;
    return 0;
}

int model_set_source_amplitude(Model &m, int i, float re, float im) {
    const int ns = (int)(m.srcHost.size() / 3);
    if (!m.rendered) { set_error("source amplitudes live on the GPU: call render_cuda first"); return -3; }
    if (i < 0 || i >= ns) { set_error("source index out of range"); return -2; }
    if (int rc = settle(m)) return rc;  // queued projections still read the amplitudes
    float2 a = make_float2(re, im);
    SR_HIP(hipMemcpy(m.srcAmp + i, &a, sizeof(float2), hipMemcpyHostToDevice));
    return 0;
}

// Describes the visibility words of one (emitter set, receiver set); no launch.
int mask_task(Model &m, int kind, int ta, int ia, int tb, int ib, MaskTask &tk, size_t &words) {
    tk = MaskTask{};
    tk.emTriBase = tk.rcTriBase = -1;
    Object *A = (kind == MK_P2F || kind == MK_SRCOBJ) ? nullptr : object_at(m, ta, ia);
    Object *B = (kind == MK_P2F || kind >= MK_NORESP) ? object_at(m, tb, ib) : nullptr;
    const int ns = (int)(m.srcHost.size() / 3), nq = (int)(m.fldHost.size() / 3);
    switch (kind) {
    case MK_P2F:
        tk.em = PointSet{m.srcP, nullptr, ns};
        tk.rc = PointSet{B->cellP, B->cellN, (int)B->ncell};
        tk.rcTriBase = B->type == OBJ_TARGET ? B->triBase : -1;
        tk.origin = ORIGIN_EMITTER; tk.useNormalEm = 0; tk.useNormalRc = (B->type == OBJ_TARGET);
        tk.tmin = m.vc.tminPoint;
        break;
    case MK_F2P:
        tk.em = PointSet{A->actP, A->actN, A->nact};
        tk.rc = PointSet{m.fldP, nullptr, nq};
        tk.emTriBase = A->triBase;
        tk.origin = ORIGIN_RECEIVER; tk.useNormalEm = 1; tk.useNormalRc = 0;
        tk.tmin = m.vc.tminPoint;
        break;
    case MK_SELF:
        tk.em = PointSet{A->actP, A->actN, A->nact};
        tk.rc = tk.em;
        tk.emTriBase = tk.rcTriBase = A->triBase;
        tk.origin = ORIGIN_LOWER_FACET; tk.useNormalEm = 1; tk.useNormalRc = 1; tk.sameObject = 1;
        tk.upperOnly = 1;  // the set is symmetric: half the rays, the rest by bit transposition
        tk.tmin = m.vc.tminFrag;
        break;
    case MK_NORESP:
        tk.em = PointSet{A->actP, A->actN, A->nact};
        tk.rc = PointSet{B->actP, B->actN, B->nact};
        tk.emTriBase = A->triBase;
        tk.rcTriBase = B->type == OBJ_TARGET ? B->triBase : -1;
        tk.origin = ORIGIN_EMITTER; tk.useNormalEm = 1; tk.useNormalRc = (B->type == OBJ_TARGET);
        tk.tmin = m.vc.tminFrag;
        break;
    case MK_RESP_FWD:
        tk.em = PointSet{A->actP, A->actN, A->nact};
        tk.rc = PointSet{B->actP, B->actN, B->nact};
        tk.emTriBase = A->triBase;
        tk.rcTriBase = B->triBase;
        tk.origin = ORIGIN_EMITTER; tk.useNormalEm = 1; tk.useNormalRc = 1;
        tk.tmin = m.vc.tminFrag;
        break;
    case MK_SRCOBJ:
        A = object_at(m, OBJ_SOURCE, ia);
        if (!A || !B) { set_error("source-object mask: no such object"); return -2; }
        tk.em = PointSet{A->actP, A->actN, A->nact};
        tk.rc = PointSet{B->actP, B->actN, B->nact};
        tk.emTriBase = -1;  // source objects never occlude
        tk.rcTriBase = B->type == OBJ_TARGET ? B->triBase : -1;
        tk.origin = ORIGIN_EMITTER; tk.useNormalEm = 1; tk.useNormalRc = (B->type == OBJ_TARGET);
        tk.tmin = m.vc.tminFrag;
        break;
    case MK_RESP_REC:
        tk.em = PointSet{B->actP, B->actN, B->nact};
        tk.rc = PointSet{A->actP, A->actN, A->nact};
        tk.emTriBase = B->triBase;
        tk.rcTriBase = A->triBase;
        tk.origin = ORIGIN_RECEIVER; tk.useNormalEm = 1; tk.useNormalRc = 1;
        tk.tmin = m.vc.tminFrag;
        break;
    default:
        set_error("unknown mask kind");
        return -2;
    }
    const size_t nrb = (size_t)(tk.rc.count + 63) / 64;
    words = std::max<size_t>(1, nrb * (size_t)tk.em.count);
    tk.errFlag = m.errFlag;
    return 0;
}

// Visibility words for one (emitter set, receiver set), built on first use.
// upperOk: the caller reads only the words at or above the diagonal blocks of a
// symmetric set (emitter >= 64 x receiver block); the others are filled in by
// transposition when a caller first needs them.
int get_mask(Model &m, int kind, int ta, int ia, int tb, int ib, uint64_t **out, bool upperOk = false) {
    auto key = std::make_tuple(kind, ta, ia, tb, ib);
    auto it = m.masks.find(key);
    if (it != m.masks.end() && it->second.valid) {
        Model::MaskBuf &mb = it->second;
        if (mb.lowerPending && !upperOk) {
            MaskTask tk{};
            size_t words = 0;
            if (mask_task(m, kind, ta, ia, tb, ib, tk, words)) return -1;
            ProfScope ps(m, PK_VIS, 0.0, 0.0, (double)words * 8.0);
            if (vis_mask_fill_lower(mb.p, tk.rc.count, m.stream)) return -1;
            mb.lowerPending = false;
        }
        *out = mb.p;
        return 0;
    }
    MaskTask tk{};
    size_t words = 0;
    if (mask_task(m, kind, ta, ia, tb, ib, tk, words)) return -1;
    if (m.noAlloc) {  // the buffer must exist already (render_cuda allocates the field-point words)
        auto f = m.masks.find(key);
        if (f == m.masks.end() || f->second.words != words) {
            set_error("allocation-free call: its visibility words are not allocated");
            return -3;
        }
    }
    Model::MaskBuf &mb = m.masks[key];
    if (mb.words != words) {
        dev_free(mb.p);
        mb.words = 0;
        if (dev_alloc(&mb.p, words)) return -1;
        mb.words = words;
    }
    uint64_t *mk = mb.p;
    tk.mask = mk;
    {
        // pairs: the (emitter, receiver) pairs the words cover (a symmetric set traces half
        // of them); bytes: the words written, the kernel's only output
        ProfScope ps(m, PK_VIS, (double)tk.em.count * tk.rc.count, 0.0, (double)words * 8.0);
        if (vis_mask_launch(tk, bvh_of(m), m.vc, m.stream, upperOk)) return -1;
    }
    mb.lowerPending = upperOk && tk.upperOnly;
    mb.valid = true;  // the traversal-overflow flag is checked when the projection finishes
    *out = mk;
    return 0;
}

// End of a projection. Projections return once their work is queued on the
// library's stream (the host queues the next call's work while this one runs);
// the visibility kernels store a traversal overflow straight into a pinned word,
// which the next call that reads results (settle) checks.
int finish_projection(Model &m) {
    m.errPending = true;
    return 0;
}

// Waits for all queued work; reports a traversal overflow raised by any
// projection since the last settle (its visibility words are then dropped and
// the results since are not valid).
int settle(Model &m) {
    if (m.stream) SR_HIP(hipStreamSynchronize(m.stream));
    if (m.side) SR_HIP(hipStreamSynchronize(m.side));
    if (!m.errPending) return 0;
    m.errPending = false;
    volatile int *err = m.errHost;  // no work is queued: the host may clear it
    if (*err) {
        *err = 0;
        for (auto &kv : m.masks) kv.second.valid = false;
        // transfer matrices built from those words are rebuilt with them
        for (auto &o : m.targets) o.tc.valid = false;
        set_error("visibility traversal stack overflow in a projection since the last result read");
        return -1;
    }
    return 0;
}

static int check_rendered(const Model &m) {
    if (!m.rendered) { set_error("call render_cuda before projecting"); return -3; }
    return 0;
}

static int gather_planes(Model &m, GatherTask t, const Phys &ph, int kind, ReduceTask rt, int nplanes,
                         long evalStride, long gvalStride, long outStride, long pcellStride, double pairs);

// ProjectSrcPointsToObjects (ModelCuda.cpp:242-278)
int model_project_src(Model &m, bool fieldObjects) {
    if (int rc = check_rendered(m)) return rc;
    const int ns = (int)(m.srcHost.size() / 3);
    Phys ph = make_phys(m, 0.f, 1.f);
    auto run = [&](Object &o, int type, int idx) -> int {
        if (o.ncell == 0 || ns == 0) return 0;
        // few occluders and one plane per source: the sum evaluates its visibility itself
        const bool inl = m.inlineVis && m.batchedSources && m.ntri >= 1 && m.ntri <= 4;
        uint64_t *mk = nullptr;
        if (!inl && get_mask(m, MK_P2F, 0, 0, type, idx, &mk)) return -1;
        P2FTask t{o.cellP, o.cellN, (int)o.ncell, o.type == OBJ_TARGET, m.srcP, m.srcAmp, ns, mk,
                  m.batchedSources ? 1 : 0, o.P, o.ncell};
        if (inl)
            t.vis = P2FVis{m.trisSorted, m.ntri, o.type == OBJ_TARGET ? o.triBase : -1, m.vc.scale, m.vc.padTri,
                           m.vc.tminPoint};
        ProfScope ps(m, PK_P2F, (double)ns * o.ncell, 30.0 * ns * o.ncell);
        return p2f_launch(t, ph, m.stream);
    };
    for (int i = 0; i < (int)m.targets.size(); i++) if (run(m.targets[i], OBJ_TARGET, i)) return -1;
    if (fieldObjects)
        for (int i = 0; i < (int)m.fields.size(); i++) if (run(m.fields[i], OBJ_FIELD, i)) return -1;
    // Source objects: each cell at the object's coherent amplitude, summed into the
    // receivers' total pressure P by the fragment->fragment kernel without
    // reciprocity (FACET_NO_RESP form, OptiX.cu:192-406; receivers at their fragment points).
    for (int si = 0; si < (int)m.sources.size(); si++) {
        Object &S = m.sources[si];
        if (!S.nact) continue;
        auto emit = [&](Object &B, int type, int idx) -> int {
            if (!B.nact) return 0;
            uint64_t *mk;
            if (get_mask(m, MK_SRCOBJ, OBJ_SOURCE, si, type, idx, &mk)) return -1;
            Phys pf = make_phys(m, S.delta, 0.9f);
            GatherTask t{};
            t.ep = S.actP; t.en = S.actN; t.ecell = S.actCell; t.eval = S.P; t.ne = S.nact;
            t.rp = B.actP; t.rn = B.actN; t.nr = B.nact; t.mask = mk;
            t.gateMode = GATE_NONE;
            ReduceTask r{};
            r.gateMode = GATE_NONE;
            r.pcell = B.P; r.rowCell = B.actCell;
            const int kind = type == OBJ_FIELD ? KIND_F2F_FIELD : KIND_F2F;
            return gather_planes(m, t, pf, kind, r, m.nplanes, S.ncell, 0, 0, B.ncell, (double)S.nact * B.nact);
        };
        for (int i = 0; i < (int)m.targets.size(); i++) if (emit(m.targets[i], OBJ_TARGET, i)) return -1;
        if (fieldObjects)
            for (int i = 0; i < (int)m.fields.size(); i++) if (emit(m.fields[i], OBJ_FIELD, i)) return -1;
    }
    return finish_projection(m);
}

int model_set_source_object_amplitude(Model &m, int i, float re, float im) {
    if (i < 0 || i >= (int)m.sources.size()) { set_error("source object index out of range"); return -2; }
    m.srcObjAmp[i] = make_float2(re, im);
    if (!m.rendered) return 0;
    if (int rc = settle(m)) return rc;  // queued projections still read the cells
    return fill_pressure(m.sources[i], m.srcObjAmp[i]);
}

// One gather + reduce over all planes in groups of <= 16.
// pairs: ordered pairs one plane of this gather evaluates (for the profiler's
// flop count: 37 shared flops per pair plus 8 per plane for the complex MAC).
// Fragment->fragment with the complex products on the FP32 matrix cores, in
// groups of up to 64 planes; a GATE_UPPER task then has its gated rows
// recomputed exactly (lower-facet emitters only).
static int gather_planes_mfma(Model &m, GatherTask t, const Phys &ph, ReduceTask rt, int nplanes, long evalStride,
                              long gvalStride, long outStride, long pcellStride, double pairs) {
    const bool gated = t.gateMode == GATE_UPPER;
    int ngated = 0;
    if (gated) {
        if (t.ep != t.rp || t.en != t.rn || !rt.out || !rt.pcell) {
            set_error("gated MFMA gather needs one object on both sides");
            return -2;
        }
        if (m.gatedCap < t.nr) {
            dev_free(m.gatedList);
            m.gatedCap = 0;
            if (dev_alloc(&m.gatedList, t.nr)) return -1;
            m.gatedCap = t.nr;
        }
        if (!m.gatedCount && dev_alloc(&m.gatedCount, 1)) return -1;
        if (counters_zero_launch(m.gatedCount, 1, m.stream)) return -1;
        if (transfer_gated_rows(t.gval, t.nr, nplanes, m.gatedList, m.gatedCount, m.stream)) return -1;
        if (word_readback(m, m.gatedCount, 0)) return -1;  // read after the main passes are queued
    }
    // one plane group: the gated rows' compact set runs on the side stream, concurrently
    // with the main passes (it reads only the inputs, positions and visibility words)
    const bool overlap = gated && nplanes <= 64;
    if (overlap && side_mark(m)) return -1;
    // ~2048 workgroups of 128 receivers; emitter ranges in whole LDS tiles of 32
    const int groups = (t.nr + 127) / 128;
    int nsplit = std::max(1, std::min((2048 + groups - 1) / groups, t.ne / 64));
    t.emPerSplit = ((t.ne + nsplit - 1) / nsplit + 31) / 32 * 32;
    t.nsplit = (t.ne + t.emPerSplit - 1) / t.emPerSplit;
    // also covers the compact set's partials: splits <= ceil(512 / groups), so split count x rows <= 512 x 128 + nr + 128
    if (ensure_scratch(m, std::max((size_t)t.nsplit * t.nr, (size_t)512 * 128 + t.nr + 128) * 64)) return -1;
    for (int b0 = 0; b0 < nplanes; b0 += 64) {
        const int nb = std::min(64, nplanes - b0);
        GatherTask g = t;
        g.nrhs = nb;
        g.evalStride = evalStride;
        g.eval = t.eval + b0 * evalStride;
        g.gval = nullptr;
        g.gateMode = GATE_NONE;
        g.partL = scr_l(m);
        g.partU = nullptr;
        {
            // flops: the complex multiply-accumulates executed on the matrix cores, 8 per
            // (pair, plane), counted by the kernel (emitter pairs no receiver of a wave sees are
            // skipped); without a counter, every pair
            ProfScope ps(m, PK_F2F, pairs * nb, pairs * 8.0 * nb);
            g.workCount = ps.work(2.0 * 32.0 * 8.0 * nb);
            if (f2f_mfma_launch(g, ph, KIND_F2F, m.stream)) return -1;
        }
        ReduceTask r = rt;
        r.partL = scr_l(m);
        r.partU = nullptr;
        r.nsplit = g.nsplit;
        r.nbStride = 64;
        r.nrhs = nb;
        r.nr = t.nr;
        r.gvalStride = gvalStride;
        r.gateMode = gated ? GATE_NONE : rt.gateMode;
        r.outStride = outStride;
        r.pcellStride = pcellStride;
        if (rt.gval) r.gval = rt.gval + b0 * gvalStride;
        if (rt.out) r.out = rt.out + b0 * outStride;
        if (rt.pcell) r.pcell = rt.pcell + b0 * pcellStride;
        ProfScope pr(m, PK_REDUCE, 0.0, 0.0);
        if (reduce_launch(r, m.stream)) return -1;
    }
    if (!gated) return 0;
    if (word_wait(m, 0, ngated)) return -1;
    if (!ngated) return 0;
    if (outStride != gvalStride) { set_error("gated MFMA gather: P_in and P_out strides differ"); return -2; }
    // the gated rows as a compact receiver set whose words hold only lower-facet emitters
    const size_t words = (size_t)((ngated + 63) / 64) * t.ne;
    if (m.gCap < ngated) {
        dev_free(m.gPc); dev_free(m.gNc); dev_free(m.gFs);
        m.gCap = 0;
        if (dev_alloc(&m.gPc, ngated) || dev_alloc(&m.gNc, ngated) || dev_alloc(&m.gFs, ngated)) return -1;
        m.gCap = ngated;
    }
    if (m.gMcCap < words) {
        dev_free(m.gMc);
        m.gMcCap = 0;
        if (dev_alloc(&m.gMc, words)) return -1;
        m.gMcCap = words;
    }
    const hipStream_t cst = overlap ? m.side : m.stream;
    if (overlap && (ensure_scratch_c(m, ((size_t)512 * 128 + t.nr + 128) * 64) || side_follow(m))) return -1;
    {
        ProfScope ps(m, PK_GATED_FIX, 0.0, 0.0, 0.0, cst);
        if (gated_compact_launch(m.gatedList, ngated, t.rp, t.rn, t.mask, t.ne, m.gPc, m.gNc, m.gFs, m.gMc, cst))
            return -1;
    }
    GatherTask c = t;
    c.evalStride = evalStride;
    c.rp = m.gPc; c.rn = m.gNc; c.nr = ngated; c.mask = m.gMc;
    c.gval = nullptr; c.gateMode = GATE_NONE; c.partL = scr_l(m); c.partU = nullptr;
    // ~512 workgroups: enough to fill the chip, few partial sums per entry
    const int cgroups = (ngated + 127) / 128;
    int cs = std::max(1, std::min((512 + cgroups - 1) / cgroups, t.ne / 64));
    c.emPerSplit = ((t.ne + cs - 1) / cs + 31) / 32 * 32;
    c.nsplit = (t.ne + c.emPerSplit - 1) / c.emPerSplit;
    if ((size_t)c.nsplit * 64 * ngated > (overlap ? m.scratchCElems : m.scratchElems)) {
        set_error("gated MFMA gather: scratch bound");
        return -2;
    }
    c.partL = overlap ? m.scratchC : scr_l(m);
    for (int b0 = 0; b0 < nplanes; b0 += 64) {
        const int nb = std::min(64, nplanes - b0);
        c.nrhs = nb;
        c.eval = t.eval + b0 * evalStride;
        {
            ProfScope ps(m, PK_GATED_FIX, 0.0, 0.0, 0.0, cst);
            if (f2f_mfma_launch(c, ph, KIND_F2F, cst, true)) return -1;
        }
        if (overlap && side_join(m)) return -1;
        ProfScope ps(m, PK_GATED_FIX, 0.0, 0.0);
        if (gated_replace_launch(c.partL, c.nsplit, ngated, m.gatedList, t.gval + b0 * gvalStride, gvalStride, nb,
                                 rt.out + b0 * outStride, rt.pcell + b0 * pcellStride, pcellStride, rt.rowCell,
                                 nullptr, m.stream))
            return -1;
    }
    return 0;
}

// Fragment->field point on the FP32 matrix cores, groups of up to 64 planes: the
// emitters' total pressures are read through the emitter->cell map as the tiles
// are loaded; partial sums are FP32 within a split of 32 emitters and FP64 across splits.
static int gather_points_mfma(Model &m, GatherTask t, const Phys &ph, int nplanes, long evalStride, double pairs) {
    int nsplit = std::max(1, std::min(2048, (t.ne + 31) / 32));
    if (m.noAlloc)  // as many splits as the fixed scratch holds
        nsplit = (int)std::max<long>(1, std::min<long>(nsplit, (long)(m.fixElems / ((size_t)64 * std::max(1, t.nr)))));
    t.emPerSplit = ((t.ne + nsplit - 1) / nsplit + 31) / 32 * 32;
    t.nsplit = (t.ne + t.emPerSplit - 1) / t.emPerSplit;
    if (ensure_scratch(m, (size_t)t.nsplit * 64 * t.nr)) return -1;  // the partials
    for (int b0 = 0; b0 < nplanes; b0 += 64) {
        const int nb = std::min(64, nplanes - b0);
        GatherTask g = t;
        g.nrhs = nb;
        g.eval = t.eval + b0 * evalStride;
        g.evalStride = evalStride;
        g.partL = scr_l(m);
        g.partU = nullptr;
        {
            ProfScope ps(m, PK_F2P, pairs * nb, pairs * 8.0 * nb);
            if (f2f_mfma_launch(g, ph, KIND_F2P, m.stream)) return -1;
        }
        ProfScope pr(m, PK_REDUCE, 0.0, 0.0);
        if (reduce_points_launch(scr_l(m), nullptr, g.nsplit, 64, nb, t.nr, m.fldVal + (size_t)b0 * t.nr, t.nr,
                                 m.stream))
            return -1;
    }
    return 0;
}

static int gather_planes(Model &m, GatherTask t, const Phys &ph, int kind, ReduceTask rt, int nplanes,
                         long evalStride, long gvalStride, long outStride, long pcellStride, double pairs) {
    if (kind == KIND_F2P && m.gatherMode == 0 && nplanes >= 16)
        return gather_points_mfma(m, t, ph, nplanes, evalStride, pairs);
    if (kind == KIND_F2F && m.gatherMode == 0 && nplanes >= 16 && !t.ecell &&
        (t.gateMode == GATE_NONE || (t.gateMode == GATE_UPPER && t.gval == t.eval && gvalStride == evalStride)))
        return gather_planes_mfma(m, t, ph, rt, nplanes, evalStride, gvalStride, outStride, pcellStride, pairs);
    const int nrb = (t.nr + 63) / 64;
    // few receivers (field points): spread the emitters thinly across many workgroups
    const int minPer = (t.nr <= 256) ? 16 : 256;
    t.nsplit = choose_splits(nrb, t.ne, minPer);
    if (m.noAlloc)  // as many splits as the fixed scratch holds (16-plane groups)
        t.nsplit = (int)std::max<long>(1, std::min<long>(t.nsplit, (long)(m.fixElems /
                                        ((size_t)gather_nb_stride(std::min(16, nplanes)) * std::max(1, t.nr)))));
    t.emPerSplit = (t.ne + t.nsplit - 1) / t.nsplit;
    t.nsplit = (t.ne + t.emPerSplit - 1) / t.emPerSplit;
    for (int b0 = 0; b0 < nplanes; b0 += 16) {
        const int nb = std::min(16, nplanes - b0);
        const int nbs = gather_nb_stride(nb);
        if (ensure_scratch(m, (size_t)t.nsplit * nbs * t.nr)) return -1;
        GatherTask g = t;
        g.nrhs = nb;
        g.evalStride = evalStride;
        g.gvalStride = gvalStride;
        g.eval = t.eval + b0 * evalStride;
        if (t.gval) g.gval = t.gval + b0 * gvalStride;
        g.partL = scr_l(m);
        g.partU = scr_u(m);
        {
            const int pk = kind == KIND_F2F ? PK_F2F : (kind == KIND_F2F_FIELD ? PK_F2F_FIELD : PK_F2P);
            ProfScope ps(m, pk, pairs * nb, pairs * (37.0 + 8.0 * nb));
            if (gather_launch(g, ph, kind, m.stream)) return -1;
        }
        ProfScope pr(m, PK_REDUCE, 0.0, 0.0);
        if (kind == KIND_F2P) {
            if (reduce_points_launch(scr_l(m), scr_u(m), g.nsplit, nbs, nb, t.nr, m.fldVal + (size_t)b0 * t.nr,
                                     t.nr, m.stream))
                return -1;
        } else {
            ReduceTask r = rt;
            r.partL = scr_l(m);
            r.partU = scr_u(m);
            r.nsplit = g.nsplit;
            r.nbStride = nbs;
            r.nrhs = nb;
            r.nr = t.nr;
            r.gvalStride = gvalStride;
            r.outStride = outStride;
            r.pcellStride = pcellStride;
            if (rt.gval) r.gval = rt.gval + b0 * gvalStride;
            if (rt.out) r.out = rt.out + b0 * outStride;
            if (rt.pcell) r.pcell = rt.pcell + b0 * pcellStride;
            if (reduce_launch(r, m.stream)) return -1;
        }
    }
    return 0;
}

// Builds (or rebuilds) the FP16 transfer matrix of target i.
static int ensure_transfer(Model &m, int i) {
    Object &o = m.targets[i];
    auto &t = o.tc;
    if (t.valid && (m.nearMode != 2 || t.nfGen == nearfield().gen)) return 0;
    const int n = o.nact;
    // Facet ranges of the active rows (sorted by facet). When every facet holds at
    // least 128 rows and padding each to whole 128-row groups adds at most a
    // quarter, the rows of S are padded so: every group is then single-facet and
    // the zero-input gate is applied inside the product. The emitters (K) are then
    // padded the same way (K position k is row k), which makes every facet's K range
    // whole chunks and the transpose of an aligned 64 x 64 block of S another one.
    const auto &fa = o.actFacetHost;
    std::vector<int2> fr;  // [fs, fe) per facet present
    for (int r = 0; r < n; r++)
        if (r == 0 || fa[r] != fa[r - 1]) fr.push_back(make_int2(r, r + 1)); else fr.back().y = r + 1;
    long padded = 0;
    bool big = !fr.empty();
    for (auto &f : fr) {
        big = big && (f.y - f.x) >= 128;
        padded += ((f.y - f.x) + 127) / 128 * 128;
    }
    const bool gateIn = big && padded <= n + n / 4;
    const int nPad = gateIn ? (int)padded : n;
    const int nRt = ((nPad + 127) / 128) * 4;
    // K steps of 8 emitters, whole chunks of 8 steps
    const int nKs = gateIn ? nPad / 8 : ((n + 7) / 8 + 7) / 8 * 8;
    if (!t.A || t.nKs != nKs || t.nRt != nRt || t.gateInGemm != gateIn || t.splitsReq != m.transferSplits) {
        free_transfer(o);
        const int nWg = nRt / 4;
        if (dev_alloc(&t.A, (size_t)nRt * nKs * 64) || dev_alloc(&t.B, (size_t)nKs * 128) ||
            dev_alloc(&t.wgSkip, nWg) || dev_alloc(&t.grpCount, nWg) || dev_alloc(&t.colScale, 64) ||
            dev_alloc(&t.colInv, 64) ||
            dev_alloc(&t.state, 2 * (1 + (size_t)m.nplanes * kColStride)) ||
            dev_alloc(&t.rowFs, n))
            return -1;
        SR_HIP(hipMemsetAsync(t.grpCount, 0, (size_t)nWg * sizeof(unsigned), m.stream));
        t.nKs = nKs;
        t.nRt = nRt;
        t.nPad = nPad;
        t.gateInGemm = gateIn;
        t.splitsReq = m.transferSplits;
        {
            std::vector<int> fs(n);
            for (int r = 0; r < n; r++)
                fs[r] = (r > 0 && fa[r] == fa[r - 1]) ? fs[r - 1] : r;
            SR_HIP(hipMemcpy(t.rowFs, fs.data(), n * sizeof(int), hipMemcpyHostToDevice));
        }
        // Runs of consecutive facets whose normals are bitwise identical: a pair
        // between two of them fails one of the two cosine culls (the cosines are
        // exact negatives of each other), so S is zero there as within one facet.
        std::vector<int2> run(fr.size());
        for (size_t q = 0; q < fr.size();) {
            const Vec3 &nq = o.frames[fa[fr[q].x]].normal;
            size_t e = q + 1;
            for (; e < fr.size(); e++) {
                const Vec3 &ne = o.frames[fa[fr[e].x]].normal;
                if (!(ne.x == nq.x && ne.y == nq.y && ne.z == nq.z)) break;
            }
            for (size_t k = q; k < e; k++) run[k] = make_int2(fr[q].x, fr[e - 1].y);
            q = e;
        }
        // the pairs within one run: zero in S, never stored or read
        t.zeroPairs = 0;
        for (size_t q = 0; q < fr.size(); q++)
            if (q == 0 || run[q].x != run[q - 1].x) t.zeroPairs += (double)(run[q].y - run[q].x) * (run[q].y - run[q].x);
        // Per 128-row group: its facet's range of K positions (single-facet groups
        // only) and that facet's run's, the K chunks wholly inside the run (zero:
        // never read) and, with the gate inside the product, the first K step of
        // the higher facets. With padded rows these ranges are in the padded
        // layout (whole 128-position groups), otherwise they are active cells.
        std::vector<int> rowAct(gateIn ? (size_t)nRt * 32 : 0, -1), actRow(gateIn ? n : 0);
        std::vector<int2> grp(nWg, make_int2(-1, -1)), grpRun(nWg, make_int2(-1, -1));
        if (gateIn) {
            std::vector<int> pb(fr.size() + 1, 0);  // padded start of each facet
            for (size_t q = 0; q < fr.size(); q++) pb[q + 1] = pb[q] + ((fr[q].y - fr[q].x) + 127) / 128 * 128;
            for (size_t q = 0; q < fr.size(); q++) {
                const int2 f = fr[q];
                for (int r = f.x; r < f.y; r++) {
                    rowAct[pb[q] + (r - f.x)] = r;
                    actRow[r] = pb[q] + (r - f.x);
                }
                // the run's first and last facets (runs are consecutive facets)
                size_t q0 = q, q1 = q;
                while (q0 > 0 && run[q0 - 1].x == run[q].x) q0--;
                while (q1 + 1 < fr.size() && run[q1 + 1].x == run[q].x) q1++;
                for (int w = pb[q] / 128; w < pb[q + 1] / 128; w++) {
                    grp[w] = make_int2(pb[q], pb[q + 1]);
                    grpRun[w] = make_int2(pb[q0], pb[q1 + 1]);
                }
            }
        } else {
            for (int w = 0; w < nWg; w++) {
                const int r0 = std::min(n - 1, w * 128), r1 = std::min(n - 1, w * 128 + 127);
                for (size_t q = 0; q < fr.size(); q++) {
                    if (r0 >= fr[q].x && r0 < fr[q].y && fa[r0] == fa[r1]) grp[w] = fr[q];
                    // rows of several facets of one run share its zero chunks too
                    if (r0 >= run[q].x && r0 < run[q].y && r1 < run[q].y) grpRun[w] = run[q];
                }
            }
        }
        std::vector<int2> skip(nWg, make_int2(0, 0));
        std::vector<int> kB(nWg, nKs);
        double steps = 0;
        for (int w = 0; w < nWg; w++) {
            if (grpRun[w].x >= 0) {
                const int rs = grpRun[w].x, re = grpRun[w].y;
                const int cLo = ((rs + 7) / 8 + 7) / 8, cHi = (re / 8) / 8;  // chunks wholly inside [rs, re)
                if (cHi > cLo) skip[w] = make_int2(cLo * 8, cHi * 8);
            }
            // the higher facets' first K step (used with padded rows only, where it
            // starts a chunk)
            if (grp[w].x >= 0) kB[w] = ((grp[w].y / 8) / 8) * 8;
            steps += 4.0 * (nKs - (skip[w].y - skip[w].x));
        }
        SR_HIP(hipMemcpy(t.wgSkip, skip.data(), nWg * sizeof(int2), hipMemcpyHostToDevice));
        t.aBytesRead = steps * 64 * 16;  // A bytes one GEMM pass reads
        // Split each group's non-zero chunks evenly (split_walk), at least 4 chunks a
        // split. Time in units of one workgroup streaming one chunk (32 KiB of S) while
        // the chip is full. One workgroup per CU already streams S at the rate two
        // do (measured: 2 splits of the lambda/7.5 trihedral's 126 groups, 252
        // workgroups, 0.139 ms per product; 4 splits 0.147 ms), so the grid runs in
        // rounds of one workgroup per CU. Estimate, per split count: the rounds x (the
        // most chunks a split walks + ~2 chunks of start-up and its 64 KiB partial-sum
        // store), or the chip's bytes (every chunk, and each split's partials written
        // and read back by the reduction) at full rate, whichever is longer; take the
        // least. When B exceeds an XCD's L2, only split counts that divide 8 or that 8
        // divides: the XCDs then each run one split of the groups they hold.
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || cus <= 0) cus = 256;
        const long slots = cus;
        const bool bigB = (double)nKs * 2048.0 > 4.0 * 1024 * 1024;
        long maxChunks = 0, sumChunks = 0;
        for (int w = 0; w < nWg; w++) {
            const long c = (nKs - (skip[w].y - skip[w].x)) / 8;
            maxChunks = std::max(maxChunks, c);
            sumChunks += c;
        }
        int nsplit = 1;
        double best = 1e300;
        for (int sp = 1; sp <= std::max(1L, maxChunks / 4); sp++) {
            if (bigB && 8 % sp != 0 && sp % 8 != 0) continue;
            const double rounds = (double)((nWg * (long)sp + slots - 1) / slots) * ((maxChunks + sp - 1) / sp + 2);
            const double bytes = ((double)sumChunks + 4.0 * nWg * sp) / slots;
            const double est = std::max(rounds, bytes);
            if (est < best) { best = est; nsplit = sp; }
        }
        if (m.transferSplits > 0) nsplit = (int)std::min<long>(m.transferSplits, std::max(1L, maxChunks));
        t.nsplit = nsplit;
        if (gateIn) {
            // which partials each (group, split) writes, by the product's own walk over K
            std::vector<unsigned char> cls((size_t)nWg * t.nsplit, 0);
            for (int w = 0; w < nWg; w++)
                for (int s = 0; s < t.nsplit; s++) {
                    const SplitWalk sw = split_walk(skip[w], nKs, t.nsplit, s);
                    unsigned char c = 0;
                    for (int o = sw.o0; o < sw.o1; o++) c |= sw.k(o) < kB[w] ? 1 : 2;
                    cls[(size_t)w * t.nsplit + s] = c;
                }
            // partial-sum rows one pass writes per plane: each partial a split writes, over its group's active rows
            t.partRows = 0;
            for (int w = 0; w < nWg; w++) {
                int act = 0;
                for (int k = 0; k < 128; k++) act += rowAct[(size_t)w * 128 + k] >= 0;
                for (int s = 0; s < t.nsplit; s++) t.partRows += (double)act * __builtin_popcount(cls[(size_t)w * t.nsplit + s]);
            }
            if (dev_alloc(&t.rowAct, rowAct.size()) || dev_alloc(&t.actRow, n) || dev_alloc(&t.kB, nWg) ||
                dev_alloc(&t.cls, cls.size()))
                return -1;
            SR_HIP(hipMemcpy(t.rowAct, rowAct.data(), rowAct.size() * sizeof(int), hipMemcpyHostToDevice));
            SR_HIP(hipMemcpy(t.actRow, actRow.data(), n * sizeof(int), hipMemcpyHostToDevice));
            SR_HIP(hipMemcpy(t.kB, kB.data(), nWg * sizeof(int), hipMemcpyHostToDevice));
            SR_HIP(hipMemcpy(t.cls, cls.data(), cls.size(), hipMemcpyHostToDevice));
            // Symmetric build items: block (rb, cb), rb < cb, whose K chunk the row
            // group keeps; its transpose (cb, rb) must then be kept too (and a
            // skipped block's transpose skipped), which holds when the runs are
            // whole padded facets. Every row of such a block precedes every
            // emitter (the padding keeps the active cells' order), so the build
            // reads only visibility words at or above the diagonal blocks. A kept
            // diagonal block (none when every facet spans 128 rows) would need the
            // others: no symmetric build then.
            auto kept = [&](int rb, int cb) { return cb * 8 < skip[rb >> 1].x || cb * 8 >= skip[rb >> 1].y; };
            const int nb = nKs / 8;  // 64-row blocks = K chunks
            std::vector<int2> items;
            bool symOk = nb * 64 == nPad;
            for (int rb = 0; rb < nb && symOk; rb++)
                for (int cb = rb; cb < nb; cb++) {
                    if (kept(rb, cb) != kept(cb, rb) || (cb == rb && kept(rb, cb))) { symOk = false; break; }
                    if (kept(rb, cb)) items.push_back(make_int2(rb, cb));
                }
            if (symOk && !items.empty()) {
                if (dev_alloc(&t.symItems, items.size())) return -1;
                SR_HIP(hipMemcpy(t.symItems, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
                t.nSym = (int)items.size();
            }
        } else {
            // two accumulators: every split writes both partials of every row
            t.partRows = 2.0 * t.nsplit * n;
            const std::vector<unsigned char> cls((size_t)nWg * t.nsplit, 3);
            if (dev_alloc(&t.cls, cls.size())) return -1;
            SR_HIP(hipMemcpy(t.cls, cls.data(), cls.size(), hipMemcpyHostToDevice));
        }
    }
    // One FP16 scale for S: |cos_s cos_i| <= 1, |G| <= |k|/(4 pi delta) e^{max(0,-alpha) D} + 1/delta^2
    // for r >= delta and |k|/(4 pi) below it.
    const double kabs = std::hypot((double)m.kf, (double)m.atten);
    const double D = 2.0 * m.vc.scale * std::sqrt(3.0);
    const double grow = m.atten < 0 ? std::exp(-m.atten * D) : 1.0;
    const double d = o.delta;
    const double smax = 0.9 * std::max(kabs / (4 * M_PI), kabs * grow / (4 * M_PI * d) + 1.0 / (d * d));
    t.sS = (float)(16384.0 / smax);
    TransferGeom g{o.actP, o.actN, n, nKs, nRt, t.wgSkip, t.rowAct, t.actRow};
    TransferPhys ph{m.kf, m.atten, (float)((double)m.atten * 1.4426950408889634), o.delta, 0.9f, t.sS,
                    (double)m.kf / (2.0 * M_PI), m.nearMode,
                    (float)(m.nearRadiusWl * 2.0 * M_PI / (double)m.kf), o.frameTab, nullptr,
                    nullptr, nullptr, nullptr, 0, m.nearQuadDeltas * o.delta};
    ph.streamS = (double)nRt * 32 * nKs * 16 >= kStreamSBytes;
    if (m.nearMode == 2) {
        NearField &nf = nearfield();
        if (!nf.trained) { set_error("learned near-field mode needs a trained network (sr_nf_train)"); return -3; }
        ph.mlp = nf.P;
    }
    const bool sym = m.nearMode == 0 && m.transferSymmetric && t.nSym > 0;
    // few occluders: the symmetric build evaluates the SELF visibility itself
    const bool fused = sym && m.transferSymmetric == 2 && m.ntri >= 1 && m.ntri <= 4;
    const SelfVis sv{m.trisSorted, m.ntri, o.triBase, m.vc.scale, m.vc.padTri, m.vc.tminFrag};
    uint64_t *mk = nullptr;
    if (!fused && get_mask(m, MK_SELF, OBJ_TARGET, i, OBJ_TARGET, i, &mk, sym)) return -1;
    int counts[2] = {0, 0};
    const double key[5] = {(double)m.nearMode, (double)m.nearRadiusWl, (double)m.nearQuadDeltas, (double)m.kf,
                           (double)o.delta};
    const bool known = std::equal(key, key + 5, t.nearKey);
    for (int pass = 0; pass < 2; pass++) {
        if (m.nearMode != 0) {
            if (!t.nearCount && dev_alloc(&t.nearCount, 2)) return -1;
            if (t.nearCap == 0) {
                // first guess: 64 near neighbours per fragment; grown to the count if it overflows
                t.nearCap = std::max(1024, 64 * n);
                if (dev_alloc(&t.nearQ, t.nearCap) || dev_alloc(&t.nearM, t.nearCap)) return -1;
            }
            ph.nearQ = t.nearQ; ph.nearM = t.nearM; ph.nearCount = t.nearCount; ph.nearCap = t.nearCap;
            if (!t.nearClear && counters_zero_launch(t.nearCount, 2, m.stream)) return -1;
            t.nearClear = false;  // the build counts into it
        }
        {
            // pair evaluations executed: every kept block, or those on and above the diagonal
            // pair evaluations executed, and the bytes of S written (every chunk the
            // product reads, padding included: the build's only large output)
            const double evals = sym ? 4096.0 * t.nSym : t.aBytesRead / 1024.0 * 256.0;
            ProfScope ps(m, PK_BUILD_S, (double)n * n - o.sameFacetPairs, evals, t.aBytesRead);
            if (sym ? transfer_build_s_sym(g, ph, mk, t.A, t.symItems, t.nSym, m.stream, fused ? &sv : nullptr)
                    : transfer_build_s(g, ph, mk, t.A, m.stream))
                return -1;
        }
        if (m.nearMode == 0) break;
        if (known) {  // same lists as the last build, which fitted
            counts[0] = t.nearKnown[0];
            counts[1] = t.nearKnown[1];
            break;
        }
        SR_HIP(hipMemcpyAsync(counts, t.nearCount, 2 * sizeof(int), hipMemcpyDeviceToHost, m.stream));
        SR_HIP(hipStreamSynchronize(m.stream));
        if (counts[0] <= t.nearCap && counts[1] <= t.nearCap) {
            std::copy(key, key + 5, t.nearKey);
            t.nearKnown[0] = counts[0];
            t.nearKnown[1] = counts[1];
            break;
        }
        dev_free(t.nearQ); dev_free(t.nearM);
        t.nearCap = std::max(counts[0], counts[1]) + std::max(counts[0], counts[1]) / 4;
        if (dev_alloc(&t.nearQ, t.nearCap) || dev_alloc(&t.nearM, t.nearCap)) return -1;
    }
    if (m.nearMode != 0) {
        ProfScope ps(m, PK_NEAR, (double)counts[0] + counts[1], 0.0);
        if (transfer_near(g, ph, counts[0], counts[1], t.A, m.stream)) return -1;
        t.nearClear = true;  // cleared by the near-pair kernels, or never counted into
    }
    t.valid = true;
    t.nfGen = nearfield().gen;
    return 0;
}

// One bounce of target i in the transfer form: P_out = S x for every row, with
// the zero-input gate (a row whose own P_in is zero in a plane keeps only its
// lower-facet terms there) applied by the reduction to the product's two
// partial sums, P += P_out. The product splits the terms at each 128-row
// group's facet boundary (rows padded to single-facet groups) or, for any
// other facet layout, with two accumulators per row.
// first: the first bounce of a projection; more: another bounce of this target
// follows in the same projection (its P_in is this bounce's P_out).
static int bounce_mixed(Model &m, int i, bool first, bool more) {
    Object &o = m.targets[i];
    auto &t = o.tc;
    const int n = o.nact;
    const int nsplit = t.nsplit;  // split K, chosen with the transfer matrix
    // Bounce state, two sets used in turn: this bounce's set, zero or holding the
    // column maxima of P_in that the previous bounce's reduction found; the other
    // set is cleared here for the next bounce.
    const int stWords = 1 + m.nplanes * kColStride;
    const bool haveMax = !first && t.maxReady;
    t.maxReady = false;
    if (!haveMax && !t.stateClear) {  // first use, or a bounce that stopped before clearing the next set
        if (counters_zero_launch(t.state, 2 * stWords, m.stream)) return -1;
        t.stateSet = 0;
    }
    int *cur = t.state + (size_t)t.stateSet * stWords, *nxt = t.state + (size_t)(t.stateSet ^ 1) * stWords;
    t.stateSet ^= 1;
    t.stateClear = false;  // until this bounce has cleared the next set
    unsigned *colMax = reinterpret_cast<unsigned *>(cur + 1);
    if (!haveMax) {
        ProfScope ps(m, PK_BUILD_B, 0.0, 0.0);
        if (transfer_col_max(o.actP, o.Pin, n, n, m.nplanes, colMax, nxt, stWords, m.stream)) return -1;
    }
    if (ensure_scratch(m, (size_t)nsplit * 64 * t.nPad)) return -1;  // partL and partU
    for (int g0 = 0; g0 < m.nplanes; g0 += 64) {
        const int np = std::min(64, m.nplanes - g0);
        const float2 *pin = o.Pin + (size_t)g0 * n;
        {
            // with the maxima from the previous reduction, the first group's B build clears the next set
            const bool clr = haveMax && g0 == 0;
            ProfScope ps(m, PK_BUILD_B, 0.0, 0.0);
            if (transfer_build_b(o.actP, pin, n, n, t.nKs, np, colMax + (size_t)g0 * kColStride, t.sS, t.colScale,
                                 t.colInv, t.B, m.stream, clr ? nxt : nullptr, clr ? stWords : 0,
                                 t.gateInGemm ? t.rowAct : nullptr))
                return -1;
        }
        TransferGemm gm{t.A, t.B, t.nPad, t.nKs, t.nRt, np, nsplit, t.wgSkip, m.scratchL,
                        t.gateInGemm ? t.kB : nullptr, m.scratchU, t.gateInGemm ? nullptr : t.rowFs};
        gm.streamS = (double)t.nRt * 32 * t.nKs * 16 >= kStreamSBytes;
        // the split-K reduction, the gate and (another bounce following) its column maxima,
        // done by each row group's last workgroup of the product
        gm.red = ReduceGateTask{m.scratchL, m.scratchU, t.cls, nsplit, t.nPad, np, n, t.rowAct, pin,
                                o.Pout + (size_t)g0 * n, o.P + (size_t)g0 * o.ncell, o.ncell, o.actCell, t.colInv,
                                o.actP,
                                more ? reinterpret_cast<unsigned *>(nxt + 1) + (size_t)g0 * kColStride : nullptr};
        gm.grpCount = m.reduceInGemm ? t.grpCount : nullptr;
        m.lastGemm = gm;
        m.lastGemm.grpCount = nullptr;  // replays time the product alone
        m.haveLastGemm = true;
        {
            const double pairs = ((double)n * n - o.sameFacetPairs) * np;
            // flops of the GEMM as executed: 2 * rows * 128 columns * 2 * K (padded)
            const double fl = 2.0 * (double)t.nRt * 32 * 128 * (double)t.nKs * 16;
            // algorithmic HBM bytes of the bounce product: the FP16 complex S entry of every
            // ordered pair of fragments not on one plane (4 B: pairs within a facet, or within
            // a run of adjacent facets with identical normals, are zero, never stored or read),
            // the FP16 complex input of every active cell and plane (4 B) and the FP32 complex
            // output P_out (8 B); partial sums, padding and re-reads are traffic, not work
            const double by = ((double)n * n - t.zeroPairs) * 4.0 + (double)n * np * 4.0 + (double)n * np * 8.0;
            ProfScope ps(m, PK_GEMM, pairs, fl, by);
            if (transfer_gemm(gm, m.stream)) return -1;
        }
        if (!gm.grpCount) {  // the same reduction as a pass of its own
            ProfScope ps(m, PK_REDUCE, 0.0, 0.0);
            if (transfer_reduce_gate(gm.red, m.stream)) return -1;
        }
    }
    if (more) t.maxReady = true;
    else t.stateClear = true;
    return 0;
}

// FACET_SELF bounces (ModelCuda.cpp:323-356, FacetCuda.cpp:124-137, FacetCuda.cu:87-116).
int model_project_t2t(Model &m, int nref) {
    if (int rc = check_rendered(m)) return rc;
    if (nref < 0) { set_error("number_reflections must be >= 0"); return -2; }
    for (auto &o : m.targets)  // PrimeReflections: P_out <- P
        if (o.nact && prime_launch(o.P, o.ncell, o.actCell, o.nact, m.nplanes, o.Pout, o.nact, m.stream)) return -1;
    for (int it = 0; it < nref; it++)
        for (int i = 0; i < (int)m.targets.size(); i++) {
            Object &o = m.targets[i];
            if (!o.nact) continue;
            std::swap(o.Pin, o.Pout);  // SwapOutputToInputPressure; P_out is fully rewritten below
            if (m.transferMode == 1) {
                if (ensure_transfer(m, i) || bounce_mixed(m, i, it == 0, it + 1 < nref)) return -1;
                continue;
            }
            uint64_t *mk;
            if (get_mask(m, MK_SELF, OBJ_TARGET, i, OBJ_TARGET, i, &mk)) return -1;
            Phys ph = make_phys(m, o.delta, 0.9f);
            GatherTask t{};
            t.ep = o.actP; t.en = o.actN; t.ecell = nullptr; t.eval = o.Pin; t.ne = o.nact;
            t.rp = o.actP; t.rn = o.actN; t.nr = o.nact; t.mask = mk;
            t.gval = o.Pin; t.gateMode = GATE_UPPER;
            ReduceTask r{};
            r.gval = o.Pin; r.gateMode = GATE_UPPER;
            r.out = o.Pout; r.pcell = o.P; r.rowCell = o.actCell;  // P += P_out (AccumulatePressure)
            const double pairs = (double)o.nact * o.nact - o.sameFacetPairs;
            if (gather_planes(m, t, ph, KIND_F2F, r, m.nplanes, o.nact, o.nact, o.nact, o.ncell, pairs)) return -1;
        }
    return finish_projection(m);
}

// FACET_NO_RESP: every target into every field object (ModelCuda.cpp:280-301).
int model_project_t2field_objects(Model &m) {
    if (int rc = check_rendered(m)) return rc;
    for (int i = 0; i < (int)m.targets.size(); i++)
        for (int j = 0; j < (int)m.fields.size(); j++) {
            Object &A = m.targets[i], &B = m.fields[j];
            if (!A.nact || !B.nact) continue;
            uint64_t *mk;
            if (get_mask(m, MK_NORESP, OBJ_TARGET, i, OBJ_FIELD, j, &mk)) return -1;
            Phys ph = make_phys(m, A.delta, 0.9f);
            GatherTask t{};
            t.ep = A.actP; t.en = A.actN; t.ecell = A.actCell; t.eval = A.P; t.ne = A.nact;
            t.rp = B.actP; t.rn = B.actN; t.nr = B.nact; t.mask = mk;
            t.gateMode = GATE_NONE;
            ReduceTask r{};
            r.gateMode = GATE_NONE;
            r.pcell = B.P; r.rowCell = B.actCell;  // field objects alias P = P_in = P_out
            if (gather_planes(m, t, ph, KIND_F2F_FIELD, r, m.nplanes, A.ncell, 0, 0, B.ncell, (double)A.nact * B.nact))
                return -1;
        }
    return finish_projection(m);
}

// FIELD_POINTS (ModelCuda.cpp:302-321): compensated FP32 gathers, FP64 totals.
int model_project_t2field_points(Model &m) {
    if (int rc = check_rendered(m)) return rc;
    const int nq = (int)(m.fldHost.size() / 3);
    if (!nq) return 0;
    for (int i = 0; i < (int)m.targets.size(); i++) {
        Object &A = m.targets[i];
        if (!A.nact) continue;
        uint64_t *mk;
        if (get_mask(m, MK_F2P, OBJ_TARGET, i, 0, 0, &mk)) return -1;
        Phys ph = make_phys(m, 0.f, 1.f);
        GatherTask t{};
        t.ep = A.actP; t.en = A.actN; t.ecell = A.actCell; t.eval = A.P; t.ne = A.nact;
        t.rp = m.fldP; t.rn = nullptr; t.nr = nq; t.mask = mk;
        t.gateMode = GATE_NONE;
        ReduceTask r{};
        if (gather_planes(m, t, ph, KIND_F2P, r, m.nplanes, A.ncell, 0, 0, 0, (double)A.nact * nq)) return -1;
    }
    return finish_projection(m);
}

// FIELD_POINTS without any allocation: the split-K partial sums go to the fixed
// buffer `scratch` (two halves of `elems` complex values) instead of the
// library's scratch, profiling records nothing, and the visibility words must
// already be allocated (render_cuda does that). Rejects work whose partial sums
// do not fit even in one split.
int model_project_t2field_points_noalloc(Model &m, float2 *scratch, size_t elems) {
    if (int rc = check_rendered(m)) return rc;
    struct Guard {
        Model &m;
        ~Guard() { m.noAlloc = false; m.fixL = m.fixU = nullptr; m.fixElems = 0; }
    } guard{m};
    m.noAlloc = true;
    m.fixElems = elems / 2;
    m.fixL = scratch;
    m.fixU = scratch + m.fixElems;
    return model_project_t2field_points(m);
}

long model_device_allocations() { return gDevAllocs.load(); }

// FACET_RESP between two target objects (OptiX.cu:427-430; never dispatched by
// the reference's orchestration). Accumulates into both objects' P_out.
int model_project_resp(Model &m, int ia, int ib) {
    if (int rc = check_rendered(m)) return rc;
    Object *A = object_at(m, OBJ_TARGET, ia), *B = object_at(m, OBJ_TARGET, ib);
    if (!A || !B || ia == ib) { set_error("RESP needs two distinct target objects"); return -2; }
    if (!A->nact || !B->nact) return 0;
    uint64_t *mf, *mr;
    if (get_mask(m, MK_RESP_FWD, OBJ_TARGET, ia, OBJ_TARGET, ib, &mf)) return -1;
    if (get_mask(m, MK_RESP_REC, OBJ_TARGET, ia, OBJ_TARGET, ib, &mr)) return -1;
    Phys ph = make_phys(m, A->delta, 0.9f);
    // forward: rows of B, emitters of A (skip when P_in_A == 0 is implicit: x = 0)
    GatherTask t{};
    t.ep = A->actP; t.en = A->actN; t.eval = A->Pin; t.ne = A->nact;
    t.rp = B->actP; t.rn = B->actN; t.nr = B->nact; t.mask = mf; t.gateMode = GATE_NONE;
    ReduceTask r{};
    r.gateMode = GATE_NONE; r.out = B->Pout;
    r.outAdd = 1;
    if (gather_planes(m, t, ph, KIND_F2F, r, m.nplanes, A->nact, 0, B->nact, 0, (double)A->nact * B->nact)) return -1;
    // reciprocal: rows of A gated by their own P_in, emitters of B
    GatherTask u{};
    u.ep = B->actP; u.en = B->actN; u.eval = B->Pin; u.ne = B->nact;
    u.rp = A->actP; u.rn = A->actN; u.nr = A->nact; u.mask = mr; u.gateMode = GATE_NONE;
    ReduceTask s{};
    s.gateMode = GATE_ALL; s.gval = A->Pin; s.out = A->Pout; s.outAdd = 1;
    if (gather_planes(m, u, ph, KIND_F2F, s, m.nplanes, B->nact, A->nact, A->nact, 0, (double)A->nact * B->nact))
        return -1;
    return finish_projection(m);
}

// P_in <- P (active cells), P_out <- 0: the state a RESP call reads.
int model_prepare_resp(Model &m, int it) {
    if (int rc = check_rendered(m)) return rc;
    Object *o = object_at(m, OBJ_TARGET, it);
    if (!o) { set_error("target index out of range"); return -2; }
    if (!o->nact) return 0;
    if (prime_launch(o->P, o->ncell, o->actCell, o->nact, m.nplanes, o->Pin, o->nact, m.stream)) return -1;
    SR_HIP(hipMemsetAsync(o->Pout, 0, (size_t)m.nplanes * o->nact * sizeof(float2), m.stream));
    SR_HIP(hipStreamSynchronize(m.stream));
    return 0;
}

int model_get_field_points(Model &m, double *out, long n, int plane) {
    const int nq = (int)(m.fldHost.size() / 3);
    if (!m.rendered) { set_error("no field points on the GPU: call render_cuda first"); return -3; }
    if (!out && n > 0) { set_error("null output buffer"); return -2; }
    if (plane < 0 || plane >= m.nplanes) { set_error("plane out of range"); return -2; }
    if (int rc = settle(m)) return rc;
    const long cnt = std::min<long>(n, nq);
    if (cnt > 0)
        SR_HIP(hipMemcpy(out, m.fldVal + (size_t)plane * nq, cnt * sizeof(double2), hipMemcpyDeviceToHost));
    return (int)cnt;
}

int model_get_object_pressure(Model &m, int type, int idx, int which, int plane, double *out, long n) {
    if (!m.rendered) { set_error("call render_cuda first"); return -3; }
    Object *o = object_at(m, type, idx);
    if (!o) { set_error("object index out of range"); return -2; }
    if (plane < 0 || plane >= m.nplanes) { set_error("plane out of range"); return -2; }
    if (n < o->ncell) { set_error("output buffer shorter than the object's cell count"); return -2; }
    if (int rc = settle(m)) return rc;
    std::vector<float2> h(o->ncell, make_float2(0.f, 0.f));
    if (which == 0 || o->type == OBJ_FIELD) {
        if (o->ncell)
            SR_HIP(hipMemcpy(h.data(), o->P + (size_t)plane * o->ncell, o->ncell * sizeof(float2), hipMemcpyDeviceToHost));
    } else {
        std::vector<float2> a(o->nact);
        const float2 *src = (which == 1) ? o->Pin : o->Pout;
        if (o->nact)
            SR_HIP(hipMemcpy(a.data(), src + (size_t)plane * o->nact, o->nact * sizeof(float2), hipMemcpyDeviceToHost));
        for (int k = 0; k < o->nact; k++) h[o->actCellHost[k]] = a[k];
    }
    for (long c = 0; c < o->ncell; c++) { out[2 * c] = h[c].x; out[2 * c + 1] = h[c].y; }
    return 0;
}

int model_get_mask(Model &m, int kind, int ta, int ia, int tb, int ib, uint64_t *out, long nwords) {
    if (int rc = check_rendered(m)) return rc;
    uint64_t *mk;
    if (get_mask(m, kind, ta, ia, tb, ib, &mk) || finish_projection(m) || settle(m)) return -1;
    SR_HIP(hipMemcpy(out, mk, nwords * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return 0;
}

// Queued behind any work in flight (stream order): scenes can follow each other
// without the host waiting in between.
int model_reset_pressures(Model &m) {
    if (int rc = check_rendered(m)) return rc;
    const long np = m.nplanes;
    std::vector<float2 *> p;
    std::vector<long> n;
    auto add = [&](float2 *q, long cnt) {
        if (q && cnt > 0) { p.push_back(q); n.push_back(cnt); }
    };
    for (auto *v : {&m.targets, &m.fields})
        for (auto &o : *v) {
            add(o.P, np * o.ncell);
            add(o.Pin, np * o.nact);
            add(o.Pout, np * o.nact);
        }
    add(reinterpret_cast<float2 *>(m.fldVal), 2 * np * (long)(m.fldHost.size() / 3));  // double2 = 2 float2
    return zero_ranges_launch(p.data(), n.data(), (int)p.size(), m.stream);
}

int model_clear_vis_cache(Model &m) {
    // the words are rewritten by later work on the same stream: no wait needed
    for (auto &kv : m.masks) kv.second.valid = false;  // buffers are kept and rewritten
    for (auto &o : m.targets) o.tc.valid = false;
    return 0;
}

int model_set_near_mode(Model &m, int mode, float radiusWl) {
    if (mode < 0 || mode > 2) { set_error("near mode: 0 point rule, 1 quadrature, 2 learned"); return -2; }
    if (!(radiusWl > 0.f)) { set_error("near radius must be > 0 wavelengths"); return -2; }
    m.nearMode = mode;
    m.nearRadiusWl = radiusWl;
    for (auto &o : m.targets) o.tc.valid = false;
    return 0;
}

int model_set_near_quadrature_radius(Model &m, float deltas) {
    if (!(deltas >= 1.0f) || deltas > 16.0f) { set_error("quadrature radius: 1 to 16 fragment sizes"); return -2; }
    m.nearQuadDeltas = deltas;
    for (auto &o : m.targets) o.tc.valid = false;
    return 0;
}

int model_set_gather_mode(Model &m, int mode) {
    if (mode != 0 && mode != 1) { set_error("gather mode: 0 FP32 MFMA products from 16 planes on, 1 VALU only"); return -2; }
    m.gatherMode = mode;
    return 0;
}

int model_set_transfer_splits(Model &m, int n) {
    if (n < 0 || n > 4096) { set_error("transfer splits: 0 (automatic) to 4096"); return -2; }
    m.transferSplits = n;
    return 0;
}

int model_set_transfer_symmetric(Model &m, int on) {
    if (on < 0 || on > 2) {
        set_error("transfer symmetric build: 0 every block, 1 half and transposes, 2 also the visibility inside");
        return -2;
    }
    if (on != m.transferSymmetric)
        for (auto &o : m.targets) o.tc.valid = false;  // built the other way: rebuild
    m.transferSymmetric = on;
    return 0;
}

int model_set_reduce_in_product(Model &m, int on) {
    if (on != 0 && on != 1) { set_error("reduce in product: 0 a pass of its own, 1 inside the product"); return -2; }
    m.reduceInGemm = on;
    return 0;
}

int model_set_inline_visibility(Model &m, int on) {
    if (on != 0 && on != 1) { set_error("inline visibility: 0 from the visibility words, 1 inside the sums"); return -2; }
    m.inlineVis = on;
    return 0;
}

int model_set_transfer_mode(Model &m, int mode) {
    if (mode != 0 && mode != 1) { set_error("transfer mode: 0 plain FP32, 1 FP16 MFMA transfer matrix"); return -2; }
    m.transferMode = mode;
    return 0;
}

void model_profile(Model &m, int on) {
    model_profile_reset(m);
    m.profiling = on != 0;
    if (m.profiling && !m.workCnt) {  // kept for the life of the process once profiling was used
        if (hipMalloc((void **)&m.workCnt, 4096 * sizeof(unsigned long long)) == hipSuccess) m.workCap = 4096;
        else m.workCnt = nullptr;
    }
    if (m.workCnt) {  // the library's stream orders it before the counting launches
        (void)hipMemsetAsync(m.workCnt, 0, m.workCap * sizeof(unsigned long long), m.stream);
        if (m.stream) (void)hipStreamSynchronize(m.stream);
    }
    m.workUsed = 0;
}

void model_profile_reset(Model &m) {
    if (m.stream) (void)hipStreamSynchronize(m.stream);
    if (m.side) (void)hipStreamSynchronize(m.side);
    for (auto &r : m.recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    m.recs.clear();
    if (m.workCnt) {
        (void)hipMemsetAsync(m.workCnt, 0, m.workCap * sizeof(unsigned long long), m.stream);
        if (m.stream) (void)hipStreamSynchronize(m.stream);
    }
    m.workUsed = 0;
}

int model_kernel_stats(Model &m, int kind, double *ms, long *launches, double *pairs, double *flops, double *bytes) {
    if (m.stream) SR_HIP(hipStreamSynchronize(m.stream));
    if (m.side) SR_HIP(hipStreamSynchronize(m.side));
    double t = 0, p = 0, f = 0, b = 0;
    long n = 0;
    std::vector<unsigned long long> work(m.workUsed);
    if (m.workUsed)
        SR_HIP(hipMemcpy(work.data(), m.workCnt, m.workUsed * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (auto &r : m.recs) {
        if (r.kind != kind) continue;
        float e = 0.f;
        SR_HIP(hipEventElapsedTime(&e, r.a, r.b));
        const double fl = r.workSlot >= 0 ? (double)work[r.workSlot] * r.flopsPerStep : r.flops;
        t += e; p += r.pairs; f += fl; b += r.bytes; n++;
    }
    *ms = t; *launches = n; *pairs = p; *flops = f;
    if (bytes) *bytes = b;
    return 0;
}

int model_time_transfer_gemm(Model &m, int reps, double *msPerLaunch) {
    if (!m.haveLastGemm || !m.rendered) { set_error("no transfer-form bounce has run since render_cuda"); return -3; }
    if (reps < 1 || reps > 10000) { set_error("reps: 1 to 10000"); return -2; }
    // the partial sums go to the current scratch (later projections may have regrown it)
    TransferGemm g = m.lastGemm;
    if ((size_t)g.nsplit * 64 * g.n > m.scratchElems) { set_error("scratch smaller than the product's partials"); return -3; }
    g.partL = m.scratchL;
    if (g.partU) g.partU = m.scratchU;
    hipEvent_t a, b;
    SR_HIP(hipEventCreate(&a));
    SR_HIP(hipEventCreate(&b));
    int rc = transfer_gemm(g, m.stream);  // warm
    if (!rc) SR_HIP(hipEventRecord(a, m.stream));
    for (int r = 0; !rc && r < reps; r++) rc = transfer_gemm(g, m.stream);
    float ms = 0.f;
    if (!rc) {
        SR_HIP(hipEventRecord(b, m.stream));
        SR_HIP(hipEventSynchronize(b));
        SR_HIP(hipEventElapsedTime(&ms, a, b));
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    *msPerLaunch = (double)ms / reps;
    return 0;
}

void model_teardown(Model &m) {
    model_profile_reset(m);
    if (m.stream) (void)hipStreamSynchronize(m.stream);
    if (m.side) (void)hipStreamSynchronize(m.side);
    for (auto &o : m.targets) free_object(o);
    for (auto &o : m.fields) free_object(o);
    for (auto &o : m.sources) free_object(o);
    m.targets.clear();
    m.fields.clear();
    m.sources.clear();
    m.srcObjAmp.clear();
    for (auto &kv : m.masks) (void)hipFree(kv.second.p);
    m.masks.clear();
    m.transferMode = 0;
    m.transferSplits = 0;
    m.transferSymmetric = 2;
    m.inlineVis = 1;
    m.reduceInGemm = 0;
    m.haveLastGemm = false;
    m.nearMode = 0;
    m.nearQuadDeltas = kNfRMinDeltas;
    m.gatherMode = 0;
    dev_free(m.gatedList); dev_free(m.gatedCount);
    m.gatedCap = 0;
    dev_free(m.gPc); dev_free(m.gNc); dev_free(m.gFs); dev_free(m.gMc);
    m.gCap = 0;
    m.gMcCap = 0;
    dev_free(m.srcP); dev_free(m.srcAmp); dev_free(m.fldP); dev_free(m.fldVal);
    dev_free(m.trisSorted); dev_free(m.nodes);
    dev_free(m.scratchL); dev_free(m.scratchU); dev_free(m.scratchC);
    m.scratchElems = 0;
    m.scratchCElems = 0;
    m.srcHost.clear();
    m.fldHost.clear();
    m.ntri = 0;
    m.rendered = false;
    m.batchedSources = false;
    m.nplanes = 1;
}

void model_release(Model &m) {
    model_teardown(m);
    if (m.stream) (void)hipStreamSynchronize(m.stream);
    if (m.side) (void)hipStreamSynchronize(m.side);
    for (auto &e : m.wordEv)
        if (e) (void)hipEventDestroy(e);
    if (m.sideIn) (void)hipEventDestroy(m.sideIn);
    if (m.sideOut) (void)hipEventDestroy(m.sideOut);
    if (m.side) (void)hipStreamDestroy(m.side);
    if (m.stream) (void)hipStreamDestroy(m.stream);
    if (m.hostWords) (void)hipHostFree(m.hostWords);
    if (m.errHost) (void)hipHostFree(m.errHost);
    if (m.workCnt) (void)hipFree(m.workCnt);
    m.stream = m.side = nullptr;
    m.sideIn = m.sideOut = nullptr;
    for (auto &e : m.wordEv) e = nullptr;
    m.hostWords = m.errHost = m.errFlag = nullptr;
    m.workCnt = nullptr;
    m.workCap = m.workUsed = 0;
}

}  // namespace sr
