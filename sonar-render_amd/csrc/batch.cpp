// Batched uneven scenes.
//
// The reference runs one scene per process (run_scripts/modules/mh.py:14-30:
// load, render, project, tear down, repeat). Here many independent scenes
// (each its own mesh, fragment size, frequency, source and field points) are
// prepared once and projected together: every phase is one launch over all
// scenes, the heavy ones persistent (workgroups pull (scene, block, emitter
// range) items from a device counter, largest scenes first, so uneven scenes
// balance), and the whole sequence can be captured once and replayed as a
// single hipGraph. Pressures and counters are reset by kernels inside the
// sequence (no memset nodes in the graph).
//
// Per scene the physics and the visibility definition are exactly those of
// the single-scene path (same kernels, same task descriptors).
#include "batch.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <numeric>
#include <vector>

#include "launch.h"
#include "model.h"
#include "model_ops.h"
#include "sums.h"
#include "batch_near.h"
#include "learned.h"
#include "nearfield.h"

namespace sr {

int vis_mask_launch(const MaskTask &tk, const Bvh &bvh, const VisConst &vc, hipStream_t st,
                    bool deferLower = false);
int vis_mask_multi_launch(const MaskTask *tasks, const Bvh *bvhs, const VisConst *vcs, const int4 *items, int nitems,
                          int emPerWave, int *counter, int grid, hipStream_t st);
int mask_transpose_multi_launch(const MaskTask *tasks, const int2 *items, int nitems, hipStream_t st);

namespace {

template <class T>
struct DevVec {
    T *p = nullptr;
    size_t n = 0;
    int upload(const std::vector<T> &h) {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = h.size();
        if (!n) return 0;
        SR_HIP(hipMalloc((void **)&p, n * sizeof(T)));
        SR_HIP(hipMemcpy(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice));
        return 0;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct SceneIn {
    std::vector<float> tris, src, pts;
    float delta, cp, f, atten;
};

struct Batch {
    std::vector<SceneIn> in;
    std::vector<std::unique_ptr<Model>> scenes;
    int bounces = 2;
    bool prepared = false;
    hipStream_t st = nullptr;
    BatchTune tune;
    // near-field correction of the bounces: 0 off (the reference's point rule),
    // 1 16x16 quadrature, 2 learned (quadrature below nearQuadDeltas fragment sizes)
    int nearMode = 0;
    float nearRadiusWl = 2.0f, nearQuadDeltas = kNfRMinDeltas;
    // per scene: CSR of the near pairs (row offsets fixed at prepare), their values
    std::vector<int *> nearRowStart, nearRows, nearCols;
    std::vector<float2 *> nearVals;
    std::vector<long> nearTotal;
    DevVec<NearBatchTask> nearTasks;
    DevVec<int2> nearRowItems, nearMlpItems, nearQuadItems, nearApplyItems;
    std::vector<DevVec<NearApplyTask>> nearApply;
    // per-scene mask buffers
    std::vector<uint64_t *> maskBufs;
    float2 *scratchL = nullptr, *scratchU = nullptr;
    size_t scratchElems = 0;
    int *counters = nullptr;
    // task tables
    DevVec<MaskTask> visTasks;
    DevVec<Bvh> visBvh;
    DevVec<VisConst> visVc;
    DevVec<int4> visItems;
    DevVec<int2> trItems;
    DevVec<P2FTask> p2fTasks;
    DevVec<Phys> p2fPhys;
    DevVec<int2> p2fItems;
    DevVec<ZeroTask> zeroTasks;
    DevVec<int2> zeroItems;
    DevVec<PrimeTask> primeTasks;
    DevVec<int2> primeItems;
    std::vector<DevVec<GatherTask>> selfTasks;
    DevVec<Phys> selfPhys;
    DevVec<int4> selfItems;
    std::vector<DevVec<ReduceTask>> redTasks;
    DevVec<int2> redItems;
    DevVec<GatherTask> f2pTasks;
    DevVec<Phys> f2pPhys;
    DevVec<int4> f2pItems;
    DevVec<PointsReduceTask> prTasks;
    DevVec<int2> prItems;
    bool att = false;
    double pairs = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

Batch &batch() {
    static Batch b;
    return b;
}

void drop_graph(Batch &b) {
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
}

void release_tables(Batch &b) {
    drop_graph(b);
    b.visTasks.release(); b.visBvh.release(); b.visVc.release(); b.visItems.release(); b.trItems.release();
    b.p2fTasks.release(); b.p2fPhys.release(); b.p2fItems.release();
    b.zeroTasks.release(); b.zeroItems.release(); b.primeTasks.release(); b.primeItems.release();
    for (auto &t : b.selfTasks) t.release();
    for (auto &t : b.redTasks) t.release();
    b.selfTasks.clear(); b.redTasks.clear();
    b.selfPhys.release(); b.selfItems.release(); b.redItems.release();
    b.f2pTasks.release(); b.f2pPhys.release(); b.f2pItems.release(); b.prTasks.release(); b.prItems.release();
    b.nearTasks.release(); b.nearRowItems.release(); b.nearMlpItems.release(); b.nearQuadItems.release();
    b.nearApplyItems.release();
    for (auto &t : b.nearApply) t.release();
    b.nearApply.clear();
}

void release_near(Batch &b) {
    for (auto *v : {&b.nearRowStart, &b.nearRows, &b.nearCols})
        for (int *p : *v) if (p) (void)hipFree(p);
    for (float2 *p : b.nearVals) if (p) (void)hipFree(p);
    b.nearRowStart.clear(); b.nearRows.clear(); b.nearCols.clear(); b.nearVals.clear(); b.nearTotal.clear();
}

// The near-pair task of scene s (fill mode when its CSR exists).
NearBatchTask near_task(const Batch &b, int s, uint64_t *mSelf, int *rowCount) {
    const Model &m = *b.scenes[s];
    const Object &o = m.targets[0];
    NearBatchTask t{};
    t.p = o.actP; t.nrm = o.actN; t.frames = o.frameTab; t.mask = mSelf; t.n = o.nact;
    const double lam = 2.0 * M_PI / (double)m.kf;
    const double R = b.nearRadiusWl * lam, Rq = (b.nearMode == 2 ? b.nearQuadDeltas : 0.0) * o.delta;
    t.radius2 = (float)(R * R);
    t.quadR2 = (float)(Rq * Rq);
    t.kr = m.kf; t.alpha = m.atten; t.delta = o.delta; t.mode = b.nearMode;
    if (!b.nearRowStart.empty()) {
        t.rowStart = b.nearRowStart[s]; t.rows = b.nearRows[s]; t.cols = b.nearCols[s]; t.vals = b.nearVals[s];
    }
    t.rowCount = rowCount;
    return t;
}

void release_all(Batch &b) {
    release_tables(b);
    release_near(b);
    for (auto *p : b.maskBufs) (void)hipFree(p);
    b.maskBufs.clear();
    if (b.scratchL) (void)hipFree(b.scratchL);
    if (b.scratchU) (void)hipFree(b.scratchU);
    if (b.counters) (void)hipFree(b.counters);
    b.scratchL = b.scratchU = nullptr;
    b.counters = nullptr;
    b.scratchElems = 0;
    for (auto &m : b.scenes) {
        m->stream = nullptr;  // the batch owns the stream
        model_release(*m);    // also the scene's side stream, events and pinned words
    }
    b.scenes.clear();
    b.prepared = false;
}

Phys phys_of(const Model &m, float deltaReg, float damp) {
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

// Builds every task table and work list from the prepared scenes and the tuning.
int build_tables(Batch &b) {
    release_tables(b);
    const int S = (int)b.scenes.size();
    std::vector<MaskTask> vt;
    std::vector<Bvh> vb;
    std::vector<VisConst> vv;
    std::vector<int4> vi;
    std::vector<int2> ti;
    std::vector<P2FTask> pt;
    std::vector<Phys> pp, sp, fp;
    std::vector<int2> pi, zi, qi, ri, rpi;
    std::vector<ZeroTask> zt;
    std::vector<PrimeTask> qt;
    std::vector<std::vector<GatherTask>> st(b.bounces);
    std::vector<std::vector<ReduceTask>> rt(b.bounces);
    std::vector<int4> si, fi;
    std::vector<GatherTask> ft;
    std::vector<PointsReduceTask> prt;
    std::vector<std::vector<NearApplyTask>> na(b.bounces);
    std::vector<NearBatchTask> nt;
    std::vector<int2> nri, nmi, nqi, nai;
    size_t scratchSelf = 0, scratchF2P = 0;
    std::vector<size_t> offSelf(S), offF2P(S);
    std::vector<int> nsSelf(S), nsF2P(S), perSelf(S), perF2P(S);
    b.pairs = 0;
    b.att = false;
    // scene order: largest first (persistent workgroups then finish together)
    std::vector<int> order(S);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(),
              [&](int x, int y) { return b.scenes[x]->targets[0].nact > b.scenes[y]->targets[0].nact; });
    for (int s = 0; s < S; s++) {
        Model &m = *b.scenes[s];
        Object &o = m.targets[0];
        const int nq = (int)(m.fldHost.size() / 3);
        if (m.atten != 0.f) b.att = true;
        const int nrb = (o.nact + 63) / 64;
        nsSelf[s] = std::max(1, std::min(64, o.nact / std::max(1, b.tune.minPerSplit)));
        perSelf[s] = (o.nact + nsSelf[s] - 1) / nsSelf[s];
        nsSelf[s] = std::max(1, (o.nact + perSelf[s] - 1) / perSelf[s]);
        nsF2P[s] = std::max(1, std::min(256, o.nact / 16));
        perF2P[s] = (o.nact + nsF2P[s] - 1) / nsF2P[s];
        nsF2P[s] = std::max(1, (o.nact + perF2P[s] - 1) / perF2P[s]);
        offSelf[s] = scratchSelf;
        scratchSelf += (size_t)(nsSelf[s] + (b.nearMode ? 1 : 0)) * o.nact;  // + the near-field split
        offF2P[s] = scratchF2P;
        scratchF2P += (size_t)nsF2P[s] * nq;
        (void)nrb;
        b.pairs += (double)o.nact + (double)b.bounces * ((double)o.nact * o.nact - o.sameFacetPairs) +
                   (double)o.nact * nq;
    }
    const size_t need = std::max(scratchSelf, scratchF2P);
    if (need > b.scratchElems) {
        if (b.scratchL) (void)hipFree(b.scratchL);
        if (b.scratchU) (void)hipFree(b.scratchU);
        b.scratchL = b.scratchU = nullptr;
        SR_HIP(hipMalloc((void **)&b.scratchL, need * sizeof(float2)));
        SR_HIP(hipMalloc((void **)&b.scratchU, need * sizeof(float2)));
        b.scratchElems = need;
    }
    for (int oi = 0; oi < S; oi++) {
        const int s = order[oi];
        Model &m = *b.scenes[s];
        Object &o = m.targets[0];
        const int nq = (int)(m.fldHost.size() / 3);
        // visibility words: sources -> cells, SELF, fragments -> field points
        const int kinds[3] = {0, 2, 1};
        for (int kk = 0; kk < 3; kk++) {
            MaskTask tk;
            size_t words;
            const int kind = kinds[kk];
            if (mask_task(m, kind, 0, 0, 0, 0, tk, words)) return -1;
            tk.mask = b.maskBufs[3 * (size_t)s + kk];  // allocated in scene order, 3 per scene
            tk.errFlag = b.counters + 2 + b.bounces;     // one traversal-overflow flag for the batch
            const int task = (int)vt.size();
            vt.push_back(tk);
            vb.push_back(Bvh{m.trisSorted, m.nodes, m.ntri});
            vv.push_back(m.vc);
            const int nrbk = (tk.rc.count + 63) / 64;
            const int nec = (tk.em.count + b.tune.visEmPerWave - 1) / b.tune.visEmPerWave;
            for (int rb = 0; rb < nrbk; rb++)
                for (int ec = 0; ec < nec; ec++) {
                    if (tk.upperOnly && (ec + 1) * b.tune.visEmPerWave <= rb * 64) continue;
                    vi.push_back(make_int4(task, rb, ec, 0));
                }
            if (tk.upperOnly)
                for (long p = 0; p < (long)nrbk * (nrbk - 1) / 2; p++) ti.push_back(make_int2(task, (int)p));
        }
        uint64_t *mP2F = vt[vt.size() - 3].mask, *mSelf = vt[vt.size() - 2].mask, *mF2P = vt[vt.size() - 1].mask;
        if (b.nearMode) {  // near pairs: fill the CSR, evaluate, one apply task per bounce (below)
            const int task = (int)nt.size();
            nt.push_back(near_task(b, s, mSelf, nullptr));
            for (int rb = 0; rb < (o.nact + 63) / 64; rb++) nri.push_back(make_int2(task, rb));
            const long tot = b.nearTotal[s];
            for (long k = 0; k < tot; k += 16) nqi.push_back(make_int2(task, (int)k));
            if (b.nearMode == 2)
                for (long k = 0; k < tot; k += 128) nmi.push_back(make_int2(task, (int)k));
            for (long e = 0; e < o.nact; e += 256) nai.push_back(make_int2(task, (int)e));
        }
        // pressure resets (P, field values)
        {
            const int z0 = (int)zt.size();
            zt.push_back(ZeroTask{(float *)o.P, 2 * o.ncell});
            zt.push_back(ZeroTask{(float *)m.fldVal, 4L * nq});
            for (int z = z0; z < z0 + 2; z++)
                for (long e = 0; e < zt[z].n; e += 256) zi.push_back(make_int2(z, (int)e));
        }
        // point -> fragment
        {
            const int task = (int)pt.size();
            pt.push_back(P2FTask{o.cellP, o.cellN, (int)o.ncell, 1, m.srcP, m.srcAmp, (int)(m.srcHost.size() / 3),
                                 mP2F, 0, o.P, o.ncell});
            pp.push_back(phys_of(m, 0.f, 1.f));
            for (long c = 0; c < o.ncell; c += 256) pi.push_back(make_int2(task, (int)c));
        }
        // bounces: P_out of bounce 0 is P; buffers alternate Pout, Pin
        float2 *buf[2] = {o.Pout, o.Pin};
        {
            const int task = (int)qt.size();
            qt.push_back(PrimeTask{o.P, o.ncell, o.actCell, o.nact, 1, buf[0], o.nact});
            for (long e = 0; e < o.nact; e += 256) qi.push_back(make_int2(task, (int)e));
        }
        const int task = (int)sp.size();
        sp.push_back(phys_of(m, o.delta, 0.9f));
        for (int bo = 0; bo < b.bounces; bo++) {
            float2 *pin = buf[bo % 2], *pout = buf[(bo + 1) % 2];
            GatherTask g{};
            g.ep = o.actP; g.en = o.actN; g.ecell = nullptr; g.eval = pin; g.evalStride = o.nact; g.ne = o.nact;
            g.rp = o.actP; g.rn = o.actN; g.nr = o.nact; g.mask = mSelf;
            g.gval = pin; g.gvalStride = o.nact; g.gateMode = GATE_UPPER;
            g.nrhs = 1; g.nsplit = nsSelf[s]; g.emPerSplit = perSelf[s];
            g.partL = b.scratchL + offSelf[s]; g.partU = b.scratchU + offSelf[s];
            st[bo].push_back(g);
            ReduceTask r{};
            r.partL = g.partL; r.partU = g.partU; r.nsplit = g.nsplit + (b.nearMode ? 1 : 0);
            r.nbStride = 1; r.nrhs = 1; r.nr = o.nact;
            if (b.nearMode) {  // the near-field correction is the last split's partial
                NearApplyTask a{};
                a.rowStart = b.nearRowStart[s]; a.cols = b.nearCols[s]; a.vals = b.nearVals[s];
                a.p = o.actP; a.nrm = o.actN; a.pin = pin; a.n = o.nact;
                a.partL = g.partL + (size_t)g.nsplit * o.nact;
                a.partU = g.partU + (size_t)g.nsplit * o.nact;
                na[bo].push_back(a);
            }
            r.gval = pin; r.gvalStride = o.nact; r.gateMode = GATE_UPPER;
            r.out = pout; r.outStride = o.nact; r.pcell = o.P; r.pcellStride = o.ncell; r.rowCell = o.actCell;
            rt[bo].push_back(r);
        }
        const int nwg = (o.nact + 255) / 256;
        for (int w = 0; w < nwg; w++)
            for (int sp2 = 0; sp2 < nsSelf[s]; sp2++) si.push_back(make_int4(task, w, sp2, 0));
        for (long e = 0; e < o.nact; e += 256) ri.push_back(make_int2(task, (int)e));
        // fragment -> field points
        {
            const int ftask = (int)ft.size();
            GatherTask g{};
            g.ep = o.actP; g.en = o.actN; g.ecell = o.actCell; g.eval = o.P; g.evalStride = o.ncell; g.ne = o.nact;
            g.rp = m.fldP; g.rn = nullptr; g.nr = nq; g.mask = mF2P; g.gateMode = GATE_NONE;
            g.nrhs = 1; g.nsplit = nsF2P[s]; g.emPerSplit = perF2P[s];
            g.partL = b.scratchL + offF2P[s]; g.partU = b.scratchU + offF2P[s];
            ft.push_back(g);
            fp.push_back(phys_of(m, 0.f, 1.f));
            const int nwgq = (nq + 255) / 256;
            for (int w = 0; w < nwgq; w++)
                for (int sp2 = 0; sp2 < nsF2P[s]; sp2++) fi.push_back(make_int4(ftask, w, sp2, 0));
            prt.push_back(PointsReduceTask{g.partL, g.partU, g.nsplit, 1, 1, nq, m.fldVal, nq});
            for (long e = 0; e < nq; e += 256) rpi.push_back(make_int2(ftask, (int)e));
        }
    }
    b.selfTasks.resize(b.bounces);
    b.redTasks.resize(b.bounces);
    if (b.visTasks.upload(vt) || b.visBvh.upload(vb) || b.visVc.upload(vv) || b.visItems.upload(vi) ||
        b.trItems.upload(ti) || b.p2fTasks.upload(pt) || b.p2fPhys.upload(pp) || b.p2fItems.upload(pi) ||
        b.zeroTasks.upload(zt) || b.zeroItems.upload(zi) || b.primeTasks.upload(qt) || b.primeItems.upload(qi) ||
        b.selfPhys.upload(sp) || b.selfItems.upload(si) || b.redItems.upload(ri) || b.f2pTasks.upload(ft) ||
        b.f2pPhys.upload(fp) || b.f2pItems.upload(fi) || b.prTasks.upload(prt) || b.prItems.upload(rpi))
        return -1;
    for (int bo = 0; bo < b.bounces; bo++)
        if (b.selfTasks[bo].upload(st[bo]) || b.redTasks[bo].upload(rt[bo])) return -1;
    if (b.nearMode) {
        b.nearApply.resize(b.bounces);
        if (b.nearTasks.upload(nt) || b.nearRowItems.upload(nri) || b.nearMlpItems.upload(nmi) ||
            b.nearQuadItems.upload(nqi) || b.nearApplyItems.upload(nai))
            return -1;
        for (int bo = 0; bo < b.bounces; bo++)
            if (b.nearApply[bo].upload(na[bo])) return -1;
    }
    return 0;
}

// Counts every scene's near pairs on the prepared visibility words (one pass at
// prepare; the counts depend only on geometry and the near settings) and
// allocates each scene's CSR.
int near_prepare(Batch &b) {
    const int S = (int)b.scenes.size();
    std::vector<int *> counts(S, nullptr);
    std::vector<NearBatchTask> nt;
    std::vector<int2> items;
    auto fail = [&](int rc) {
        for (int *p : counts) if (p) (void)hipFree(p);
        return rc;
    };
    for (int s = 0; s < S; s++) {
        Model &m = *b.scenes[s];
        const Object &o = m.targets[0];
        MaskTask tk;
        size_t words;
        if (mask_task(m, 2, 0, 0, 0, 0, tk, words)) return fail(-1);
        tk.mask = b.maskBufs[3 * (size_t)s + 1];
        tk.errFlag = b.counters + 2 + b.bounces;
        if (vis_mask_launch(tk, Bvh{m.trisSorted, m.nodes, m.ntri}, m.vc, b.st)) return fail(-1);
        if (hipMalloc((void **)&counts[s], std::max(1, o.nact) * sizeof(int)) != hipSuccess) {
            set_error("batch near field: out of device memory");
            return fail(-1);
        }
        nt.push_back(near_task(b, s, tk.mask, counts[s]));
        for (int rb = 0; rb < (o.nact + 63) / 64; rb++) items.push_back(make_int2(s, rb));
    }
    DevVec<NearBatchTask> dt;
    DevVec<int2> di;
    if (dt.upload(nt) || di.upload(items)) return fail(-1);
    int rc = near_rows_multi_launch(dt.p, di.p, (int)di.n, b.st);
    if (!rc && hipStreamSynchronize(b.st) != hipSuccess) { set_error("batch near field: count pass failed"); rc = -1; }
    for (int s = 0; !rc && s < S; s++) {
        const int n = b.scenes[s]->targets[0].nact;
        std::vector<int> c(n), start(n + 1, 0);
        if (n && hipMemcpy(c.data(), counts[s], n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
            rc = -1;
            break;
        }
        for (int r = 0; r < n; r++) start[r + 1] = start[r] + c[r];
        const long tot = start[n];
        int *rs = nullptr, *rw = nullptr, *cl = nullptr;
        float2 *vl = nullptr;
        if (hipMalloc((void **)&rs, (n + 1) * sizeof(int)) != hipSuccess ||
            hipMalloc((void **)&rw, std::max(1L, tot) * sizeof(int)) != hipSuccess ||
            hipMalloc((void **)&cl, std::max(1L, tot) * sizeof(int)) != hipSuccess ||
            hipMalloc((void **)&vl, std::max(1L, tot) * sizeof(float2)) != hipSuccess ||
            hipMemcpy(rs, start.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("batch near field: out of device memory");
            rc = -1;
        }
        b.nearRowStart.push_back(rs); b.nearRows.push_back(rw); b.nearCols.push_back(cl); b.nearVals.push_back(vl);
        b.nearTotal.push_back(tot);
    }
    dt.release();
    di.release();
    return fail(rc);
}

// The whole batch as one launch sequence on the batch stream.
int enqueue(Batch &b) {
    hipStream_t s = b.st;
    const int nCounters = 3 + b.bounces;  // vis, f2p, one per bounce, error flag
    if (counters_zero_launch(b.counters, nCounters, s)) return -1;
    if (zero_multi_launch(b.zeroTasks.p, b.zeroItems.p, (int)b.zeroItems.n, s)) return -1;
    if (vis_mask_multi_launch(b.visTasks.p, b.visBvh.p, b.visVc.p, b.visItems.p, (int)b.visItems.n,
                              b.tune.visEmPerWave, b.counters + 0, b.tune.visGrid, s))
        return -1;
    if (mask_transpose_multi_launch(b.visTasks.p, b.trItems.p, (int)b.trItems.n, s)) return -1;
    if (b.nearMode) {  // the near pairs of every scene, in CSR order, and their values
        if (near_rows_multi_launch(b.nearTasks.p, b.nearRowItems.p, (int)b.nearRowItems.n, s) ||
            near_vals_multi_launch(b.nearTasks.p, nearfield().P, b.nearMlpItems.p, (int)b.nearMlpItems.n,
                                   b.nearQuadItems.p, (int)b.nearQuadItems.n, s))
            return -1;
    }
    if (p2f_multi_launch(b.p2fTasks.p, b.p2fPhys.p, b.p2fItems.p, (int)b.p2fItems.n, b.att, s)) return -1;
    if (prime_multi_launch(b.primeTasks.p, b.primeItems.p, (int)b.primeItems.n, s)) return -1;
    for (int bo = 0; bo < b.bounces; bo++) {
        if (gather_multi_launch(b.selfTasks[bo].p, b.selfPhys.p, b.selfItems.p, (int)b.selfItems.n,
                                b.counters + 2 + bo, KIND_F2F, 1, b.att, b.tune.gatherGrid, s))
            return -1;
        if (b.nearMode &&
            near_apply_multi_launch(b.nearApply[bo].p, b.nearApplyItems.p, (int)b.nearApplyItems.n, s))
            return -1;
        if (reduce_multi_launch(b.redTasks[bo].p, b.redItems.p, (int)b.redItems.n, s)) return -1;
    }
    if (gather_multi_launch(b.f2pTasks.p, b.f2pPhys.p, b.f2pItems.p, (int)b.f2pItems.n, b.counters + 1, KIND_F2P,
                            1, b.att, b.tune.gatherGrid, s))
        return -1;
    if (reduce_points_multi_launch(b.prTasks.p, b.prItems.p, (int)b.prItems.n, s)) return -1;
    return 0;
}

int capture(Batch &b) {
    drop_graph(b);
    SR_HIP(hipStreamBeginCapture(b.st, hipStreamCaptureModeGlobal));
    int rc = enqueue(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b.st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    SR_HIP(e);
    b.graph = g;
    SR_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    return 0;
}

}  // namespace

int batch_clear() {
    Batch &b = batch();
    release_all(b);
    b.in.clear();
    return 0;
}

int batch_add_scene(const float *tris, int ntri, float delta, float cp, float f, float atten, const float *src,
                    int nsrc, const float *pts, int npts) {
    Batch &b = batch();
    if (b.prepared) { set_error("batch already prepared; clear it first"); return -3; }
    if (ntri <= 0 || !tris || nsrc <= 0 || !src || npts < 0 || (npts > 0 && !pts) || !(delta > 0.f) || !(cp > 0.f)) {
        set_error("batch_add_scene: bad arguments");
        return -2;
    }
    SceneIn s;
    s.tris.assign(tris, tris + (size_t)9 * ntri);
    s.src.assign(src, src + (size_t)3 * nsrc);
    if (npts) s.pts.assign(pts, pts + (size_t)3 * npts);
    s.delta = delta; s.cp = cp; s.f = f; s.atten = atten;
    b.in.push_back(std::move(s));
    return (int)b.in.size() - 1;
}

int batch_set_near(int mode, float radiusWavelengths, float quadDeltas) {
    Batch &b = batch();
    if (b.prepared) { set_error("batch already prepared"); return -3; }
    if (mode < 0 || mode > 2 || !(radiusWavelengths > 0.f) || !(quadDeltas >= 1.f) || quadDeltas > 16.f) {
        set_error("batch near field: mode 0-2, radius > 0 wavelengths, quadrature radius 1-16 fragment sizes");
        return -2;
    }
    b.nearMode = mode;
    b.nearRadiusWl = radiusWavelengths;
    b.nearQuadDeltas = quadDeltas;
    return 0;
}

int batch_set_bounces(int n) {
    Batch &b = batch();
    if (n < 0) { set_error("bounces must be >= 0"); return -2; }
    if (b.prepared) { set_error("batch already prepared"); return -3; }
    b.bounces = n;
    return 0;
}

int batch_prepare() {
    Batch &b = batch();
    if (b.prepared) return 0;
    if (b.in.empty()) { set_error("empty batch"); return -3; }
    if (!b.st) SR_HIP(hipStreamCreateWithFlags(&b.st, hipStreamNonBlocking));
    for (auto &s : b.in) {
        auto m = std::make_unique<Model>();
        m->stream = b.st;
        if (model_set_initial_conditions(*m, s.cp, s.f, s.atten, 0.f) ||
            model_load_geometry(*m, s.tris.data(), (int)s.tris.size() / 9, OBJ_TARGET, s.delta) ||
            model_load_points(*m, s.src.data(), (int)s.src.size() / 3, true) ||
            model_load_points(*m, s.pts.data(), (int)s.pts.size() / 3, false) || model_render(*m)) {
            b.scenes.push_back(std::move(m));
            release_all(b);
            return -1;
        }
        b.scenes.push_back(std::move(m));
    }
    // a failure from here on releases everything prepared so far
    auto fail = [&]() {
        release_all(b);
        return -1;
    };
    for (auto &mp : b.scenes) {
        for (int kind : {0, 2, 1}) {
            MaskTask tk;
            size_t words;
            if (mask_task(*mp, kind, 0, 0, 0, 0, tk, words)) return fail();
            uint64_t *p = nullptr;
            if (hipMalloc((void **)&p, words * sizeof(uint64_t)) != hipSuccess) {
                set_error("batch_prepare: out of device memory for the visibility words");
                return fail();
            }
            b.maskBufs.push_back(p);
        }
    }
    if (hipMalloc((void **)&b.counters, (3 + b.bounces) * sizeof(int)) != hipSuccess) {
        set_error("batch_prepare: out of device memory");
        return fail();
    }
    if (b.nearMode) {
        if (b.nearMode == 2 && !nearfield().trained) {
            set_error("batch learned near field needs a trained network (sr_nf_train)");
            return fail();
        }
        if (near_prepare(b)) return fail();
    }
    if (build_tables(b)) return fail();
    b.prepared = true;
    return 0;
}

int batch_run(int useGraph, float *ms) {
    Batch &b = batch();
    if (!b.prepared && batch_prepare()) return -1;
    auto t0 = std::chrono::steady_clock::now();
    if (useGraph) {
        if (!b.exec && capture(b)) return -1;
        SR_HIP(hipGraphLaunch(b.exec, b.st));
    } else {
        if (enqueue(b)) return -1;
    }
    SR_HIP(hipStreamSynchronize(b.st));
    if (ms) *ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    int err = 0;
    SR_HIP(hipMemcpy(&err, b.counters + 2 + b.bounces, sizeof(int), hipMemcpyDeviceToHost));
    if (err) { set_error("visibility traversal stack overflow in a batched scene"); return -4; }
    return 0;
}

// Tries each candidate tiling on the prepared batch (graph replays, best of 3)
// and keeps the fastest. report: (ms, visGrid, gatherGrid, visEmPerWave, minPerSplit) per candidate.
int batch_autotune(float *report, int nreport) {
    Batch &b = batch();
    if (!b.prepared && batch_prepare()) return -1;
    std::vector<BatchTune> cands;
    for (int grid : {512, 1024, 2048})
        for (int em : {16, 64})
            for (int mps : {64, 256}) {
                BatchTune t;
                t.visGrid = grid;
                t.gatherGrid = grid;
                t.visEmPerWave = em;
                t.minPerSplit = mps;
                cands.push_back(t);
            }
    float best = 1e30f;
    BatchTune bestT = b.tune;
    int k = 0;
    for (auto &t : cands) {
        b.tune = t;
        if (build_tables(b)) return -1;
        float m = 1e30f, x = 0;
        for (int rep = 0; rep < 3; rep++) {
            if (batch_run(1, &x)) return -1;
            m = std::min(m, x);
        }
        if (report && k < nreport) {
            float *r = report + 5 * k;
            r[0] = m; r[1] = (float)t.visGrid; r[2] = (float)t.gatherGrid; r[3] = (float)t.visEmPerWave;
            r[4] = (float)t.minPerSplit;
        }
        k++;
        if (m < best) { best = m; bestT = t; }
    }
    b.tune = bestT;
    if (build_tables(b)) return -1;
    return k;
}

int batch_field_points(int scene, double *out, int n) {
    Batch &b = batch();
    if (!b.prepared) { set_error("batch not prepared"); return -3; }
    if (scene < 0 || scene >= (int)b.scenes.size()) { set_error("scene index out of range"); return -2; }
    return model_get_field_points(*b.scenes[scene], out, n, 0);
}

double batch_pairs() { return batch().pairs; }

int batch_get_tune(int *o) {
    const BatchTune &t = batch().tune;
    o[0] = t.visGrid; o[1] = t.gatherGrid; o[2] = t.visEmPerWave; o[3] = t.minPerSplit;
    return 0;
}

int batch_set_tune(const int *in) {
    Batch &b = batch();
    if (in[0] <= 0 || in[1] <= 0 || in[2] <= 0 || in[3] <= 0) { set_error("tile parameters must be > 0"); return -2; }
    b.tune = BatchTune{in[0], in[1], in[2], in[3]};
    if (b.prepared) return build_tables(b);
    return 0;
}

}  // namespace sr
