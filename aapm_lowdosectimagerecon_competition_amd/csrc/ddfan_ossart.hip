// ddfan_ossart.hip -- device-resident OS-SART(+FISTA) of OSSART_AAPM.m:49-117:
// the per-reconstruction set-up, the batched iteration (several reconstructions
// of different sizes share every launch) and its hipGraph capture, behind
// ddfan_ossart (one plan) and ddfan_batch_* (many plans).
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <memory>
#include <new>
#include <vector>

#include "ddfan_guard.h"
#include "ddfan_plan.h"

using namespace ddf;

namespace ddf {

int ossart_prepare(OsJob& J, hipStream_t st) {
    ddfan_plan* P = J.P;
    CHECK(P && J.meas && J.x, DDFAN_EINVAL, "NULL argument");
    CHECK(J.numIter >= 0, DDFAN_EINVAL, "numIter must be >= 0");
    CHECK(J.numOS >= 1 && J.numOS <= P->PN, DDFAN_EINVAL, "numOS must be in [1, %d]", P->PN);
    TRY(check_proj_shape(P));
    const GeomDev& g = P->g;
    const int DNU = g.DNU, XN = g.XN, YN = g.YN, SLN = g.SLN, mode = P->mode;
    const int numOS = J.numOS;
    const size_t nvox = (size_t)XN * YN * SLN, npix = (size_t)XN * YN;
    J.cnt.resize(numOS);
    for (int k = 0; k < numOS; ++k) J.cnt[k] = subset_views(P->PN, numOS, k, nullptr);
    J.maxc = J.cnt[0];
    const int maxc = J.maxc;
    // The plan's own stream may still use the buffers reallocated below.
    HIPCALL(hipStreamSynchronize(P->stream));
    HIPCALL(P->rowSum.reserve(sizeof(float) * (size_t)numOS * maxc * DNU));
    HIPCALL(P->colSum.reserve(sizeof(float) * (size_t)numOS * npix));
    HIPCALL(P->rbuf.reserve(sizeof(float) * (size_t)maxc * DNU * SLN));
    HIPCALL(P->satP.reserve(std::max(P->satP.bytes, sat_sino_bytes(mode, P->nzcb, P->ZTb, DNU, maxc))));
    HIPCALL(P->momTab.reserve(sizeof(float) * std::max(J.numIter, 1)));
    HIPCALL(P->counter.reserve(sizeof(int)));
    if (J.fista) HIPCALL(P->xprev.reserve(sizeof(float) * nvox));
    TRY(ensure_vol_sat(P));

    // OS-SART weights from slice 0, replicated to every slice (OSSART_AAPM.m:59-66):
    // rowSum_k = A_k(mask*1), colSum_k = A_k^T(1), both on a one-slice geometry
    // with the slice-0 angles.
    GeomDev g0 = g;
    g0.SLN = 1;
    J.h_cs0.resize((size_t)P->PN * 4);
    for (int v = 0; v < P->PN; ++v) {
        J.h_cs0[2 * v] = P->h_cs[2 * ((size_t)v * SLN)];
        J.h_cs0[2 * v + 1] = P->h_cs[2 * ((size_t)v * SLN) + 1];
    }
    source_table(J.h_cs0.data(), P->PN, P->g.x0, P->g.y0, J.h_cs0.data() + 2 * (size_t)P->PN);
    TRY(upload(P->cs0, J.h_cs0.data(), sizeof(float) * J.h_cs0.size(), st));
    const float2* src0 = P->cs0.as<float2>() + P->PN;
    HIPCALL(P->onesV.reserve(sizeof(float) * npix));
    HIPCALL(P->onesP.reserve(sizeof(float) * (size_t)maxc * DNU));
    HIPCALL(P->satX0.reserve(sat_vol_bytes(mode, 1, 1, XN, YN)));
    HIPCALL(P->satY0.reserve(sat_vol_bytes(mode, 1, 1, YN, XN)));
    HIPCALL(P->satP0.reserve(sat_sino_bytes(mode, 1, 1, DNU, maxc)));
    HIPCALL(launch_fill(P->onesV.as<float>(), 1.0f, npix, st));
    HIPCALL(launch_fill(P->onesP.as<float>(), 1.0f, (size_t)maxc * DNU, st));
    HIPCALL(launch_sat_vol(mode, 1, g0, 1, P->onesV.as<float>(), P->mask.as<uint8_t>(), P->satX0.p,
                           P->satY0.p, st));
    // One table of the ones sinogram serves every subset: a subset of c views
    // reads its first c view rows, which do not depend on the view count.
    HIPCALL(launch_sat_sino(mode, 1, DNU, 1, 1, maxc, P->onesP.as<float>(), P->satP0.p, st));
    // The subsets' weights as batched launches (up to kBatch subsets each):
    // every subset's arithmetic is that of its own launch.
    std::vector<ProjJob> pj;
    std::vector<BackJob> bj;
    pj.reserve(numOS);
    bj.reserve(numOS);
    for (int k = 0; k < numOS; ++k) {
        const Subset sub{k, numOS, J.cnt[k]};
        pj.push_back(ProjJob{g0, sub, 1, 1, 0, PROJ_PLAIN, 1, P->cs0.as<float2>(),
                             P->bx.as<float>(), P->by.as<float>(), P->xds.as<float>(),
                             P->yds.as<float>(), P->satX0.p, P->satY0.p,
                             P->rowSum.as<float>() + (size_t)k * maxc * DNU, nullptr, nullptr, 0});
        bj.push_back(BackJob{g0, sub, 1, 1, BACK_PLAIN, src0,
                             P->mask.as<uint8_t>(), P->satP0.p,
                             P->colSum.as<float>() + (size_t)k * npix, nullptr, nullptr, nullptr,
                             nullptr, 0});
    }
    HIPCALL(launch_proj_batch(mode, pj.data(), numOS, st));
    HIPCALL(launch_back_batch(mode, bj.data(), numOS, st));
    // FISTA momenta (OSSART_AAPM.m:111-115): MATLAB double scalars, applied to
    // single arrays as single((t0-1)/t1).
    J.h_mom.assign(std::max(J.numIter, 1), 0.0f);
    double t0 = 1.0;
    for (int it = 0; it < J.numIter; ++it) {
        const double t1 = (1.0 + sqrt(1.0 + 4.0 * t0 * t0)) / 2.0;
        J.h_mom[it] = (float)((t0 - 1.0) / t1);
        t0 = t1;
    }
    HIPCALL(hipMemcpyAsync(P->momTab.p, J.h_mom.data(), sizeof(float) * J.h_mom.size(),
                           hipMemcpyHostToDevice, st));
    // iteration counter = 0 (the all-zero bit pattern of 0.0f)
    HIPCALL(launch_fill(P->counter.as<float>(), 0.0f, 1, st));
    if (J.fista)
        HIPCALL(hipMemcpyAsync(P->xprev.p, J.x, sizeof(float) * nvox, hipMemcpyDeviceToDevice, st));
    return DDFAN_OK;
}

hipError_t enqueue_batch_iteration(OsJob* const* jobs, int n, int mode, const BatchShape& shape,
                                   hipStream_t st, ddfan_plan* TP) {
    int S = 0;
    for (int i = 0; i < n; ++i) S = std::max(S, jobs[i]->numOS);
    std::vector<SatVolJob> sv;
    std::vector<ProjJob> pj;
    std::vector<SatSinoJob> ss;
    std::vector<BackJob> bj;
    std::vector<size_t> tb;  // volume-table bytes of each job of the round
    tb.reserve(n);
    sv.reserve(n);
    pj.reserve(n);
    ss.reserve(n);
    bj.reserve(n);
    auto T = [&](int kind, auto&& f) -> hipError_t { return TP ? TP->timed(kind, f, st) : f(); };
    for (int r = 0; r < S; ++r) {
        sv.clear();
        pj.clear();
        ss.clear();
        bj.clear();
        tb.clear();
        for (int i = 0; i < n; ++i) {
            const OsJob& J = *jobs[i];
            const long long nO = J.numOS;
            // subset k runs in round floor(k*S/nO); k*S/nO is increasing since S >= nO
            const long long k = ((long long)r * nO + S - 1) / S;
            if (k >= nO || (k * S) / nO != r) continue;
            ddfan_plan* P = J.P;
            const GeomDev& g = P->g;
            const size_t npix = (size_t)g.XN * g.YN;
            const Subset sub{(int)k, (int)nO, J.cnt[k]};
            const int epi = (J.fista && k == nO - 1) ? BACK_UPDATE_FISTA : BACK_UPDATE;
            sv.push_back(SatVolJob{g, P->nzc, P->ZT, J.x, P->mask.as<uint8_t>(), P->satX.p, P->satY.p});
            tb.push_back(P->satX.bytes + P->satY.bytes);
            pj.push_back(ProjJob{g, sub, P->nzc, P->ZT, P->slabTile, PROJ_RATIO, P->viewsPerWave,
                                 P->cs.as<float2>(),
                                 P->bx.as<float>(), P->by.as<float>(), P->xds.as<float>(),
                                 P->yds.as<float>(), P->satX.p, P->satY.p, P->rbuf.as<float>(),
                                 J.meas, P->rowSum.as<float>() + (size_t)k * J.maxc * g.DNU,
                                 P->projStage, P->h_cs.data()});
            ss.push_back(SatSinoJob{g.DNU, g.SLN, P->nzcb, P->ZTb, J.cnt[k], P->rbuf.as<float>(),
                                    P->satP.p});
            bj.push_back(BackJob{g, sub, P->nzcb, P->ZTb, epi, P->src.as<float2>(),
                                 P->mask.as<uint8_t>(), P->satP.p, J.x,
                                 P->colSum.as<float>() + (size_t)k * npix, P->xprev.as<float>(),
                                 P->momTab.as<float>(), P->counter.as<int>(), P->backStage});
        }
        const int m = (int)sv.size();
        if (m == 0) continue;
        // A job with large tables runs its four kernels back to back, so the
        // tables it has just built are read before other jobs' tables evict
        // them from the caches; the rest share batched launches per kernel kind.
        int ms = 0;  // shared jobs compacted to the front
        for (int i = 0; i < m; ++i) {
            if ((long long)tb[i] < shape.soloBytes) {
                if (ms != i) {
                    std::swap(sv[ms], sv[i]);
                    std::swap(pj[ms], pj[i]);
                    std::swap(ss[ms], ss[i]);
                    std::swap(bj[ms], bj[i]);
                    std::swap(tb[ms], tb[i]);
                }
                ++ms;
            }
        }
        for (int i = ms; i < m; ++i) {
            hipError_t e = T(0, [&] { return launch_sat_vol_batch(mode, &sv[i], 1, st); });
            if (e != hipSuccess) return e;
            e = T(1, [&] { return launch_proj_batch(mode, &pj[i], 1, st, shape.projFill); });
            if (e != hipSuccess) return e;
            e = T(2, [&] { return launch_sat_sino_batch(mode, &ss[i], 1, st); });
            if (e != hipSuccess) return e;
            e = T(3, [&] { return launch_back_batch(mode, &bj[i], 1, st, shape.backFill); });
            if (e != hipSuccess) return e;
        }
        if (ms == 0) continue;
        hipError_t e = T(0, [&] { return launch_sat_vol_batch(mode, sv.data(), ms, st); });
        if (e != hipSuccess) return e;
        e = T(1, [&] {
            return launch_proj_batch(mode, pj.data(), ms, st, shape.projFill);
        });
        if (e != hipSuccess) return e;
        e = T(2, [&] { return launch_sat_sino_batch(mode, ss.data(), ms, st); });
        if (e != hipSuccess) return e;
        e = T(3, [&] { return launch_back_batch(mode, bj.data(), ms, st, shape.backFill); });
        if (e != hipSuccess) return e;
    }
    std::vector<int*> ctr(n);
    for (int i = 0; i < n; ++i) ctr[i] = jobs[i]->P->counter.as<int>();
    return launch_counter_inc_batch(ctr.data(), n, st);
}

int IterGraph::capture(OsJob* const* jobs, int n, int mode, const BatchShape& shape,
                       hipStream_t st) {
    reset();
    HIPCALL(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const hipError_t ce = enqueue_batch_iteration(jobs, n, mode, shape, st, nullptr);
    const hipError_t ee = hipStreamEndCapture(st, &graph);
    if (ce != hipSuccess) {
        reset();
        HIPCALL(ce);
    }
    HIPCALL(ee);
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        reset();
        HIPCALL(ie);
    }
    return DDFAN_OK;
}

int IterGraph::launch(hipStream_t st) {
    CHECK(exec, DDFAN_EINVAL, "graph not captured");
    HIPCALL(hipGraphLaunch(exec, st));
    return DDFAN_OK;
}

}  // namespace ddf

namespace {

// Waits for a stream when leaving scope: host staging and graphs that work in
// flight still reads are destroyed only after it.
struct SyncOnExit {
    hipStream_t s;
    ~SyncOnExit() {
        if (s) (void)hipStreamSynchronize(s);
    }
};

// Runs `iters` iterations of every job on `st`: eagerly, or as replays of one
// captured iteration; returns after the stream has drained.
int run_iterations(OsJob* const* jobs, int n, int mode, const BatchShape& shape, int iters,
                   bool graph, hipStream_t st, ddfan_plan* timingPlan) {
    if (iters <= 0 || n == 0) return DDFAN_OK;
    if (graph) {
        IterGraph G;
        SyncOnExit sync{st};  // before G is destroyed
        TRY(G.capture(jobs, n, mode, shape, st));
        for (int it = 0; it < iters; ++it) TRY(G.launch(st));
        HIPCALL(hipStreamSynchronize(st));
        return DDFAN_OK;
    }
    for (int it = 0; it < iters; ++it)
        HIPCALL(enqueue_batch_iteration(jobs, n, mode, shape, st, timingPlan));
    HIPCALL(hipStreamSynchronize(st));
    return DDFAN_OK;
}

}  // namespace

struct ddfan_batch {
    int device = 0;
    int mode = 0;
    int nstreams = 0;  // DDFAN_BATCH_STREAMS: groups run side by side (0: automatic)
    BatchShape shape;
    hipStream_t stream = nullptr;
    std::vector<hipStream_t> extra;  // streams of groups 1.. (group 0 uses `stream`)
    std::vector<OsJob> jobs;
    ~ddfan_batch() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        for (hipStream_t s : extra) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
};

namespace {

// Groups of an automatic split.  Measured on MI355X, 10 iterations of 48-view
// subsets: 16 jobs of 256^2 x 8-16 slices, 720 views, with throughput-tuned
// shared launches: one group 0.224 s, two 0.185 s, four 0.193-0.201 s (one
// after another 0.404 s); the uneven mix of 256^2-768^2 x 8-123 slices, whose
// jobs run their kernels back to back: one group 6.13 s, two 5.88-5.94 s,
// four 6.03-6.10 s (one after another 6.10-6.12 s): one group's launch tails
// and set-up overlap the other group's kernels.
constexpr int kAutoStreams = 2;

// Runs every job of the batch in `ng` groups, one stream each, side by side:
// jobs go to the group with the least work so far (largest first; work =
// voxel-views x iterations), each group replays its own batch iteration
// (captured again whenever a job of the group has finished) on its stream,
// and the host launches the groups' iterations in turn so all streams stay
// busy.  A job's launches are those of its group's batch iteration, so its
// arithmetic is that of running alone.
int run_groups(ddfan_batch* B, int ng, bool graph) {
    const int n = (int)B->jobs.size();
    std::vector<hipStream_t> st(ng);
    st[0] = B->stream;
    for (int k = 1; k < ng; ++k) st[k] = B->extra[k - 1];
    // everything after the set-up on B->stream
    hipEvent_t ready = nullptr;
    HIPCALL(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    struct EvGuard {
        hipEvent_t e;
        ~EvGuard() { (void)hipEventDestroy(e); }
    } evg{ready};
    HIPCALL(hipEventRecord(ready, B->stream));
    for (int k = 1; k < ng; ++k) HIPCALL(hipStreamWaitEvent(st[k], ready, 0));
    std::vector<int> order(n), group(n);
    std::vector<double> work(n), load(ng, 0.0);
    for (int i = 0; i < n; ++i) {
        const GeomDev& g = B->jobs[i].P->g;
        work[i] = (double)g.XN * g.YN * g.SLN * B->jobs[i].P->PN * std::max(B->jobs[i].numIter, 0);
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
    for (int i : order) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        group[i] = k;
        load[k] += work[i];
    }
    std::vector<int> left(n);
    for (int i = 0; i < n; ++i) left[i] = B->jobs[i].numIter;
    // graphs are destroyed after every stream has drained (the launches in
    // flight read their executables)
    std::vector<std::unique_ptr<IterGraph>> graphs;
    struct Drain {
        std::vector<hipStream_t>& s;
        ~Drain() {
            for (hipStream_t x : s) (void)hipStreamSynchronize(x);
        }
    } drain{st};
    std::vector<IterGraph*> cur(ng, nullptr);
    std::vector<std::vector<OsJob*>> members(ng);
    for (;;) {
        bool any = false;
        for (int k = 0; k < ng; ++k) {
            std::vector<OsJob*> act;
            for (int i = 0; i < n; ++i)
                if (group[i] == k && left[i] > 0) act.push_back(&B->jobs[i]);
            if (act.empty()) continue;
            any = true;
            if (!graph) {
                HIPCALL(enqueue_batch_iteration(act.data(), (int)act.size(), B->mode, B->shape,
                                                st[k], nullptr));
            } else {
                if (act != members[k]) {
                    graphs.push_back(std::make_unique<IterGraph>());
                    TRY(graphs.back()->capture(act.data(), (int)act.size(), B->mode, B->shape, st[k]));
                    cur[k] = graphs.back().get();
                    members[k] = act;
                }
                TRY(cur[k]->launch(st[k]));
            }
            for (int i = 0; i < n; ++i)
                if (group[i] == k && left[i] > 0) --left[i];
        }
        if (!any) break;
    }
    for (int k = 0; k < ng; ++k) HIPCALL(hipStreamSynchronize(st[k]));
    return DDFAN_OK;
}

// Median of `reps` timed runs of f on `st` (after one untimed run), in ms.
template <class F>
int median_ms(hipStream_t st, int reps, F&& f, double* ms) {
    hipEvent_t a = nullptr, b = nullptr;
    HIPCALL(hipEventCreate(&a));
    struct EvPair {
        hipEvent_t& a;
        hipEvent_t& b;
        ~EvPair() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } guard{a, b};
    HIPCALL(hipEventCreate(&b));
    HIPCALL(f());
    std::vector<double> t;
    for (int r = 0; r < reps; ++r) {
        HIPCALL(hipEventRecord(a, st));
        HIPCALL(f());
        HIPCALL(hipEventRecord(b, st));
        HIPCALL(hipEventSynchronize(b));
        float x = 0.0f;
        HIPCALL(hipEventElapsedTime(&x, a, b));
        t.push_back(x);
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return DDFAN_OK;
}

// Launch parameters of one plan (restored if batch tuning fails part-way).
struct Shape {
    int ZT, V, pstage, ZTb, bstage;
};
Shape shape_of(const ddfan_plan* P) {
    return Shape{P->ZT, P->viewsPerWave, P->projStage, P->ZTb, P->backStage};
}
int set_shape(ddfan_plan* P, const Shape& s) {
    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_PROJ_CHUNK, s.ZT));
    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_VIEWS_PER_WAVE, s.V));
    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_PROJ_STAGE, s.pstage));
    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_CHUNK, s.ZTb));
    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_STAGE, s.bstage));
    return DDFAN_OK;
}

// Throughput shapes of the shared launches.  A plan's own autotuning times
// its launches alone, where a small reconstruction leaves most of the GPU idle
// and the shape with the most waves wins even if each wave does little; in a
// batch the shared launches fill the GPU together, so work per wave decides.
// This times the shared reconstructions' launches together (volume tables +
// projector, sinogram table + backprojector, on two of each one's subsets)
// for one projector shape (slice chunk, views or tasks per wave, staging) and
// one backprojector shape (slice chunk, staged views) applied to all of them
// (a plan whose slices allow a narrower chunk only takes that one),
// keeps the fastest, and applies it when it beats the shapes in place.
// Outputs go to scratch buffers; the iterates are only read.  Slab tiles,
// which group the projector's sums, are left as they are, so results do not
// change.
int batch_tune(ddfan_batch* B, int reps, double* out_ms) {
    std::vector<OsJob*> sh;
    for (OsJob& J : B->jobs)
        if ((long long)(J.P->satX.bytes + J.P->satY.bytes) < B->shape.soloBytes) sh.push_back(&J);
    const int m = (int)sh.size();
    if (out_ms)
        for (int i = 0; i < 4; ++i) out_ms[i] = 0.0;
    if (m == 0) return DDFAN_OK;
    const int mode = B->mode;
    hipStream_t st = B->stream;
    std::vector<Shape> saved(m);
    std::vector<std::vector<int>> ks(m), cnt(m);  // timed subsets and their view counts
    std::vector<DBuf> sino(m), vol(m);
    int ZTcap = 1;  // widest chunk any shared plan allows (each plan takes min(chunk, its own))
    for (int i = 0; i < m; ++i) {
        ddfan_plan* P = sh[i]->P;
        saved[i] = shape_of(P);
        ZTcap = std::max(ZTcap, std::min(P->ZTmax, 16));
        const int nO = sh[i]->numOS;
        ks[i] = nO > 1 ? std::vector<int>{0, nO / 2} : std::vector<int>{0};
        int maxc = 0;
        for (int k : ks[i]) {
            cnt[i].push_back(subset_views(P->PN, nO, k, nullptr));
            maxc = std::max(maxc, cnt[i].back());
        }
        const GeomDev& g = P->g;
        HIPCALL(sino[i].reserve(sizeof(float) * (size_t)maxc * g.DNU * g.SLN));
        HIPCALL(vol[i].reserve(sizeof(float) * (size_t)g.XN * g.YN * g.SLN));
        HIPCALL(launch_fill(sino[i].as<float>(), 0.0f, (size_t)maxc * g.DNU * g.SLN, st));
    }
    int nsub = 0;
    for (int i = 0; i < m; ++i) nsub = std::max(nsub, (int)ks[i].size());
    auto proj = [&]() -> hipError_t {
        std::vector<SatVolJob> sv;
        std::vector<ProjJob> pj;
        for (int s = 0; s < nsub; ++s) {
            sv.clear();
            pj.clear();
            for (int i = 0; i < m; ++i) {
                if (s >= (int)ks[i].size()) continue;
                ddfan_plan* P = sh[i]->P;
                const Subset sub{ks[i][s], sh[i]->numOS, cnt[i][s]};
                sv.push_back(SatVolJob{P->g, P->nzc, P->ZT, sh[i]->x, P->mask.as<uint8_t>(),
                                       P->satX.p, P->satY.p});
                pj.push_back(ProjJob{P->g, sub, P->nzc, P->ZT, P->slabTile, PROJ_PLAIN,
                                     P->viewsPerWave, P->cs.as<float2>(), P->bx.as<float>(),
                                     P->by.as<float>(), P->xds.as<float>(), P->yds.as<float>(),
                                     P->satX.p, P->satY.p, sino[i].as<float>(), nullptr, nullptr,
                                     P->projStage, P->h_cs.data()});
            }
            hipError_t e = launch_sat_vol_batch(mode, sv.data(), (int)sv.size(), st);
            if (e == hipSuccess)
                e = launch_proj_batch(mode, pj.data(), (int)pj.size(), st, B->shape.projFill);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    auto back = [&]() -> hipError_t {
        std::vector<SatSinoJob> ss;
        std::vector<BackJob> bj;
        for (int s = 0; s < nsub; ++s) {
            ss.clear();
            bj.clear();
            for (int i = 0; i < m; ++i) {
                if (s >= (int)ks[i].size()) continue;
                ddfan_plan* P = sh[i]->P;
                const Subset sub{ks[i][s], sh[i]->numOS, cnt[i][s]};
                ss.push_back(SatSinoJob{P->g.DNU, P->g.SLN, P->nzcb, P->ZTb, cnt[i][s],
                                        sino[i].as<float>(), P->satP.p});
                bj.push_back(BackJob{P->g, sub, P->nzcb, P->ZTb, BACK_PLAIN, P->src.as<float2>(),
                                     P->mask.as<uint8_t>(), P->satP.p, vol[i].as<float>(), nullptr,
                                     nullptr, nullptr, nullptr, P->backStage});
            }
            hipError_t e = launch_sat_sino_batch(mode, ss.data(), (int)ss.size(), st);
            if (e == hipSuccess)
                e = launch_back_batch(mode, bj.data(), (int)bj.size(), st, B->shape.backFill);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    // Tables sized for the current chunks of every shared plan.
    auto fit_tables = [&]() -> int {
        for (int i = 0; i < m; ++i) {
            ddfan_plan* P = sh[i]->P;
            TRY(ensure_vol_sat(P));
            int maxc = 0;
            for (int c : cnt[i]) maxc = std::max(maxc, c);
            const size_t b = sat_sino_bytes(mode, P->nzcb, P->ZTb, P->g.DNU, maxc);
            if (b > P->satP.bytes) {
                HIPCALL(hipStreamSynchronize(st));
                HIPCALL(P->satP.reserve(b));
            }
        }
        return DDFAN_OK;
    };
    int rc = [&]() -> int {
        TRY(fit_tables());
        double p0, b0;
        TRY(median_ms(st, reps, proj, &p0));
        TRY(median_ms(st, reps, back, &b0));
        double bestP = p0, bestB = b0;
        Shape best = shape_of(sh[0]->P);
        bool projUniform = false, backUniform = false;
        std::vector<int> chunks;
        for (int z = ZTcap; z >= 1; z >>= 1) chunks.push_back(z);
        std::vector<std::pair<int, int>> shapes;  // (views or tasks per wave, stage)
        for (int v : {4, 2, 1})
            for (int ps : {1, 0}) shapes.push_back({v, ps});
        for (int ps : {2, 3})
            for (int v : {12, 9, 6, 5, 4, 3}) shapes.push_back({v, ps});
        // a candidate chunk z gives each plan min(z, its widest chunk)
        auto zc = [&](int i, int z) { return std::min(z, sh[i]->P->ZTmax); };
        for (int z : chunks)
            for (const auto& s : shapes) {
                bool ok = true;
                for (int i = 0; i < m && ok; ++i)
                    ok = proj_shape_ok(mode, sh[i]->P->g, zc(i, z), s.first, s.second);
                if (!ok) continue;
                for (int i = 0; i < m; ++i) {
                    ddfan_plan* P = sh[i]->P;
                    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_PROJ_CHUNK, zc(i, z)));
                    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_VIEWS_PER_WAVE, s.first));
                    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_PROJ_STAGE, s.second));
                }
                TRY(fit_tables());
                double ms;
                TRY(median_ms(st, reps, proj, &ms));
                if (ms < bestP) {
                    bestP = ms;
                    best.ZT = z;  // each plan: zc(i, best.ZT)
                    best.V = s.first;
                    best.pstage = s.second;
                    projUniform = true;
                }
            }
        for (int i = 0; i < m; ++i) {  // projector: the winner, or each plan's own shape
            ddfan_plan* P = sh[i]->P;
            const Shape& s = projUniform ? best : saved[i];
            TRY(ddfan_plan_set_param(P, DDFAN_PARAM_PROJ_CHUNK, projUniform ? zc(i, s.ZT) : s.ZT));
            TRY(ddfan_plan_set_param(P, DDFAN_PARAM_VIEWS_PER_WAVE, s.V));
            TRY(ddfan_plan_set_param(P, DDFAN_PARAM_PROJ_STAGE, s.pstage));
        }
        TRY(fit_tables());
        HIPCALL(proj());  // a projection of each iterate into the sinograms the backprojector reads
        for (int z : chunks)
            for (int bs : {1, 2, 4, 0}) {
                if (bs * z > 32) continue;  // also bs * zc(i, z) <= 32 for every plan
                for (int i = 0; i < m; ++i) {
                    ddfan_plan* P = sh[i]->P;
                    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_CHUNK, zc(i, z)));
                    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_STAGE, bs));
                }
                TRY(fit_tables());
                double ms;
                TRY(median_ms(st, reps, back, &ms));
                if (ms < bestB) {
                    bestB = ms;
                    best.ZTb = z;
                    best.bstage = bs;
                    backUniform = true;
                }
            }
        for (int i = 0; i < m; ++i) {
            ddfan_plan* P = sh[i]->P;
            const Shape& s = backUniform ? best : saved[i];
            TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_CHUNK, backUniform ? zc(i, s.ZTb) : s.ZTb));
            TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_STAGE, s.bstage));
        }
        TRY(fit_tables());
        HIPCALL(hipStreamSynchronize(st));
        if (out_ms) {
            out_ms[0] = p0 / nsub;
            out_ms[1] = b0 / nsub;
            out_ms[2] = bestP / nsub;
            out_ms[3] = bestB / nsub;
        }
        return DDFAN_OK;
    }();
    (void)hipStreamSynchronize(st);  // before the scratch buffers are freed
    if (rc != DDFAN_OK)
        for (int i = 0; i < m; ++i) (void)set_shape(sh[i]->P, saved[i]);
    return rc;
}

}  // namespace

extern "C" {

int ddfan_batch_tune(ddfan_batch* B, int reps, double* out_ms) {
    return ddf::guarded([&]() -> int {
        CHECK(B, DDFAN_EINVAL, "batch is NULL");
        HIPCALL(hipSetDevice(B->device));
        reps = std::max(1, std::min(reps, 20));
        return batch_tune(B, reps, out_ms);
    });
}

int ddfan_ossart(ddfan_plan* P, const float* d_meas, float* d_x, int numIter, int numOS,
                 int useFISTA, int flags) {
    return ddf::guarded([&]() -> int {
        CHECK(P && d_meas && d_x, DDFAN_EINVAL, "NULL argument");
        HIPCALL(hipSetDevice(P->device));
        OsJob J;
        J.P = P;
        J.meas = d_meas;
        J.x = d_x;
        J.numIter = numIter;
        J.numOS = numOS;
        J.fista = useFISTA ? 1 : 0;
        SyncOnExit sync{P->stream};  // J's staging outlives the uploads
        TRY(ossart_prepare(J, P->stream));
        OsJob* jp = &J;
        const bool graph = (flags & DDFAN_OSSART_GRAPH) != 0;
        return run_iterations(&jp, 1, P->mode, BatchShape{}, numIter, graph, P->stream,
                              (P->timing && !graph) ? P : nullptr);
    });
}

int ddfan_batch_create(int sat_mode, ddfan_batch** out) {
    return ddf::guarded([&]() -> int {
        CHECK(out, DDFAN_EINVAL, "out is NULL");
        *out = nullptr;
        CHECK(sat_mode >= 0 && sat_mode <= 2, DDFAN_EINVAL, "sat_mode must be 0, 1 or 2");
        int dev;
        TRY(check_device(&dev));
        ddfan_batch* B = new (std::nothrow) ddfan_batch();
        CHECK(B, DDFAN_ENOMEM, "host allocation failed");
        B->device = dev;
        B->mode = sat_mode;
        const hipError_t e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete B;
            HIPCALL(e);
        }
        *out = B;
        return DDFAN_OK;
    });
}

int ddfan_batch_add(ddfan_batch* B, ddfan_plan* P, const float* d_meas, float* d_x, int numIter,
                    int numOS, int useFISTA, int* index) {
    return ddf::guarded([&]() -> int {
        CHECK(B && P && d_meas && d_x, DDFAN_EINVAL, "NULL argument");
        CHECK(P->device == B->device, DDFAN_EINVAL, "plan and batch are on different devices");
        CHECK(P->mode == B->mode, DDFAN_EINVAL, "plan SAT mode %d differs from the batch's %d",
              P->mode, B->mode);
        CHECK(numIter >= 0, DDFAN_EINVAL, "numIter must be >= 0");
        CHECK(numOS >= 1 && numOS <= P->PN, DDFAN_EINVAL, "numOS must be in [1, %d]", P->PN);
        for (const OsJob& J : B->jobs)
            CHECK(J.P != P, DDFAN_EINVAL, "a plan holds the state of one reconstruction: add it once");
        OsJob J;
        J.P = P;
        J.meas = d_meas;
        J.x = d_x;
        J.numIter = numIter;
        J.numOS = numOS;
        J.fista = useFISTA ? 1 : 0;
        B->jobs.push_back(std::move(J));
        if (index) *index = (int)B->jobs.size() - 1;
        return DDFAN_OK;
    });
}

int ddfan_batch_set_param(ddfan_batch* B, int param, long long value) {
    return ddf::guarded([&]() -> int {
        CHECK(B, DDFAN_EINVAL, "batch is NULL");
        switch (param) {
            case DDFAN_BATCH_PROJ_FILL:
                CHECK(value >= 1, DDFAN_EINVAL, "projector launch limit must be >= 1");
                B->shape.projFill = value;
                return DDFAN_OK;
            case DDFAN_BATCH_BACK_FILL:
                CHECK(value >= 1, DDFAN_EINVAL, "backprojector launch limit must be >= 1");
                B->shape.backFill = value;
                return DDFAN_OK;
            case DDFAN_BATCH_SOLO_BYTES:
                CHECK(value >= 0, DDFAN_EINVAL, "table-size threshold must be >= 0");
                B->shape.soloBytes = value;
                return DDFAN_OK;
            case DDFAN_BATCH_STREAMS:
                CHECK(value >= 0 && value <= 8, DDFAN_EINVAL, "streams must be in [0, 8]");
                B->nstreams = (int)value;
                return DDFAN_OK;
            default:
                return fail(DDFAN_EINVAL, "unknown batch parameter %d", param);
        }
    });
}

int ddfan_batch_size(const ddfan_batch* B, int* n) {
    return ddf::guarded([&]() -> int {
        CHECK(B && n, DDFAN_EINVAL, "NULL argument");
        *n = (int)B->jobs.size();
        return DDFAN_OK;
    });
}

int ddfan_batch_run(ddfan_batch* B, int flags) {
    return ddf::guarded([&]() -> int {
        CHECK(B, DDFAN_EINVAL, "batch is NULL");
        HIPCALL(hipSetDevice(B->device));
        const int n = (int)B->jobs.size();
        if (n == 0) return DDFAN_OK;
        SyncOnExit sync{B->stream};
        for (OsJob& J : B->jobs) TRY(ossart_prepare(J, B->stream));
        int ng = B->nstreams;
        if (ng == 0) ng = kAutoStreams;
        ng = std::min(ng, n);
        while ((int)B->extra.size() < ng - 1) {
            hipStream_t s = nullptr;
            HIPCALL(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            B->extra.push_back(s);
        }
        if (ng > 1) return run_groups(B, ng, (flags & DDFAN_OSSART_GRAPH) != 0);
        // Segments of constant membership: a job leaves after its last iteration.
        std::vector<int> left(n);
        for (int i = 0; i < n; ++i) left[i] = B->jobs[i].numIter;
        const bool graph = (flags & DDFAN_OSSART_GRAPH) != 0;
        for (;;) {
            std::vector<OsJob*> act;
            int seg = 0;
            for (int i = 0; i < n; ++i)
                if (left[i] > 0) {
                    act.push_back(&B->jobs[i]);
                    seg = seg == 0 ? left[i] : std::min(seg, left[i]);
                }
            if (act.empty()) break;
            TRY(run_iterations(act.data(), (int)act.size(), B->mode, B->shape, seg, graph, B->stream,
                               nullptr));
            for (int i = 0; i < n; ++i)
                if (left[i] > 0) left[i] -= seg;
        }
        return DDFAN_OK;
    });
}

int ddfan_batch_destroy(ddfan_batch* B) {
    return ddf::guarded([&]() -> int {
        if (!B) return DDFAN_OK;
        (void)hipSetDevice(B->device);
        delete B;
        return DDFAN_OK;
    });
}

}  // extern "C"
