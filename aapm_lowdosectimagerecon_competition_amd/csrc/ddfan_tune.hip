// ddfan_tune.hip -- launch-shape autotuner.
//
// For one plan shape and one launch size (views per operator call, spaced
// PN / views apart like an OS-SART subset) it times,
// with HIP events on the plan's stream, the projector over slice chunks x
// views per wave x LDS staging (one slab pass), then slab tilings of the
// winner, and the backprojector over slice chunks x LDS-staged view groups,
// and keeps the fastest.  Results live in
// a process-wide table keyed by (SAT mode, XN, YN, SLN, DNU, widest uniform
// chunk, views), which can be saved to and loaded from a text file and is
// consulted by ddfan_plan_apply_tuned (the request loop applies it to every
// admitted reconstruction).  No parameter changes results: slice chunks,
// views per wave and staging give bit-identical output, slab passes change
// only the summation grouping of the projector.
#include <hip/hip_runtime.h>
#include <stdio.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "ddfan_guard.h"
#include "ddfan_plan.h"

using namespace ddf;

namespace {

struct TuneKey {
    int mode, XN, YN, SLN, DNU, ZTmax, views;
    bool operator<(const TuneKey& o) const {
        return std::tie(mode, XN, YN, SLN, DNU, ZTmax, views) <
               std::tie(o.mode, o.XN, o.YN, o.SLN, o.DNU, o.ZTmax, o.views);
    }
};
struct TuneVal {
    int ZT, V, tile, ZTb, stage, pstage;
    double projMs, backMs;
    int bpass = -1;  // backprojector views per pass (-1: not recorded, left as it is)
};

bool stage_fits(int stage, int ZTb) {
    return (stage == 0 || stage == 1 || stage == 2 || stage == 4) && stage * ZTb <= 32;
}

std::mutex g_tuneMu;
std::map<TuneKey, TuneVal> g_tune;

TuneKey key_of(const ddfan_plan* P, int views) {
    return TuneKey{P->mode, P->g.XN, P->g.YN, P->g.SLN, P->g.DNU, P->ZTmax, views};
}

int apply(ddfan_plan* P, const TuneVal& v) {
    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_PROJ_CHUNK, v.ZT));
    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_VIEWS_PER_WAVE, v.V));
    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_SLAB_TILE, v.tile));
    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_CHUNK, v.ZTb));
    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_STAGE, v.stage));
    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_PROJ_STAGE, v.pstage));
    if (v.bpass >= 0) TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_PASS, v.bpass));
    return DDFAN_OK;
}

// Median of `reps` timed calls of f (after one untimed call), in ms.
template <class F>
int time_ms(hipStream_t st, int reps, F&& f, double* ms) {
    hipEvent_t a, b;
    HIPCALL(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return fail(DDFAN_EDEVICE, "hipEventCreate failed");
    }
    int rc = [&]() -> int {
        TRY(f());
        std::vector<double> t;
        for (int r = 0; r < reps; ++r) {
            HIPCALL(hipEventRecord(a, st));
            TRY(f());
            HIPCALL(hipEventRecord(b, st));
            HIPCALL(hipEventSynchronize(b));
            float x = 0.0f;
            HIPCALL(hipEventElapsedTime(&x, a, b));
            t.push_back(x);
        }
        std::sort(t.begin(), t.end());
        *ms = t[t.size() / 2];
        return DDFAN_OK;
    }();
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

}  // namespace

extern "C" {

int ddfan_plan_autotune(ddfan_plan* P, int views, int reps, int flags, int* out_params,
                        double* out_ms) {
    return ddf::guarded([&]() -> int {
        CHECK(P, DDFAN_EINVAL, "plan is NULL");
        if (views <= 0) views = P->PN;
        CHECK(views <= P->PN, DDFAN_EINVAL, "views %d exceeds the plan's %d", views, P->PN);
        reps = std::max(1, std::min(reps, 20));
        const TuneKey key = key_of(P, views);
        TuneVal best{};
        bool have = false;
        if (flags & DDFAN_TUNE_USE_CACHE) {
            std::lock_guard<std::mutex> lk(g_tuneMu);
            auto it = g_tune.find(key);
            if (it != g_tune.end()) {
                best = it->second;
                have = true;
            }
        }
        if (!have) {
            HIPCALL(hipSetDevice(P->device));
            const GeomDev& g = P->g;
            const size_t nvox = (size_t)g.XN * g.YN * g.SLN;
            const size_t nprj = (size_t)views * g.DNU * g.SLN;
            DBuf vol, prj;
            HIPCALL(vol.reserve(nvox * sizeof(float)));
            HIPCALL(prj.reserve(nprj * sizeof(float)));
            HIPCALL(launch_fill(vol.as<float>(), 1.0f, nvox, P->stream));
            HIPCALL(launch_fill(prj.as<float>(), 1.0f, nprj, P->stream));
            // the plan's own settings are restored if tuning fails part-way
            const TuneVal saved{P->ZT,        P->viewsPerWave, P->slabTile, P->ZTb,
                                P->backStage, P->projStage,    0.0,         0.0, P->backPass};
            int rc = [&]() -> int {
                std::vector<int> chunks;
                for (int z = std::min(P->ZTmax, 16); z >= 1; z >>= 1) chunks.push_back(z);
                // launches of `views` views spaced PN / views apart: an OS-SART
                // subset (consecutive views when views = PN); the spacing decides
                // how much of a table row neighbouring views of a launch share
                const int stride = std::max(1, P->PN / views);
                // spaced launches: three subsets (first views 0, stride/3, 2 stride/3),
                // since where a subset's views fall against the slab-orientation
                // changes at 45 degrees decides the whole-row projector's blocks
                const int nsub = stride >= 3 ? 3 : 1;
                auto proj = [&]() {
                    for (int k = 0; k < nsub; ++k)
                        TRY(ddfan_plan_proj(P, vol.as<float>(), prj.as<float>(), k * stride / 3,
                                            stride, views));
                    return (int)DDFAN_OK;
                };
                auto back = [&]() {
                    return ddfan_plan_back(P, prj.as<float>(), vol.as<float>(), 0, stride, views);
                };
                double bestP = 1e30;
                TRY(ddfan_plan_set_param(P, DDFAN_PARAM_SLAB_TILE, 0));
                // the whole-row projector (stages 2, 3) only for spaced views, whose
                // neighbours share no stretch of a row
                std::vector<std::pair<int, int>> shapes;  // (views or tasks per wave, stage)
                for (int v : {4, 2, 1})
                    for (int ps : {1, 0}) shapes.push_back({v, ps});
                if (stride > 1)
                    for (int ps : {2, 3})
                        for (int v : {12, 9, 6, 5, 4, 3}) shapes.push_back({v, ps});
                for (int z : chunks) {
                    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_PROJ_CHUNK, z));
                    for (const auto& sh : shapes) {
                        const int v = sh.first, ps = sh.second;
                        if (!proj_shape_ok(P->mode, g, z, v, ps)) continue;
                        TRY(ddfan_plan_set_param(P, DDFAN_PARAM_VIEWS_PER_WAVE, v));
                        TRY(ddfan_plan_set_param(P, DDFAN_PARAM_PROJ_STAGE, ps));
                        double ms;
                        TRY(time_ms(P->stream, reps, proj, &ms));
                        if (ms < bestP) {
                            bestP = ms;
                            best.ZT = z;
                            best.V = v;
                            best.tile = 0;
                            best.pstage = ps;
                        }
                    }
                }
                TRY(ddfan_plan_set_param(P, DDFAN_PARAM_PROJ_CHUNK, best.ZT));
                TRY(ddfan_plan_set_param(P, DDFAN_PARAM_VIEWS_PER_WAVE, best.V));
                TRY(ddfan_plan_set_param(P, DDFAN_PARAM_PROJ_STAGE, best.pstage));
                const int nmax = std::max(g.XN, g.YN);
                for (int t : {256, 128, 64}) {
                    if (t >= nmax) continue;
                    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_SLAB_TILE, t));
                    double ms;
                    TRY(time_ms(P->stream, reps, proj, &ms));
                    if (ms < bestP) {
                        bestP = ms;
                        best.tile = t;
                    }
                }
                best.projMs = bestP / nsub;  // per call
                double bestB = 1e30;
                for (int z : chunks) {
                    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_STAGE, 0));
                    TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_CHUNK, z));
                    for (int st : {1, 2, 4, 0}) {
                        if (!stage_fits(st, z)) continue;
                        TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_STAGE, st));
                        double ms;
                        TRY(time_ms(P->stream, reps, back, &ms));
                        if (ms < bestB) {
                            bestB = ms;
                            best.ZTb = z;
                            best.stage = st;
                        }
                    }
                }
                TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_CHUNK, best.ZTb));
                TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_STAGE, best.stage));
                best.bpass = P->backPass;
                // view passes of a long plain backprojection
                if (views > 96) {
                    for (int bp : {0, 288, 192, 144, 96}) {
                        if (bp >= views) continue;
                        TRY(ddfan_plan_set_param(P, DDFAN_PARAM_BACK_PASS, bp));
                        double ms;
                        TRY(time_ms(P->stream, reps, back, &ms));
                        if (ms < bestB) {
                            bestB = ms;
                            best.bpass = bp;
                        }
                    }
                }
                best.backMs = bestB;
                return DDFAN_OK;
            }();
            HIPCALL(hipStreamSynchronize(P->stream));
            if (rc != DDFAN_OK) {
                (void)apply(P, saved);
                return rc;
            }
            std::lock_guard<std::mutex> lk(g_tuneMu);
            g_tune[key] = best;
        }
        TRY(apply(P, best));
        if (out_params) {
            out_params[0] = best.ZT;
            out_params[1] = best.V;
            out_params[2] = best.tile;
            out_params[3] = best.ZTb;
            out_params[4] = best.stage;
            out_params[5] = best.pstage;
        }
        if (out_ms) {
            out_ms[0] = best.projMs;
            out_ms[1] = best.backMs;
        }
        return DDFAN_OK;
    });
}

int ddfan_plan_apply_tuned(ddfan_plan* P, int views, int* found) {
    return ddf::guarded([&]() -> int {
        CHECK(P && found, DDFAN_EINVAL, "NULL argument");
        if (views <= 0) views = P->PN;
        TuneVal v{};
        {
            std::lock_guard<std::mutex> lk(g_tuneMu);
            auto it = g_tune.find(key_of(P, views));
            *found = it != g_tune.end();
            if (!*found) return DDFAN_OK;
            v = it->second;
        }
        return apply(P, v);
    });
}

// One line per entry (bstage: backprojector staged views, pstage: projector
// staging, bpass: backprojector views per pass; tables written before bpass
// existed have 15 fields and leave it as it is):
// mode XN YN SLN DNU ZTmax views  ZT V tile ZTb bstage pstage  proj_ms back_ms  bpass
int ddfan_tune_save(const char* path) {
    return ddf::guarded([&]() -> int {
        CHECK(path, DDFAN_EINVAL, "path is NULL");
        FILE* f = fopen(path, "w");
        CHECK(f, DDFAN_EINVAL, "cannot write %s", path);
        fprintf(f, "# ddfan launch-shape table: mode XN YN SLN DNU ZTmax views  ZT V tile ZTb bstage "
                   "pstage  proj_ms back_ms  bpass\n");
        std::lock_guard<std::mutex> lk(g_tuneMu);
        for (const auto& kv : g_tune) {
            const TuneKey& k = kv.first;
            const TuneVal& v = kv.second;
            fprintf(f, "%d %d %d %d %d %d %d  %d %d %d %d %d %d  %.4f %.4f  %d\n", k.mode, k.XN,
                    k.YN, k.SLN, k.DNU, k.ZTmax, k.views, v.ZT, v.V, v.tile, v.ZTb, v.stage,
                    v.pstage, v.projMs, v.backMs, v.bpass);
        }
        const int bad = ferror(f);
        fclose(f);
        CHECK(!bad, DDFAN_EINVAL, "write to %s failed", path);
        return DDFAN_OK;
    });
}

int ddfan_tune_load(const char* path, int* entries) {
    return ddf::guarded([&]() -> int {
        CHECK(path, DDFAN_EINVAL, "path is NULL");
        FILE* f = fopen(path, "r");
        CHECK(f, DDFAN_EINVAL, "cannot read %s", path);
        char line[512];
        int n = 0, lineNo = 0, bad = 0;
        std::vector<std::pair<TuneKey, TuneVal>> rows;
        while (fgets(line, sizeof line, f)) {
            ++lineNo;
            if (line[0] == '#' || line[0] == '\n') continue;
            TuneKey k{};
            TuneVal v{};
            int tokens = 0;
            for (const char* c = line; *c;) {
                while (*c == ' ' || *c == '\t' || *c == '\n' || *c == '\r') ++c;
                if (!*c) break;
                ++tokens;
                while (*c && *c != ' ' && *c != '\t' && *c != '\n' && *c != '\r') ++c;
            }
            v.bpass = -1;
            if ((tokens != 15 && tokens != 16) ||
                sscanf(line, "%d %d %d %d %d %d %d %d %d %d %d %d %d %lf %lf %d", &k.mode, &k.XN,
                       &k.YN, &k.SLN, &k.DNU, &k.ZTmax, &k.views, &v.ZT, &v.V, &v.tile, &v.ZTb,
                       &v.stage, &v.pstage, &v.projMs, &v.backMs, &v.bpass) != tokens ||
                v.bpass < -1) {
                bad = lineNo;
                break;
            }
            const bool pow2 = v.ZT >= 1 && v.ZT <= 16 && (v.ZT & (v.ZT - 1)) == 0 && v.ZTb >= 1 &&
                              v.ZTb <= 16 && (v.ZTb & (v.ZTb - 1)) == 0;
            const bool pshape = (v.pstage == 0 || v.pstage == 1)
                                    ? (v.V == 1 || v.V == 2 || v.V == 4)
                                    : (v.pstage == 2 || v.pstage == 3) &&
                                          ((v.V >= 3 && v.V <= 6) || v.V == 9 || v.V == 12) &&
                                          v.ZT % (v.pstage == 2 ? 4 : 2) == 0;
            if (!pow2 || v.ZT > k.ZTmax || v.ZTb > k.ZTmax || !pshape || v.tile < 0 ||
                !stage_fits(v.stage, v.ZTb)) {
                bad = lineNo;
                break;
            }
            rows.push_back({k, v});
            ++n;
        }
        fclose(f);
        CHECK(!bad, DDFAN_EINVAL, "%s:%d: malformed launch-shape entry", path, bad);
        std::lock_guard<std::mutex> lk(g_tuneMu);
        for (const auto& r : rows) g_tune[r.first] = r.second;
        if (entries) *entries = n;
        return DDFAN_OK;
    });
}

int ddfan_tune_clear(void) {
    return ddf::guarded([&]() -> int {
        std::lock_guard<std::mutex> lk(g_tuneMu);
        g_tune.clear();
        return DDFAN_OK;
    });
}

}  // extern "C"
