// ddfan_api.hip -- C ABI: drop-in entry points, plans, OS-SART drivers.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <vector>

#include "ddfan_guard.h"
#include "ddfan_plan.h"

using namespace ddf;

namespace {
thread_local std::string t_err;
std::atomic<int> g_default_sat{DDFAN_SAT_F32};
}  // namespace

namespace ddf {

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
    return code;
}

int check_device(int* dev) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return fail(DDFAN_ENODEV, "no GPU visible");
    HIPCALL(hipGetDevice(dev));
    return DDFAN_OK;
}

int upload(DBuf& b, const void* src, size_t bytes, hipStream_t st) {
    HIPCALL(b.reserve(bytes));
    HIPCALL(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st));
    return DDFAN_OK;
}

int ensure_vol_sat(ddfan_plan* P) {
    const size_t bx = sat_vol_bytes(P->mode, P->nzc, P->ZT, P->g.XN, P->g.YN);
    const size_t by = sat_vol_bytes(P->mode, P->nzc, P->ZT, P->g.YN, P->g.XN);
    if (bx > P->satX.bytes || by > P->satY.bytes) {
        HIPCALL(hipStreamSynchronize(P->stream));
        HIPCALL(P->satX.reserve(bx));
        HIPCALL(P->satY.reserve(by));
    }
    return DDFAN_OK;
}

}  // namespace ddf

namespace ddf {

int check_subset(int PN, int first, int stride, int count) {
    CHECK(count >= 0 && stride >= 1 && first >= 0, DDFAN_EINVAL,
          "bad view subset (first %d stride %d count %d)", first, stride, count);
    if (count > 0)
        CHECK((long long)first + (long long)(count - 1) * stride < PN, DDFAN_EINVAL,
              "view subset (first %d stride %d count %d) exceeds %d views", first, stride, count,
              PN);
    return DDFAN_OK;
}

// One slice split per GPU: startIdx[k] is the first slice of shard k
// (DD2MutiSlices.m:95-101).  On one device the shards are the identity.
int check_start_idx(const int* startIdx, int nStart, int SLN) {
    if (!startIdx || nStart <= 0) return DDFAN_OK;
    CHECK(startIdx[0] == 0, DDFAN_EINVAL, "startIdx[0] must be 0 (got %d)", startIdx[0]);
    for (int k = 1; k < nStart; ++k)
        CHECK(startIdx[k] >= startIdx[k - 1] && startIdx[k] <= SLN, DDFAN_EINVAL,
              "startIdx must be non-decreasing within [0, SLN]");
    return DDFAN_OK;
}

}  // namespace ddf

// =================================================================== plan
namespace ddf {

void fill_geom(GeomDev& g, const ddfan_geom* G, int SLN) {
    g.x0 = G->x0;
    g.y0 = G->y0;
    g.dx = G->dx;
    g.invdx = 1.0f / G->dx;  // MultiSlicesProjGPU.cu:149
    g.ocxP = (G->XN - 1.0f) * 0.5f - G->imgXCenter / G->dx;  // :225-226
    g.ocyP = (G->YN - 1.0f) * 0.5f - G->imgYCenter / G->dx;
    g.ocxB = (float)((G->XN - 1.0) * 0.5 - (double)(G->imgXCenter / G->dx));  // Back :198-200
    g.ocyB = (float)((G->YN - 1.0) * 0.5 - (double)(G->imgYCenter / G->dx));
    g.inv_sid = 1.0f / sqrtf(G->x0 * G->x0 + G->y0 * G->y0);  // Back :141
    float S2D, dbeta;
    back_scalars(G->x0, G->y0, G->xds, G->yds, &S2D, &dbeta);
    g.invdb = (float)(1.0 / (double)dbeta);  // Back :150
    g.detCntIdx = G->detCntIdx;
    g.XN = G->XN;
    g.YN = G->YN;
    g.SLN = SLN;
    g.DNU = G->DNU;
    g.backMid = back_mid_range(G->XN, G->YN, g.dx, g.ocxB, g.ocyB, g.x0, g.y0);
}

int default_sat_mode() { return g_default_sat.load(); }

int validate_geom(const ddfan_geom* G) {
    CHECK(G, DDFAN_EINVAL, "geometry is NULL");
    CHECK(G->DNU >= 2, DDFAN_EINVAL, "DNU must be >= 2 (got %d)", G->DNU);
    CHECK(G->XN >= 1 && G->YN >= 1 && G->SLN >= 1, DDFAN_EINVAL, "XN, YN, SLN must be >= 1");
    CHECK(G->xds && G->yds, DDFAN_EINVAL, "xds/yds are NULL");
    CHECK(G->dx > 0.0f && isfinite(G->dx), DDFAN_EINVAL, "dx must be positive");
    CHECK((long long)G->XN * G->YN * G->SLN < (1LL << 40), DDFAN_EINVAL, "volume too large");
    return DDFAN_OK;
}

int check_proj_shape(const ddfan_plan* P) {
    CHECK(proj_shape_ok(P->mode, P->g, P->ZT, P->viewsPerWave, P->projStage), DDFAN_EINVAL,
          "projector launch shape (chunk %d, %d per wave, stage %d) does not fit this plan: "
          "stages 0 and 1 take 1, 2 or 4 views per wave; the whole-row projector (stage 2: 4 "
          "slices per block, 3: 2) takes 3, 4, 5, 6, 9 or 12 tasks per wave, a chunk that is a multiple "
          "of its slices and images up to 775 voxels wide",
          P->ZT, P->viewsPerWave, P->projStage);
    return DDFAN_OK;
}


}  // namespace ddf

extern "C" {

const char* ddfan_status_string(int s) {
    switch (s) {
        case DDFAN_OK: return "ok";
        case DDFAN_EINVAL: return "invalid argument";
        case DDFAN_ENOMEM: return "out of memory";
        case DDFAN_EBOUNDS: return "input out of bounds";
        case DDFAN_EDEVICE: return "device error";
        case DDFAN_ENODEV: return "no device";
        case DDFAN_EBUSY: return "rejected by memory budget";
        case DDFAN_EIO: return "file missing, unreadable or malformed";
        default: return "unknown status";
    }
}

const char* ddfan_last_error(void) { return t_err.c_str(); }

int ddfan_version(void) { return 1; }

int ddfan_device_count(int* count) {
    return ddf::guarded([&]() -> int {
        CHECK(count, DDFAN_EINVAL, "count is NULL");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
        return DDFAN_OK;
    });
}

int ddfan_set_default_sat_mode(int m) {
    return ddf::guarded([&]() -> int {
        CHECK(m >= 0 && m <= 2, DDFAN_EINVAL, "sat_mode must be 0, 1 or 2");
        g_default_sat.store(m);
        return DDFAN_OK;
    });
}

// ------------------------------------------------------------ tables
int ddfan_detector_arc(int DNU, float detCntIdx, float detCellWidth, float sdd, float sid,
                       float* xds, float* yds) {
    return ddf::guarded([&]() -> int {
        CHECK(DNU >= 1 && xds && yds, DDFAN_EINVAL, "bad detector arguments");
        detector_arc(DNU, detCntIdx, detCellWidth, sdd, sid, xds, yds);
        return DDFAN_OK;
    });
}

int ddfan_recon_mask(int XN, int YN, unsigned char* mask) {
    return ddf::guarded([&]() -> int {
        CHECK(XN >= 1 && YN >= 1 && mask, DDFAN_EINVAL, "bad mask arguments");
        recon_mask(XN, YN, mask);
        return DDFAN_OK;
    });
}

int ddfan_subset_views(int numViews, int numOS, int k, int* out, int* count) {
    return ddf::guarded([&]() -> int {
        CHECK(numViews >= 0 && numOS >= 1 && k >= 0 && k < numOS && count, DDFAN_EINVAL,
              "bad subset arguments");
        *count = subset_views(numViews, numOS, k, out);
        return DDFAN_OK;
    });
}

int ddfan_rebin_starts(const float* zPos, int SLN, float SD, float SO, int DetHeight,
                       float PerDetW, float PerDetH, int DefTimes, float pitch, int* start) {
    return ddf::guarded([&]() -> int {
        CHECK(zPos && start && SLN >= 0 && DefTimes >= 1 && DetHeight >= 1, DDFAN_EINVAL,
              "bad rebin arguments");
        const RebinParams rp = rebin_params(SD, SO, DetHeight, PerDetW, PerDetH, DefTimes, pitch);
        for (int s = 0; s < SLN; ++s) start[s] = rebin_start(zPos[s], rp.h, rp.deltaZ);
        return DDFAN_OK;
    });
}

// ------------------------------------------------------------ memory
int ddfan_dev_alloc(size_t bytes, void** ptr) {
    return ddf::guarded([&]() -> int {
        CHECK(ptr, DDFAN_EINVAL, "ptr is NULL");
        int dev;
        TRY(check_device(&dev));
        HIPCALL(hipMalloc(ptr, bytes ? bytes : 16));
        return DDFAN_OK;
    });
}
int ddfan_dev_free(void* ptr) {
    return ddf::guarded([&]() -> int {
        if (ptr) HIPCALL(hipFree(ptr));
        return DDFAN_OK;
    });
}
int ddfan_h2d(void* dst, const void* src, size_t bytes) {
    return ddf::guarded([&]() -> int {
        HIPCALL(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return DDFAN_OK;
    });
}
int ddfan_d2h(void* dst, const void* src, size_t bytes) {
    return ddf::guarded([&]() -> int {
        HIPCALL(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return DDFAN_OK;
    });
}
int ddfan_d2d(void* dst, const void* src, size_t bytes) {
    return ddf::guarded([&]() -> int {
        HIPCALL(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
        return DDFAN_OK;
    });
}
int ddfan_dev_mem_info(size_t* fb, size_t* tb) {
    return ddf::guarded([&]() -> int {
        int dev;
        TRY(check_device(&dev));
        HIPCALL(hipMemGetInfo(fb, tb));
        return DDFAN_OK;
    });
}

// ------------------------------------------------------------ plan
int ddfan_plan_footprint(const ddfan_geom* G, int PN, int sat_mode, int slice_chunk, int numOS,
                         size_t* bytes) {
    return ddf::guarded([&]() -> int {
        TRY(validate_geom(G));
        CHECK(bytes && PN >= 1 && sat_mode >= 0 && sat_mode <= 2, DDFAN_EINVAL, "bad arguments");
        const int ZT = slice_chunk > 0 ? slice_chunk : 16;
        const int nzc = (G->SLN + ZT - 1) / ZT;
        size_t b = sizeof(float) * 4 * (size_t)PN * G->SLN + sizeof(float) * 4 * (G->DNU + 1) +
                   (size_t)G->XN * G->YN;
        b += 2 * sat_vol_bytes(sat_mode, nzc, ZT, G->XN > G->YN ? G->XN : G->YN,
                               G->XN > G->YN ? G->XN : G->YN);
        if (numOS > 0) {
            const int cnt = (PN + numOS - 1) / numOS;
            const size_t nvox = (size_t)G->XN * G->YN * G->SLN;
            b += sat_sino_bytes(sat_mode, nzc, ZT, G->DNU, cnt);
            b += sizeof(float) * ((size_t)cnt * G->DNU * G->SLN + nvox + (size_t)numOS * cnt * G->DNU +
                                  (size_t)numOS * G->XN * G->YN);
            b += sat_sino_bytes(sat_mode, 1, 1, G->DNU, cnt) + 2 * sat_vol_bytes(sat_mode, 1, 1, G->XN, G->YN);
            b += sizeof(float) * 4 * (size_t)PN;  // slice 0's (cos, sin) and source tables
        } else {
            b += sat_sino_bytes(sat_mode, nzc, ZT, G->DNU, PN);
        }
        *bytes = b;
        return DDFAN_OK;
    });
}

int ddfan_plan_create(const ddfan_geom* G, const float* hangs, int PN, const unsigned char* mask,
                      int sat_mode, int slice_chunk, ddfan_plan** out) {
    return ddf::guarded([&]() -> int {
        CHECK(out, DDFAN_EINVAL, "out is NULL");
        *out = nullptr;
        TRY(validate_geom(G));
        CHECK(hangs && PN >= 1, DDFAN_EINVAL, "need PN >= 1 angles");
        // regularizeAngle subtracts 2*pi until the angle is in range; bounded
        // magnitudes keep that loop short (and finite).
        for (size_t i = 0; i < (size_t)PN * G->SLN; ++i)
            CHECK(isfinite(hangs[i]) && fabsf(hangs[i]) < 1.0e6f, DDFAN_EINVAL,
                  "angle %zu is not finite or exceeds 1e6 rad", i);
        CHECK(sat_mode >= 0 && sat_mode <= 2, DDFAN_EINVAL, "sat_mode must be 0, 1 or 2");
        CHECK(slice_chunk == 0 || slice_chunk == 1 || slice_chunk == 2 || slice_chunk == 4 ||
                  slice_chunk == 8 || slice_chunk == 16,
              DDFAN_EINVAL, "slice_chunk must be 0, 1, 2, 4, 8 or 16");
        int dev;
        TRY(check_device(&dev));
        ddfan_plan* P = new (std::nothrow) ddfan_plan();
        CHECK(P, DDFAN_ENOMEM, "host allocation failed");
        P->device = dev;
        int rc = [&]() -> int {
            HIPCALL(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
            const int SLN = G->SLN, DNU = G->DNU;
            fill_geom(P->g, G, SLN);
            P->PN = PN;
            P->mode = sat_mode;
            P->ZTmax = choose_slice_chunk(hangs, PN, SLN, slice_chunk);
            // Launch-shape defaults measured on MI355X at 512x512x64, 2304 views:
            // the projector stages its table rows through LDS on chunks of 8
            // slices with 4 views per wave in one slab pass (narrower chunks read
            // per lane), the backprojector runs on the widest chunk with its
            // detector windows staged one view at a time (narrow chunks stage
            // 4 / chunk views per group, so each wave copies a row).
            P->ZT = default_proj_chunk(P->ZTmax, SLN);
            P->projStage = P->ZT >= 4 ? 1 : 0;
            P->nzc = (SLN + P->ZT - 1) / P->ZT;
            P->ZTb = P->ZTmax;
            P->nzcb = (SLN + P->ZTb - 1) / P->ZTb;
            P->backStage = std::max(1, 4 / P->ZTb);
            P->viewsPerWave = 4;
            P->slabTile = 0;
            // a full-view plain backprojection runs in launches of 192 views
            // (measured at 512x512x64, 2304 views: two-float 19.2 -> 17.2 ms,
            // f32 17.0 -> 15.8, f64 21.0 -> 19.8; best 144-192)
            P->backPass = 192;
            P->h_hangs.assign(hangs, hangs + (size_t)PN * SLN);
            P->h_cs.resize((size_t)PN * SLN * 2);
            cossin_table(hangs, (size_t)PN * SLN, P->h_cs.data());
            std::vector<float> hbx(DNU + 1), hby(DNU + 1);
            dd_boundaries(DNU + 1, G->xds, hbx.data());
            dd_boundaries(DNU + 1, G->yds, hby.data());
            std::vector<uint8_t> hm;
            if (!mask) hm.assign((size_t)G->XN * G->YN, 1);
            std::vector<float> hsrc(P->h_cs.size());
            source_table(P->h_cs.data(), (size_t)PN * SLN, G->x0, G->y0, hsrc.data());
            TRY(upload(P->cs, P->h_cs.data(), sizeof(float) * P->h_cs.size(), P->stream));
            TRY(upload(P->src, hsrc.data(), sizeof(float) * hsrc.size(), P->stream));
            TRY(upload(P->bx, hbx.data(), sizeof(float) * (DNU + 1), P->stream));
            TRY(upload(P->by, hby.data(), sizeof(float) * (DNU + 1), P->stream));
            TRY(upload(P->xds, G->xds, sizeof(float) * DNU, P->stream));
            TRY(upload(P->yds, G->yds, sizeof(float) * DNU, P->stream));
            TRY(upload(P->mask, mask ? mask : hm.data(), (size_t)G->XN * G->YN, P->stream));
            HIPCALL(P->satX.reserve(sat_vol_bytes(sat_mode, P->nzc, P->ZT, G->XN, G->YN)));
            HIPCALL(P->satY.reserve(sat_vol_bytes(sat_mode, P->nzc, P->ZT, G->YN, G->XN)));
            HIPCALL(hipStreamSynchronize(P->stream));
            return DDFAN_OK;
        }();
        if (rc != DDFAN_OK) {
            delete P;
            return rc;
        }
        *out = P;
        return DDFAN_OK;
    });
}

int ddfan_plan_destroy(ddfan_plan* P) {
    return ddf::guarded([&]() -> int {
        if (!P) return DDFAN_OK;
        (void)hipSetDevice(P->device);
        delete P;
        return DDFAN_OK;
    });
}

int ddfan_plan_query(const ddfan_plan* P, int* zt, int* nzc, int* mode) {
    return ddf::guarded([&]() -> int {
        CHECK(P, DDFAN_EINVAL, "plan is NULL");
        if (zt) *zt = P->ZTmax;
        if (nzc) *nzc = (P->g.SLN + P->ZTmax - 1) / P->ZTmax;
        if (mode) *mode = P->mode;
        return DDFAN_OK;
    });
}

int ddfan_plan_set_param(ddfan_plan* P, int param, int value) {
    return ddf::guarded([&]() -> int {
        CHECK(P, DDFAN_EINVAL, "plan is NULL");
        switch (param) {
            case DDFAN_PARAM_SLAB_TILE:
                CHECK(value >= 0, DDFAN_EINVAL, "slab tile must be >= 0");
                P->slabTile = value;
                return DDFAN_OK;
            case DDFAN_PARAM_VIEWS_PER_WAVE:
                CHECK(value == 1 || value == 2 || (value >= 3 && value <= 6) || value == 9 ||
                          value == 12,
                      DDFAN_EINVAL,
                      "views per wave must be 1, 2 or 4 (tasks per wave 3, 4, 5, 6, 9 or 12 for "
                      "the whole-row projector)");
                P->viewsPerWave = value;
                return DDFAN_OK;
            case DDFAN_PARAM_PROJ_CHUNK:
            case DDFAN_PARAM_BACK_CHUNK: {
                CHECK(value >= 1 && value <= P->ZTmax && (value & (value - 1)) == 0, DDFAN_EINVAL,
                      "slice chunk must be a power of two <= %d (slices with identical angles)",
                      P->ZTmax);
                HIPCALL(hipSetDevice(P->device));
                HIPCALL(hipStreamSynchronize(P->stream));
                const int n = (P->g.SLN + value - 1) / value;
                if (param == DDFAN_PARAM_PROJ_CHUNK) {
                    P->ZT = value;
                    P->nzc = n;
                } else {
                    P->ZTb = value;
                    P->nzcb = n;
                    while (P->backStage * value > 32) P->backStage /= 2;  // LDS budget
                }
                return DDFAN_OK;
            }
            case DDFAN_PARAM_BACK_PASS:
                CHECK(value >= 0, DDFAN_EINVAL, "views per backprojector pass must be >= 0");
                P->backPass = value;
                return DDFAN_OK;
            case DDFAN_PARAM_PROJ_STAGE:
                CHECK(value >= 0 && value <= 3, DDFAN_EINVAL, "projector stage must be 0, 1, 2 or 3");
                HIPCALL(hipSetDevice(P->device));
                HIPCALL(hipStreamSynchronize(P->stream));
                P->projStage = value;
                return DDFAN_OK;
            case DDFAN_PARAM_BACK_STAGE:
                CHECK((value == 0 || value == 1 || value == 2 || value == 4) && value * P->ZTb <= 32,
                      DDFAN_EINVAL, "staged views must be 0, 1, 2 or 4 with views x chunk <= 32");
                HIPCALL(hipSetDevice(P->device));
                HIPCALL(hipStreamSynchronize(P->stream));
                P->backStage = value;
                return DDFAN_OK;
            case DDFAN_PARAM_BACK_MID_RANGE:
                return fail(DDFAN_EINVAL, "DDFAN_PARAM_BACK_MID_RANGE is read-only");
            default:
                return fail(DDFAN_EINVAL, "unknown plan parameter %d", param);
        }
    });
}

int ddfan_plan_get_param(const ddfan_plan* P, int param, int* value) {
    return ddf::guarded([&]() -> int {
        CHECK(P && value, DDFAN_EINVAL, "NULL argument");
        switch (param) {
            case DDFAN_PARAM_SLAB_TILE: *value = P->slabTile; return DDFAN_OK;
            case DDFAN_PARAM_VIEWS_PER_WAVE: *value = P->viewsPerWave; return DDFAN_OK;
            case DDFAN_PARAM_PROJ_CHUNK: *value = P->ZT; return DDFAN_OK;
            case DDFAN_PARAM_BACK_CHUNK: *value = P->ZTb; return DDFAN_OK;
            case DDFAN_PARAM_BACK_STAGE: *value = P->backStage; return DDFAN_OK;
            case DDFAN_PARAM_PROJ_STAGE: *value = P->projStage; return DDFAN_OK;
            case DDFAN_PARAM_BACK_PASS: *value = P->backPass; return DDFAN_OK;
            case DDFAN_PARAM_BACK_MID_RANGE: *value = P->g.backMid; return DDFAN_OK;
            default: return fail(DDFAN_EINVAL, "unknown plan parameter %d", param);
        }
    });
}

int ddfan_plan_sync(ddfan_plan* P) {
    return ddf::guarded([&]() -> int {
        CHECK(P, DDFAN_EINVAL, "plan is NULL");
        HIPCALL(hipSetDevice(P->device));
        HIPCALL(hipStreamSynchronize(P->stream));
        return DDFAN_OK;
    });
}

int ddfan_plan_proj(ddfan_plan* P, const float* d_vol, float* d_prj, int first, int stride,
                    int count) {
    return ddf::guarded([&]() -> int {
        CHECK(P && d_vol && d_prj, DDFAN_EINVAL, "NULL argument");
        TRY(check_subset(P->PN, first, stride, count));
        if (count == 0) return DDFAN_OK;
        TRY(check_proj_shape(P));
        HIPCALL(hipSetDevice(P->device));
        TRY(ensure_vol_sat(P));
        const Subset sub{first, stride, count};
        HIPCALL(P->timed(0, [&] {
            return launch_sat_vol(P->mode, P->ZT, P->g, P->nzc, d_vol, P->mask.as<uint8_t>(), P->satX.p,
                                  P->satY.p, P->stream);
        }));
        HIPCALL(P->timed(1, [&] {
            return launch_proj(P->mode, P->ZT, PROJ_PLAIN, P->g, P->nzc, P->cs.as<float2>(),
                               P->bx.as<float>(), P->by.as<float>(), P->xds.as<float>(),
                               P->yds.as<float>(), P->satX.p, P->satY.p, sub, d_prj, nullptr, nullptr,
                               P->slabTile, P->viewsPerWave, P->projStage, P->stream,
                               P->h_cs.data());
        }));
        return DDFAN_OK;
    });
}

int ddfan_plan_vol_tables(ddfan_plan* P, const float* d_vol, void* d_satX, void* d_satY,
                          size_t* bytes_x, size_t* bytes_y) {
    return ddf::guarded([&]() -> int {
        CHECK(P && bytes_x && bytes_y, DDFAN_EINVAL, "NULL argument");
        const size_t bx = sat_vol_bytes(P->mode, P->nzc, P->ZT, P->g.XN, P->g.YN) - 256;
        const size_t by = sat_vol_bytes(P->mode, P->nzc, P->ZT, P->g.YN, P->g.XN) - 256;
        *bytes_x = bx;
        *bytes_y = by;
        if (!d_satX && !d_satY) return DDFAN_OK;
        CHECK(d_vol && d_satX && d_satY, DDFAN_EINVAL, "NULL buffer");
        HIPCALL(hipSetDevice(P->device));
        TRY(ensure_vol_sat(P));
        HIPCALL(launch_sat_vol(P->mode, P->ZT, P->g, P->nzc, d_vol, P->mask.as<uint8_t>(),
                               P->satX.p, P->satY.p, P->stream));
        HIPCALL(hipMemcpyAsync(d_satX, P->satX.p, bx, hipMemcpyDeviceToDevice, P->stream));
        HIPCALL(hipMemcpyAsync(d_satY, P->satY.p, by, hipMemcpyDeviceToDevice, P->stream));
        HIPCALL(hipStreamSynchronize(P->stream));
        return DDFAN_OK;
    });
}

int ddfan_plan_sino_table(ddfan_plan* P, const float* d_prj, int count, void* d_out,
                          size_t* bytes) {
    return ddf::guarded([&]() -> int {
        CHECK(P && bytes, DDFAN_EINVAL, "NULL argument");
        CHECK(count >= 1 && count <= P->PN, DDFAN_EINVAL, "count must be in [1, %d]", P->PN);
        const size_t b = sat_sino_bytes(P->mode, P->nzcb, P->ZTb, P->g.DNU, count) - 256;
        *bytes = b;
        if (!d_out) return DDFAN_OK;
        CHECK(d_prj, DDFAN_EINVAL, "NULL buffer");
        HIPCALL(hipSetDevice(P->device));
        const size_t need = sat_sino_bytes(P->mode, P->nzcb, P->ZTb, P->g.DNU, count);
        if (need > P->satP.bytes) {
            HIPCALL(hipStreamSynchronize(P->stream));
            HIPCALL(P->satP.reserve(need));
        }
        HIPCALL(launch_sat_sino(P->mode, P->ZTb, P->g.DNU, P->g.SLN, P->nzcb, count, d_prj,
                                P->satP.p, P->stream));
        HIPCALL(hipMemcpyAsync(d_out, P->satP.p, b, hipMemcpyDeviceToDevice, P->stream));
        HIPCALL(hipStreamSynchronize(P->stream));
        return DDFAN_OK;
    });
}

int ddfan_plan_back(ddfan_plan* P, const float* d_prj, float* d_vol, int first, int stride,
                    int count) {
    return ddf::guarded([&]() -> int {
        CHECK(P && d_vol && d_prj, DDFAN_EINVAL, "NULL argument");
        TRY(check_subset(P->PN, first, stride, count));
        HIPCALL(hipSetDevice(P->device));
        const Subset sub{first, stride, count};
        const size_t need = sat_sino_bytes(P->mode, P->nzcb, P->ZTb, P->g.DNU, std::max(count, 1));
        if (need > P->satP.bytes) {
            HIPCALL(hipStreamSynchronize(P->stream));
            HIPCALL(P->satP.reserve(need));
        }
        HIPCALL(P->timed(2, [&] {
            return launch_sat_sino(P->mode, P->ZTb, P->g.DNU, P->g.SLN, P->nzcb, count, d_prj, P->satP.p,
                                   P->stream);
        }));
        HIPCALL(P->timed(3, [&] {
            return launch_back(P->mode, P->ZTb, BACK_PLAIN, P->g, P->nzcb, P->src.as<float2>(),
                               P->mask.as<uint8_t>(), P->satP.p, sub, d_vol, nullptr, nullptr, nullptr,
                               nullptr, P->backStage, P->stream, P->backPass);
        }));
        return DDFAN_OK;
    });
}

int ddfan_plan_timing(ddfan_plan* P, int enable) {
    return ddf::guarded([&]() -> int {
        CHECK(P, DDFAN_EINVAL, "plan is NULL");
        HIPCALL(hipSetDevice(P->device));
        HIPCALL(hipStreamSynchronize(P->stream));
        P->clear_events();
        P->timing = enable != 0;
        return DDFAN_OK;
    });
}

int ddfan_plan_timing_read(ddfan_plan* P, int kind, double* total_ms, int* launches) {
    return ddf::guarded([&]() -> int {
        CHECK(P && total_ms && launches, DDFAN_EINVAL, "NULL argument");
        HIPCALL(hipSetDevice(P->device));
        HIPCALL(hipStreamSynchronize(P->stream));
        double t = 0.0;
        int n = 0;
        for (auto& e : P->events) {
            if (e.kind != kind) continue;
            float ms = 0.0f;
            HIPCALL(hipEventElapsedTime(&ms, e.a, e.b));
            t += ms;
            ++n;
        }
        *total_ms = t;
        *launches = n;
        return DDFAN_OK;
    });
}

int ddfan_plan_trace_proj(ddfan_plan* P, int first, int stride, int count, int32_t* d_trace) {
    return ddf::guarded([&]() -> int {
        CHECK(P && d_trace, DDFAN_EINVAL, "NULL argument");
        TRY(check_subset(P->PN, first, stride, count));
        HIPCALL(hipSetDevice(P->device));
        HIPCALL(launch_proj_trace(P->g, P->cs.as<float2>(), P->bx.as<float>(), P->by.as<float>(),
                                  Subset{first, stride, count}, d_trace, P->stream));
        HIPCALL(hipStreamSynchronize(P->stream));
        return DDFAN_OK;
    });
}

int ddfan_plan_trace_back(ddfan_plan* P, int first, int stride, int count, int32_t* d_trace) {
    return ddf::guarded([&]() -> int {
        CHECK(P && d_trace, DDFAN_EINVAL, "NULL argument");
        TRY(check_subset(P->PN, first, stride, count));
        HIPCALL(hipSetDevice(P->device));
        HIPCALL(launch_back_trace(P->g, P->src.as<float2>(), Subset{first, stride, count}, d_trace,
                                  P->stream));
        HIPCALL(hipStreamSynchronize(P->stream));
        return DDFAN_OK;
    });
}

// ------------------------------------------------------------ drop-in entry points
int ddfan_proj(float x0, float y0, int DNU, const float* xds, const float* yds, float imgXCenter,
               float imgYCenter, const float* hangs, int PN, int XN, int YN, int SLN,
               const float* hvol, float* hprj, float dx, const unsigned char* mask,
               const int* startIdx, int nStart) {
    return ddf::guarded([&]() -> int {
        CHECK(hvol && hprj && hangs, DDFAN_EINVAL, "NULL buffer");
        ddfan_geom G{x0, y0, DNU, xds, yds, 0.0f, imgXCenter, imgYCenter, XN, YN, SLN, dx};
        TRY(validate_geom(&G));
        CHECK(PN >= 0, DDFAN_EINVAL, "PN must be >= 0 (got %d)", PN);
        TRY(check_start_idx(startIdx, nStart, SLN));
        if (PN == 0) return DDFAN_OK;
        ddfan_plan* P = nullptr;
        TRY(ddfan_plan_create(&G, hangs, PN, mask, g_default_sat.load(), 0, &P));
        const size_t nv = (size_t)XN * YN * SLN, np = (size_t)PN * DNU * SLN;
        int rc = [&]() -> int {
            DBuf dv, dp;
            HIPCALL(dv.reserve(sizeof(float) * nv));
            HIPCALL(dp.reserve(sizeof(float) * np));
            HIPCALL(hipMemcpyAsync(dv.p, hvol, sizeof(float) * nv, hipMemcpyHostToDevice, P->stream));
            TRY(ddfan_plan_proj(P, dv.as<float>(), dp.as<float>(), 0, 1, PN));
            HIPCALL(hipMemcpyAsync(hprj, dp.p, sizeof(float) * np, hipMemcpyDeviceToHost, P->stream));
            HIPCALL(hipStreamSynchronize(P->stream));
            return DDFAN_OK;
        }();
        ddfan_plan_destroy(P);
        return rc;
    });
}

int ddfan_back(float x0, float y0, int DNU, const float* xds, const float* yds, float detCntIdx,
               float imgXCenter, float imgYCenter, const float* hangs, int PN, int XN, int YN,
               int SLN, float* hvol, const float* hprj, float dx, const unsigned char* mask,
               const int* startIdx, int nStart) {
    return ddf::guarded([&]() -> int {
        CHECK(hvol && hprj && hangs, DDFAN_EINVAL, "NULL buffer");
        ddfan_geom G{x0, y0, DNU, xds, yds, detCntIdx, imgXCenter, imgYCenter, XN, YN, SLN, dx};
        TRY(validate_geom(&G));
        CHECK(PN >= 0, DDFAN_EINVAL, "PN must be >= 0 (got %d)", PN);
        TRY(check_start_idx(startIdx, nStart, SLN));
        const size_t nv = (size_t)XN * YN * SLN, np = (size_t)PN * DNU * SLN;
        if (PN == 0) {
            memset(hvol, 0, sizeof(float) * nv);
            return DDFAN_OK;
        }
        ddfan_plan* P = nullptr;
        TRY(ddfan_plan_create(&G, hangs, PN, mask, g_default_sat.load(), 0, &P));
        int rc = [&]() -> int {
            DBuf dv, dp;
            HIPCALL(dv.reserve(sizeof(float) * nv));
            HIPCALL(dp.reserve(sizeof(float) * np));
            HIPCALL(hipMemcpyAsync(dp.p, hprj, sizeof(float) * np, hipMemcpyHostToDevice, P->stream));
            TRY(ddfan_plan_back(P, dp.as<float>(), dv.as<float>(), 0, 1, PN));
            HIPCALL(hipMemcpyAsync(hvol, dv.p, sizeof(float) * nv, hipMemcpyDeviceToHost, P->stream));
            HIPCALL(hipStreamSynchronize(P->stream));
            return DDFAN_OK;
        }();
        ddfan_plan_destroy(P);
        return rc;
    });
}

// ------------------------------------------------------------ rebin
namespace {
int rebin_common(float* Proj, float* Views, const void* raw, bool rawOnDevice, int rawKind,
                 int rawOrder, float countScale, int layout, long long TotalViews,
                 const float* zPos, int SLN,
                 float SD, float SO, float BVAngle, int W, int H, float PerDetW, float PerDetH,
                 int DefTimes, float DetCenterW, float pitch, bool outOnDevice,
                 unsigned char* d_cut, int32_t* d_idv) {
    CHECK(Proj && Views && raw && (zPos || SLN == 0), DDFAN_EINVAL, "NULL buffer");
    CHECK(SLN >= 0 && W >= 1 && W <= 2048 && H >= 2 && DefTimes >= 1 && TotalViews >= 0,
          DDFAN_EINVAL, "bad rebin shape");
    CHECK(rawKind >= 0 && rawKind <= 2, DDFAN_EINVAL, "raw_kind must be 0, 1 or 2");
    CHECK(rawOrder == 0 || (rawOrder == 1 && rawOnDevice), DDFAN_EINVAL,
          "raw_order must be 0, or 1 with device-resident raw data");
    CHECK(layout == 0 || layout == 1, DDFAN_EINVAL, "layout must be 0 or 1");
    CHECK(isfinite(BVAngle) && fabsf(BVAngle) < 1.0e5f && isfinite(SD) && isfinite(SO) &&
              SD > 0.0f && SO > 0.0f && PerDetH > 0.0f && PerDetW > 0.0f && pitch > 0.0f,
          DDFAN_EINVAL, "bad rebin geometry");
    CHECK((double)TotalViews * (2.0 * 3.141592653589793 / DefTimes) < 1.0e6, DDFAN_EINVAL,
          "view angles beyond 1e6 rad");
    if (SLN == 0) return DDFAN_OK;
    const RebinParams rp = rebin_params(SD, SO, H, PerDetW, PerDetH, DefTimes, pitch);
    std::vector<int> start(SLN);
    long long vmin = (long long)1 << 62, vmax = -1;  // 1-based raw views read
    for (int s = 0; s < SLN; ++s) {
        start[s] = rebin_start(zPos[s], rp.h, rp.deltaZ);
        if (start[s] >= 0) {
            const long long lo = std::max(1, start[s]);
            const long long hi = (long long)start[s] + DefTimes - 1;
            CHECK(hi <= TotalViews, DDFAN_EBOUNDS,
                  "slice %d needs raw views up to %lld but only %lld exist", s, hi, TotalViews);
            if (hi >= lo) {
                vmin = std::min(vmin, lo);
                vmax = std::max(vmax, hi);
            }
        }
    }
    int dev;
    TRY(check_device(&dev));
    std::vector<float> SM(W);
    rebin_sm_table(W, SO, DetCenterW, rp.PerDetA, SM.data());
    RebinDev p{};
    p.SD = SD;
    p.SO = SO;
    p.BVAngle = BVAngle;
    p.PerDetH = PerDetH;
    p.DetCenterHreal = rp.DetCenterHreal;
    p.delta = rp.delta;
    p.deltaZ = rp.deltaZ;
    p.twopi = (float)(2 * 3.141592653589793);
    p.W = W;
    p.H = H;
    p.DefTimes = DefTimes;
    p.SLN = SLN;
    // Raw-view block order: with nondecreasing slice starts the slices that
    // read raw view i are a contiguous range lo..hi; used when the grid it
    // needs (views x the widest range) is at most twice the work.
    std::vector<int> lo, hi;
    {
        bool sorted = true;
        for (int s = 1; s < SLN; ++s) sorted = sorted && start[s] >= start[s - 1];
        const long long imin = start[0], imax = (long long)start[SLN - 1] + DefTimes - 1;
        const long long NI = imax - imin + 1;
        if (sorted && NI <= 0x3fffffffLL) {
            lo.resize(NI);
            hi.resize(NI);
            int a = 0, b = -1, R = 0;
            for (long long k = 0; k < NI; ++k) {
                const long long i = imin + k;
                while (b + 1 < SLN && start[b + 1] <= i) ++b;              // last s with start <= i
                while (a < SLN && (long long)start[a] + DefTimes - 1 < i) ++a;  // first s whose window reaches i
                lo[k] = a;
                hi[k] = b;
                R = std::max(R, b - a + 1);
            }
            p.imin = imin;
            p.NI = (int)NI;
            p.R = R;
            if (R == 0 || (double)NI * R > 2.0 * SLN * DefTimes + 64.0) {
                lo.clear();
                hi.clear();
            }
        }
    }
    hipStream_t st = nullptr;
    HIPCALL(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    int rc = [&]() -> int {
        DBuf dz, dstart, dsm, draw, dP, dV, dlo, dhi, dtmp;
        TRY(upload(dz, zPos, sizeof(float) * SLN, st));
        TRY(upload(dstart, start.data(), sizeof(int) * SLN, st));
        TRY(upload(dsm, SM.data(), sizeof(float) * W, st));
        if (!lo.empty()) {
            TRY(upload(dlo, lo.data(), sizeof(int) * lo.size(), st));
            TRY(upload(dhi, hi.data(), sizeof(int) * hi.size(), st));
            p.lo = dlo.as<int>();
            p.hi = dhi.as<int>();
        }
        const size_t esz = rawKind == 0 ? 4 : 2;
        p.totalViews = TotalViews;
        const size_t perView = (size_t)W * H * esz;
        const void* rawDev = raw;
        if (vmax < vmin) {  // no slice reads raw data
            vmin = 1;
            vmax = 0;
        }
        if (rawOnDevice) {
            p.viewOffset = 0;
        } else {
            p.viewOffset = vmin - 1;
            const size_t nbytes = perView * (size_t)(vmax - vmin + 1);
            HIPCALL(draw.reserve(nbytes));
            if (nbytes)
                HIPCALL(hipMemcpyAsync(draw.p, (const char*)raw + perView * (size_t)(vmin - 1),
                                       nbytes, hipMemcpyHostToDevice, st));
            rawDev = draw.p;
        }
        const size_t nOut = (size_t)SLN * DefTimes * W, nViews = (size_t)SLN * DefTimes;
        float* outP = Proj;
        float* outV = Views;
        if (!outOnDevice) {
            HIPCALL(dP.reserve(sizeof(float) * nOut));
            HIPCALL(dV.reserve(sizeof(float) * nViews));
            outP = dP.as<float>();
            outV = dV.as<float>();
        }
        if (layout == 1) HIPCALL(dtmp.reserve(sizeof(float) * nOut));
        HIPCALL(launch_rebin(p, rawKind, rawOrder, layout, rawDev, countScale, dz.as<float>(),
                             dstart.as<int>(), dsm.as<float>(), outP, outV, d_cut, d_idv,
                             dtmp.as<float>(), st));
        if (!outOnDevice) {
            HIPCALL(hipMemcpyAsync(Proj, outP, sizeof(float) * nOut, hipMemcpyDeviceToHost, st));
            HIPCALL(hipMemcpyAsync(Views, outV, sizeof(float) * nViews, hipMemcpyDeviceToHost, st));
        }
        HIPCALL(hipStreamSynchronize(st));
        return DDFAN_OK;
    }();
    (void)hipStreamDestroy(st);
    return rc;
}
}  // namespace

int ddfan_helical_to_fan(float* Proj, float* Views, const float* proj, long long TotalViews,
                         const float* zPos, int SLN, float SD, float SO, float BVAngle,
                         int DetWidth, int DetHeight, float PerDetW, float PerDetH, int DefTimes,
                         float DetCenterW, float SpiralPitchFactor) {
    return ddf::guarded([&]() -> int {
        return rebin_common(Proj, Views, proj, false, 0, 0, 1.0f, 0, TotalViews, zPos, SLN, SD, SO,
                            BVAngle, DetWidth, DetHeight, PerDetW, PerDetH, DefTimes, DetCenterW,
                            SpiralPitchFactor, false, nullptr, nullptr);
    });
}

int ddfan_helical_to_fan_dev(float* d_Proj, float* d_Views, const void* d_raw, int raw_kind,
                             int raw_order, float count_scale, int layout, long long TotalViews,
                             const float* zPos, int SLN, float SD, float SO, float BVAngle,
                             int DetWidth, int DetHeight, float PerDetW, float PerDetH,
                             int DefTimes, float DetCenterW, float SpiralPitchFactor,
                             unsigned char* d_cut, int32_t* d_idv) {
    return ddf::guarded([&]() -> int {
        CHECK(raw_kind == 0 || count_scale > 0.0f, DDFAN_EINVAL, "count_scale must be positive");
        return rebin_common(d_Proj, d_Views, d_raw, true, raw_kind, raw_order, count_scale, layout,
                            TotalViews,
                            zPos, SLN, SD, SO, BVAngle, DetWidth, DetHeight, PerDetW, PerDetH,
                            DefTimes, DetCenterW, SpiralPitchFactor, true, d_cut, d_idv);
    });
}

// ------------------------------------------------------------ plain host OS-SART
int ddfan_ossart_host_plain(float x0, float y0, int DNU, const float* xds, const float* yds,
                            float detCntIdx, float imgXCenter, float imgYCenter,
                            const float* hangs, int numViews, int XN, int YN, int SLN,
                            const float* meas, float* x, float dx, const unsigned char* mask,
                            int numIter, int numOS, int useFISTA, int sat_mode) {
    return ddf::guarded([&]() -> int {
        CHECK(hangs && meas && x, DDFAN_EINVAL, "NULL buffer");
        CHECK(numOS >= 1 && numOS <= numViews && numIter >= 0, DDFAN_EINVAL, "bad iteration counts");
        CHECK(sat_mode >= 0 && sat_mode <= 2, DDFAN_EINVAL, "bad sat_mode");
        const int saved = g_default_sat.exchange(sat_mode);
        struct Restore {
            int v;
            ~Restore() { g_default_sat.store(v); }
        } restore{saved};
        const size_t npix = (size_t)XN * YN, nvox = npix * SLN;
        const int maxc = subset_views(numViews, numOS, 0, nullptr);
        std::vector<std::vector<int>> idx(numOS);
        std::vector<std::vector<float>> rowSum(numOS), colSum(numOS);
        std::vector<float> hsub((size_t)maxc * SLN), prj((size_t)maxc * DNU * SLN), tp(nvox);
        std::vector<float> ones(std::max(nvox, (size_t)maxc * DNU * SLN), 1.0f), lasImg;
        // weights: full-slice calls, slice 1 kept and replicated (OSSART_AAPM.m:59-66)
        for (int k = 0; k < numOS; ++k) {
            const int n = subset_views(numViews, numOS, k, nullptr);
            idx[k].resize(n);
            subset_views(numViews, numOS, k, idx[k].data());
            for (int j = 0; j < n; ++j)
                for (int z = 0; z < SLN; ++z) hsub[(size_t)j * SLN + z] = hangs[(size_t)idx[k][j] * SLN + z];
            TRY(ddfan_proj(x0, y0, DNU, xds, yds, imgXCenter, imgYCenter, hsub.data(), n, XN, YN, SLN,
                           ones.data(), prj.data(), dx, mask, nullptr, 0));
            rowSum[k].resize((size_t)n * DNU);
            for (size_t r = 0; r < (size_t)n * DNU; ++r) rowSum[k][r] = prj[r * SLN];
            TRY(ddfan_back(x0, y0, DNU, xds, yds, detCntIdx, imgXCenter, imgYCenter, hsub.data(), n, XN,
                           YN, SLN, tp.data(), ones.data(), dx, mask, nullptr, 0));
            colSum[k].resize(npix);
            for (size_t p = 0; p < npix; ++p) colSum[k][p] = tp[p * SLN];
        }
        double t0 = 1.0;
        for (int it = 0; it < numIter; ++it) {
            if (useFISTA) lasImg.assign(x, x + nvox);
            for (int k = 0; k < numOS; ++k) {
                const int n = (int)idx[k].size();
                for (int j = 0; j < n; ++j)
                    for (int z = 0; z < SLN; ++z) hsub[(size_t)j * SLN + z] = hangs[(size_t)idx[k][j] * SLN + z];
                TRY(ddfan_proj(x0, y0, DNU, xds, yds, imgXCenter, imgYCenter, hsub.data(), n, XN, YN,
                               SLN, x, prj.data(), dx, mask, nullptr, 0));
                for (int j = 0; j < n; ++j)
                    for (int u = 0; u < DNU; ++u) {
                        const float w = rowSum[k][(size_t)j * DNU + u];
                        if (w == 0.0f) continue;
                        const float* b = meas + ((size_t)idx[k][j] * DNU + u) * SLN;
                        float* r = prj.data() + ((size_t)j * DNU + u) * SLN;
                        for (int z = 0; z < SLN; ++z) r[z] = (b[z] - r[z]) / w;
                    }
                TRY(ddfan_back(x0, y0, DNU, xds, yds, detCntIdx, imgXCenter, imgYCenter, hsub.data(), n,
                               XN, YN, SLN, tp.data(), prj.data(), dx, mask, nullptr, 0));
                for (size_t p = 0; p < npix; ++p) {
                    const float w = colSum[k][p];
                    for (int z = 0; z < SLN; ++z) {
                        float t = tp[p * SLN + z];
                        if (w != 0.0f) t = t / w;
                        x[p * SLN + z] = x[p * SLN + z] + t;
                    }
                }
            }
            if (useFISTA) {
                const double t1 = (1.0 + sqrt(1.0 + 4.0 * t0 * t0)) / 2.0;
                const float m = (float)((t0 - 1.0) / t1);
                for (size_t o = 0; o < nvox; ++o) x[o] = x[o] + m * (x[o] - lasImg[o]);
                t0 = t1;
            }
        }
        return DDFAN_OK;
    });
}

}  // extern "C"
