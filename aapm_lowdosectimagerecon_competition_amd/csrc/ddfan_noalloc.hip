// ddfan_noalloc.hip -- the DD projector on device-resident buffers without
// any device allocation, for callers that cannot allocate (code under stream
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Its working tables (angle table, detector boundaries and centres, mask and
// the two volume summed-area tables) live in one file-scope device array,
// `scratch`, of kScratchBytes per GPU, held from the library's first such
// call on that GPU until the process ends.  In host code `scratch` names a
// host-side stand-in, not device memory: the device address comes from
// hipGetSymbolAddress once per device (with that device current) and only
// that address is handed to copies and kernels.  Calls on one GPU share its
// array, so they are serialised by a per-device mutex; each call writes every
// table entry its kernels read before reading it.  The kernels and launch
// shape are those of ddfan_proj with the same default SAT precision, so the
// results are bit-identical to it.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "ddfan_guard.h"
#include "ddfan_plan.h"

using namespace ddf;

namespace {

constexpr size_t kScratchBytes = (size_t)512 << 20;  // 512 MiB per GPU

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

struct DeviceScratch {
    unsigned char* ptr = nullptr;  // device address of `scratch` on this GPU
    std::mutex use;                // one call at a time uses the array
};

std::mutex g_mu;
std::map<int, DeviceScratch> g_scratch;

// The device address of `scratch` on device `dev` (current), looked up on the
// first call on that device.
int device_scratch(int dev, DeviceScratch** out) {
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceScratch& s = g_scratch[dev];
    if (!s.ptr) {
        void* p = nullptr;
        HIPCALL(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        CHECK(p, DDFAN_EDEVICE, "no device address for the scratch array");
        s.ptr = static_cast<unsigned char*>(p);
    }
    *out = &s;
    return DDFAN_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

int ddfan_proj_noalloc(float x0, float y0, int DNU, const float* xds, const float* yds,
                       float imgXCenter, float imgYCenter, const float* hangs, int PN, int XN,
                       int YN, int SLN, const float* d_vol, float* d_prj, float dx,
                       const unsigned char* mask, const int* startIdx, int nStart) {
    return ddf::guarded([&]() -> int {
        CHECK(d_vol && d_prj && hangs, DDFAN_EINVAL, "NULL buffer");
        ddfan_geom G{x0, y0, DNU, xds, yds, 0.0f, imgXCenter, imgYCenter, XN, YN, SLN, dx};
        TRY(validate_geom(&G));
        CHECK(PN >= 0, DDFAN_EINVAL, "PN must be >= 0 (got %d)", PN);
        TRY(check_start_idx(startIdx, nStart, SLN));
        if (PN == 0) return DDFAN_OK;
        for (size_t i = 0; i < (size_t)PN * SLN; ++i)
            CHECK(isfinite(hangs[i]) && fabsf(hangs[i]) < 1.0e6f, DDFAN_EINVAL,
                  "angle %zu is not finite or exceeds 1e6 rad", i);
        const int mode = default_sat_mode();
        // the launch shape of a plan's defaults (ddfan_plan_create)
        const int ZT = default_proj_chunk(choose_slice_chunk(hangs, PN, SLN, 0), SLN);
        const int nzc = (SLN + ZT - 1) / ZT;
        const int stage = ZT >= 4 ? 1 : 0;
        // scratch layout
        const size_t oCs = 0;
        const size_t oBx = oCs + align256(sizeof(float) * 2 * (size_t)PN * SLN);
        const size_t oBy = oBx + align256(sizeof(float) * (DNU + 1));
        const size_t oXd = oBy + align256(sizeof(float) * (DNU + 1));
        const size_t oYd = oXd + align256(sizeof(float) * DNU);
        const size_t oMk = oYd + align256(sizeof(float) * DNU);
        const size_t oSx = oMk + (mask ? align256((size_t)XN * YN) : 0);
        const size_t oSy = oSx + align256(sat_vol_bytes(mode, nzc, ZT, XN, YN));
        const size_t need = oSy + sat_vol_bytes(mode, nzc, ZT, YN, XN);
        CHECK(need <= kScratchBytes, DDFAN_EINVAL,
              "projection needs %zu bytes of scratch; the allocation-free projector has %zu",
              need, kScratchBytes);
        int dev;
        TRY(check_device(&dev));
        DeviceScratch* S = nullptr;
        TRY(device_scratch(dev, &S));
        // host tables (host staging reused by the calling thread)
        thread_local std::vector<float> hcs, hbx, hby;
        hcs.resize((size_t)PN * SLN * 2);
        hbx.resize(DNU + 1);
        hby.resize(DNU + 1);
        cossin_table(hangs, (size_t)PN * SLN, hcs.data());
        dd_boundaries(DNU + 1, xds, hbx.data());
        dd_boundaries(DNU + 1, yds, hby.data());
        GeomDev g{};
        fill_geom(g, &G, SLN);

        std::lock_guard<std::mutex> use(S->use);
        unsigned char* base = S->ptr;
        const hipStream_t st = nullptr;
        HIPCALL(hipMemcpyAsync(base + oCs, hcs.data(), sizeof(float) * hcs.size(),
                               hipMemcpyHostToDevice, st));
        HIPCALL(hipMemcpyAsync(base + oBx, hbx.data(), sizeof(float) * (DNU + 1),
                               hipMemcpyHostToDevice, st));
        HIPCALL(hipMemcpyAsync(base + oBy, hby.data(), sizeof(float) * (DNU + 1),
                               hipMemcpyHostToDevice, st));
        HIPCALL(hipMemcpyAsync(base + oXd, xds, sizeof(float) * DNU, hipMemcpyHostToDevice, st));
        HIPCALL(hipMemcpyAsync(base + oYd, yds, sizeof(float) * DNU, hipMemcpyHostToDevice, st));
        const uint8_t* dmask = nullptr;  // no mask: the tables take the samples unscaled
        if (mask) {
            HIPCALL(hipMemcpyAsync(base + oMk, mask, (size_t)XN * YN, hipMemcpyHostToDevice, st));
            dmask = base + oMk;
        }
        HIPCALL(launch_sat_vol(mode, ZT, g, nzc, d_vol, dmask, base + oSx, base + oSy, st));
        HIPCALL(launch_proj(mode, ZT, PROJ_PLAIN, g, nzc, reinterpret_cast<const float2*>(base + oCs),
                            reinterpret_cast<const float*>(base + oBx),
                            reinterpret_cast<const float*>(base + oBy),
                            reinterpret_cast<const float*>(base + oXd),
                            reinterpret_cast<const float*>(base + oYd), base + oSx, base + oSy,
                            Subset{0, 1, PN}, d_prj, nullptr, nullptr, 0, 4, stage, st));
        HIPCALL(hipStreamSynchronize(st));
        return DDFAN_OK;
    });
}

}  // extern "C"
