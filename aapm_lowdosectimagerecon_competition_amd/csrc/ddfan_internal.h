// ddfan_internal.h -- device-side data descriptors and kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace ddf {

// Summed-area-table precision variants.
enum SatMode : int {
    SAT_F32 = 0,       // FP32 sequential prefix (the reference's arithmetic)
    SAT_TWOFLOAT = 1,  // FP32 compensated (TwoSum) prefix stored as hi/lo pairs
    SAT_F64 = 2,       // FP64 prefix stored as double
};

// Bytes of one table entry: (S, v) float pair, (hi, lo, v, 0) floats, (S, v) doubles.
inline size_t sat_elem_bytes(int mode) { return mode == SAT_F32 ? 8 : 16; }

// Scalars every kernel needs, passed by value.
struct GeomDev {
    float x0, y0, dx, invdx;
    float ocxP, ocyP;  // projector image-centre indices (MultiSlicesProjGPU.cu:225-226)
    float ocxB, ocyB;  // backprojector image-centre indices (MultiSlicesBackGPU.cu:198-200)
    float inv_sid;     // 1/|source| (MultiSlicesBackGPU.cu:141)
    float invdb;       // 1/dbeta (MultiSlicesBackGPU.cu:150)
    float detCntIdx;
    int XN, YN, SLN, DNU;
    int backMid;  // back_mid_range(): the backprojector's footprint needs no range check
};

// Views first, first+stride, ... (count of them) of the plan's angle table.
struct Subset {
    int first, stride, count;
};

// Summed-area tables are stored as [chunk][row][zt][k] of entries pairing the
// prefix S[k] with the sample v[k]: `zt` runs over the ZT slices of a chunk
// that share one geometry, k = 0..N, rows are slabs (volume) or views
// (sinogram).

// Fused epilogues.
enum ProjEpi : int { PROJ_PLAIN = 0, PROJ_RATIO = 1 };
enum BackEpi : int { BACK_PLAIN = 0, BACK_UPDATE = 1, BACK_UPDATE_FISTA = 2 };

// Batched launches.  Every operator launch takes a list of independent jobs
// (one per reconstruction); jobs sharing the kernel's template shape go into
// one launch of up to kBatch jobs, whose blocks are laid out job after job.
constexpr int kBatch = 8;

struct SatVolJob {  // both volume tables of one reconstruction
    GeomDev g;
    int nzc, ZT;
    const float* vol;
    const uint8_t* mask;
    void* satX;
    void* satY;
};
struct SatSinoJob {  // detector-axis table of one sinogram subset
    int DNU, SLN, nzc, ZT, PNs;
    const float* r;
    void* P;
};
struct ProjJob {
    GeomDev g;
    Subset sub;
    int nzc, ZT, slabTile, epi;
    int V;  // views per wave: 1, 2 or 4 (stages 0, 1); tasks per wave 3-6, 9 or 12 (stages 2, 3)
    const float2* cossin;
    const float* bx;
    const float* by;
    const float* xds;
    const float* yds;
    const void* satX;
    const void* satY;
    float* out;
    const float* meas;
    const float* rowSum;
    int stage = 0;  // 0: per-lane gathers (k_proj); 1: LDS-staged runs (k_proj_lds);
                    // 2, 3: whole rows in LDS, 4 or 2 slices per block (k_proj_rows)
    // Host copy of the (cos, sin) table, [view][slice] pairs: sizes the
    // whole-row kernel's grid by the subset's runs of one slab orientation
    // (without it, one block per view).
    const float* hcs = nullptr;
};
struct BackJob {
    GeomDev g;
    Subset sub;
    int nzc, ZT, epi;
    const float2* src;  // source positions [view][slice] (source_table)
    const uint8_t* mask;
    const void* P;
    float* vol;
    const float* colSum;
    float* xprev;
    const float* momTab;
    const int* counter;
    int NV = 0;  // views per LDS-staged group (k_back_lds); 0 = per-lane gathers (k_back)
    int tabViews = 0;  // views per slice chunk of P (0: sub.count)
    int accum = 0;     // 1: continue the sums already in vol (a later view pass)
};

hipError_t launch_sat_vol_batch(int mode, const SatVolJob* jobs, int n, hipStream_t st);
hipError_t launch_sat_sino_batch(int mode, const SatSinoJob* jobs, int n, hipStream_t st);
// Launch-group limit of the batched launches (see launch_proj_batch): no limit
// beyond kBatch jobs per launch unless a caller passes one.
constexpr long long kNoFillLimit = 1LL << 40;
hipError_t launch_proj_batch(int mode, const ProjJob* jobs, int n, hipStream_t st,
                             long long fillWaves = kNoFillLimit);
hipError_t launch_back_batch(int mode, const BackJob* jobs, int n, hipStream_t st,
                             long long fillBlocks = kNoFillLimit);
hipError_t launch_counter_inc_batch(int* const* counters, int n, hipStream_t st);

// Single-job forms.
hipError_t launch_sat_vol(int mode, int ZT, const GeomDev& g, int nzc, const float* vol,
                          const uint8_t* mask, void* satX, void* satY, hipStream_t st);
hipError_t launch_sat_sino(int mode, int ZT, int DNU, int SLN, int nzc, int PNs, const float* r,
                           void* P, hipStream_t st);
hipError_t launch_proj(int mode, int ZT, int epi, const GeomDev& g, int nzc, const float2* cossin,
                       const float* bx, const float* by, const float* xds, const float* yds,
                       const void* satX, const void* satY, Subset sub, float* out,
                       const float* meas, const float* rowSum, int slabTile, int viewsPerWave,
                       int stage, hipStream_t st, const float* hcs = nullptr);
// viewPass > 0 (plain epilogue only): launches of at most viewPass views each,
// every one continuing the sums of the previous (same results as one launch).
hipError_t launch_back(int mode, int ZT, int epi, const GeomDev& g, int nzc, const float2* src,
                       const uint8_t* mask, const void* P, Subset sub, float* vol,
                       const float* colSum, float* xprev, const float* momTab,
                       const int* iterCounter, int stageViews, hipStream_t st, int viewPass = 0);

hipError_t launch_counter_inc(int* counter, hipStream_t st);
// The projector launch shape fits the geometry: stages 0 and 1 take 1, 2 or 4
// views per wave; the whole-row kernel (stage 2: 4 slices per block, 3: 2)
// takes 3, 4, 5, 6, 9 or 12 tasks per wave, a slice chunk that is a multiple of its
// slices and rows of at most 776 entries.
bool proj_shape_ok(int mode, const GeomDev& g, int ZT, int V, int stage);
hipError_t launch_fill(float* p, float v, size_t n, hipStream_t st);

// Index traces for the bit-exactness tests (same device functions as the
// kernels above): projector [j][u][z][slab][2], backprojector [y][x][z][j][2].
hipError_t launch_proj_trace(const GeomDev& g, const float2* cossin, const float* bx,
                             const float* by, Subset sub, int32_t* trace, hipStream_t st);
hipError_t launch_back_trace(const GeomDev& g, const float2* src, Subset sub, int32_t* trace,
                             hipStream_t st);

// Helical-to-fan rebin parameters (HelicalToFanFunc_mex.cpp:160-164).
struct RebinDev {
    float SD, SO, BVAngle, PerDetH, DetCenterHreal, delta, deltaZ, twopi;
    int W, H, DefTimes, SLN;
    long long viewOffset;  // raw view index (0-based) held at raw[0] (raw order 0)
    long long totalViews;  // raw views held (raw order 1 counts them from the end)
    // Block order by raw view (slice starts nondecreasing): the slices reading
    // raw view imin + k are lo[k]..hi[k]; block (k, r) is slice lo[k] + r, r < R.
    // lo == nullptr: block s * DefTimes + a.
    const int* lo;
    const int* hi;
    long long imin;
    int NI, R;
};

// rawKind 0: float raw line integrals; 1: uint16 counts converted to
// -log(max(c,1)/(gain*I0)) on the fly; 2: uint16 line integrals times
// countScale.  rawOrder 0: raw[(view*W+ch)*H+row] (the rebin's own input,
// HelicalToFanFunc_mex.cpp:45,74); 1: the same per-frame layout with the
// frames in reverse view order, f = TotalViews-1-view (the projection files
// in file order, which readProj.m:43-54 reads and flips).  outLayout 0: reference
// layout Proj[(s*DefTimes+v)*W+j], Views[s*DefTimes+v]; 1: sinogram layout
// prj[(v*W+j)*SLN+s], hangs[v*SLN+s].
// outLayout 1 goes through `tmp` (SLN * DefTimes * W floats, the reference
// layout) and a transpose; tmp is unused for layout 0.
hipError_t launch_rebin(const RebinDev& p, int rawKind, int rawOrder, int outLayout,
                        const void* raw, float countScale, const float* zPos, const int* start,
                        const float* SM, float* Proj, float* Views, uint8_t* cut,
                        int32_t* idvTrace, float* tmp, hipStream_t st);

}  // namespace ddf
