// ddfan_kernels.hip -- gfx950 kernels of the DD fan-beam path.
//
//   k_sat_vol    volume summed-area tables, both slab orientations, mask fused
//                (MultiSlicesProjGPU.cu:26-97,202-209)
//   k_proj       branchless DD projector (MultiSlicesProjGPU.cu:100-200) with an
//                optional OS-SART ratio epilogue (OSSART_AAPM.m:81-88)
//   k_sat_sino   detector-axis summed-area table of a sinogram
//                (MultiSlicesBackGPU.cu:12-66)
//   k_back       voxel-driven DD backprojector (MultiSlicesBackGPU.cu:115-184)
//                with optional OS-SART normalise/update and FISTA epilogues
//                (OSSART_AAPM.m:91-97,111-115)
//   k_rebin      helical-to-fan gather (HelicalToFanFunc_mex.cpp:41-139)
//
// No texture units: every fetch is an explicit clamp + FP32/FP64 lerp from
// global memory with the reference's texel-centred convention.  Geometry that
// decides an integer goes through ddfan_index.h and is bit-identical to the
// host restatement.  Slices of a chunk (ZT of them) share one view angle, so
// the per-ray index arithmetic is done once per chunk and the ZT slices ride
// in registers; a chunk of ZT = 1 is the general per-slice-angle path.
#include "ddfan_index.h"
#include "ddfan_internal.h"

namespace ddf {

template <int MODE>
struct Sat;
template <>
struct Sat<SAT_F32> {
    using T = float;
    static constexpr int W = 1;
};
template <>
struct Sat<SAT_TWOFLOAT> {
    using T = float;
    static constexpr int W = 2;
};
template <>
struct Sat<SAT_F64> {
    using T = double;
    static constexpr int W = 1;
};

// gfx950 hands consecutive workgroups to its 8 XCDs round-robin.  Blocks that
// share an XCD (and its L2) get a contiguous range of linear work items, so
// the concurrently resident blocks of one XCD sweep the same slice chunk.
__device__ __forceinline__ int xcd_linear(int b, int nblocks) {
    const int per = nblocks >> 3;  // grids are padded to a multiple of 8
    return (b & 7) * per + (b >> 3);
}
static inline int xcd_grid(int nwork) { return ((nwork + 7) / 8) * 8; }

template <int N>
__device__ __forceinline__ void ldv(float (&d)[N], const float* __restrict__ p) {
    if constexpr (N % 4 == 0) {
#pragma unroll
        for (int q = 0; q < N / 4; ++q) {
            const float4 t = reinterpret_cast<const float4*>(p)[q];
            d[4 * q] = t.x;
            d[4 * q + 1] = t.y;
            d[4 * q + 2] = t.z;
            d[4 * q + 3] = t.w;
        }
    } else if constexpr (N % 2 == 0) {
#pragma unroll
        for (int q = 0; q < N / 2; ++q) {
            const float2 t = reinterpret_cast<const float2*>(p)[q];
            d[2 * q] = t.x;
            d[2 * q + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int q = 0; q < N; ++q) d[q] = p[q];
    }
}

template <int N>
__device__ __forceinline__ void ldv(double (&d)[N], const double* __restrict__ p) {
    if constexpr (N % 2 == 0) {
#pragma unroll
        for (int q = 0; q < N / 2; ++q) {
            const double2 t = reinterpret_cast<const double2*>(p)[q];
            d[2 * q] = t.x;
            d[2 * q + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int q = 0; q < N; ++q) d[q] = p[q];
    }
}

// acc[zt] += (T(cR) - T(cL)) * scale for the ZT slices of one SAT row, where
// T(c) is the clamp-to-edge linear fetch at texel pair (i, i+1) with weight t.
template <int MODE, int ZT>
__device__ __forceinline__ void sat_diff_accum(float (&acc)[ZT],
                                               const typename Sat<MODE>::T* __restrict__ row,
                                               int N, int iL, float tL, int iR, float tR,
                                               float scale) {
    constexpr int STR = ZT * Sat<MODE>::W;
    const int l0 = min(max(iL, 0), N), l1 = min(max(iL + 1, 0), N);
    const int r0 = min(max(iR, 0), N), r1 = min(max(iR + 1, 0), N);
    tL = (l0 == l1) ? 0.0f : tL;
    tR = (r0 == r1) ? 0.0f : tR;
    if constexpr (MODE == SAT_F32) {
        float a0[ZT], a1[ZT], b0[ZT], b1[ZT];
        ldv(a0, row + (size_t)l0 * STR);
        ldv(a1, row + (size_t)l1 * STR);
        ldv(b0, row + (size_t)r0 * STR);
        ldv(b1, row + (size_t)r1 * STR);
#pragma unroll
        for (int q = 0; q < ZT; ++q) {
            const float TL = fmaf(tL, a1[q] - a0[q], a0[q]);
            const float TR = fmaf(tR, b1[q] - b0[q], b0[q]);
            acc[q] = fmaf(TR - TL, scale, acc[q]);
        }
    } else if constexpr (MODE == SAT_TWOFLOAT) {
        // T = hi[i0] + (lo[i0] + t*((hi[i1]-hi[i0]) + (lo[i1]-lo[i0]))) kept as a
        // (hi, lo) pair so T(cR)-T(cL) cancels without losing the low word.
        float h0[ZT], h1[ZT], g0[ZT], g1[ZT];
        ldv(h0, row + (size_t)l0 * STR);
        ldv(h1, row + (size_t)l1 * STR);
        ldv(g0, row + (size_t)r0 * STR);
        ldv(g1, row + (size_t)r1 * STR);
        float m0[ZT], m1[ZT], n0[ZT], n1[ZT];
        ldv(m0, row + (size_t)l0 * STR + ZT);
        ldv(m1, row + (size_t)l1 * STR + ZT);
        ldv(n0, row + (size_t)r0 * STR + ZT);
        ldv(n1, row + (size_t)r1 * STR + ZT);
#pragma unroll
        for (int q = 0; q < ZT; ++q) {
            const float TLlo = fmaf(tL, (h1[q] - h0[q]) + (m1[q] - m0[q]), m0[q]);
            const float TRlo = fmaf(tR, (g1[q] - g0[q]) + (n1[q] - n0[q]), n0[q]);
            const float D = (g0[q] - h0[q]) + (TRlo - TLlo);
            acc[q] = fmaf(D, scale, acc[q]);
        }
    } else {
        double a0[ZT], a1[ZT], b0[ZT], b1[ZT];
        ldv(a0, row + (size_t)l0 * STR);
        ldv(a1, row + (size_t)l1 * STR);
        ldv(b0, row + (size_t)r0 * STR);
        ldv(b1, row + (size_t)r1 * STR);
        const double dtL = tL, dtR = tR;
#pragma unroll
        for (int q = 0; q < ZT; ++q) {
            const double TL = fma(dtL, a1[q] - a0[q], a0[q]);
            const double TR = fma(dtR, b1[q] - b0[q], b0[q]);
            acc[q] = fmaf((float)(TR - TL), scale, acc[q]);
        }
    }
}

// ------------------------------------------------------------------ SAT
template <int MODE, int ZT>
struct SatWriter {
    using T = typename Sat<MODE>::T;
    static constexpr int W = Sat<MODE>::W;
    float hi = 0.0f, lo = 0.0f;
    double d = 0.0;
    __device__ __forceinline__ void add(float v) {
        if constexpr (MODE == SAT_F32) {
            hi = hi + v;
        } else if constexpr (MODE == SAT_TWOFLOAT) {
            const float s = hi + v;  // TwoSum(hi, v)
            const float bb = s - hi;
            const float e = (hi - (s - bb)) + (v - bb);
            hi = s;
            lo = lo + e;
        } else {
            d = d + (double)v;
        }
    }
    // out points at element zt of prefix entry k.
    __device__ __forceinline__ void store(T* __restrict__ out) const {
        if constexpr (MODE == SAT_F32) {
            out[0] = hi;
        } else if constexpr (MODE == SAT_TWOFLOAT) {
            const float H = hi + lo;  // FastTwoSum normalisation, |hi| >= |lo|
            const float L = lo - (H - hi);
            out[0] = H;
            out[ZT] = L;
        } else {
            out[0] = d;
        }
    }
};

template <int MODE, int ZT>
__global__ __launch_bounds__(256) void k_sat_vol(const GeomDev g, const int nzc,
                                                 const float* __restrict__ vol,
                                                 const uint8_t* __restrict__ mask,
                                                 typename Sat<MODE>::T* __restrict__ satX,
                                                 typename Sat<MODE>::T* __restrict__ satY) {
    constexpr int STR = ZT * Sat<MODE>::W;
    const int SLNp = nzc * ZT;
    const bool isY = blockIdx.y == 1;  // SAT along y: rows are image columns x
    const int nrows = isY ? g.XN : g.YN;
    const int Nk = isY ? g.YN : g.XN;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nrows * SLNp) return;
    const int z = (int)(idx % SLNp);
    const int row = (int)(idx / SLNp);
    const int zc = z / ZT, zt = z - zc * ZT;
    typename Sat<MODE>::T* out =
        (isY ? satY : satX) + ((size_t)zc * nrows + row) * (size_t)(Nk + 1) * STR + zt;
    const bool real = z < g.SLN;
    SatWriter<MODE, ZT> w;
    w.store(out);
    for (int k = 0; k < Nk; ++k) {
        float v = 0.0f;
        if (real) {
            const int x = isY ? row : k, y = isY ? k : row;
            const size_t p = (size_t)y * g.XN + x;
            v = vol[p * g.SLN + z] * (float)mask[p];
        }
        w.add(v);
        w.store(out + (size_t)(k + 1) * STR);
    }
}

template <int MODE, int ZT>
__global__ __launch_bounds__(256) void k_sat_sino(const int DNU, const int SLN, const int nzc,
                                                  const int PNs, const float* __restrict__ r,
                                                  typename Sat<MODE>::T* __restrict__ P) {
    constexpr int STR = ZT * Sat<MODE>::W;
    const int SLNp = nzc * ZT;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)PNs * SLNp) return;
    const int z = (int)(idx % SLNp);
    const int j = (int)(idx / SLNp);
    const int zc = z / ZT, zt = z - zc * ZT;
    typename Sat<MODE>::T* out = P + ((size_t)zc * PNs + j) * (size_t)(DNU + 1) * STR + zt;
    const bool real = z < SLN;
    SatWriter<MODE, ZT> w;
    w.store(out);
    for (int u = 0; u < DNU; ++u) {
        const float v = real ? r[((size_t)j * DNU + u) * SLN + z] : 0.0f;
        w.add(v);
        w.store(out + (size_t)(u + 1) * STR);
    }
}

// ------------------------------------------------------------------ projector
template <int MODE, int ZT, int EPI>
__global__ __launch_bounds__(64) void k_proj(
    const GeomDev g, const int nzc, const int nub, const int nwork,
    const float2* __restrict__ cossin, const float* __restrict__ bx, const float* __restrict__ by,
    const float* __restrict__ xds, const float* __restrict__ yds,
    const typename Sat<MODE>::T* __restrict__ satX, const typename Sat<MODE>::T* __restrict__ satY,
    const Subset sub, float* __restrict__ out, const float* __restrict__ meas,
    const float* __restrict__ rowSum) {
    using T = typename Sat<MODE>::T;
    constexpr int STR = ZT * Sat<MODE>::W;
    const int lin = xcd_linear(blockIdx.x, gridDim.x);
    if (lin >= nwork) return;
    const int ub = lin % nub;
    const int rest = lin / nub;
    const int j = rest % sub.count;
    const int zc = rest / sub.count;

    const int u = ub * 64 + threadIdx.x;
    const bool act = u < g.DNU;
    const int uu = act ? u : g.DNU - 1;
    const int vg = sub.first + j * sub.stride;
    const int z0 = zc * ZT;
    const float2 cs = cossin[(size_t)vg * g.SLN + min(z0, g.SLN - 1)];
    const ProjRay ray = proj_ray(g.x0, g.y0, cs.x, cs.y, bx[uu], by[uu], bx[uu + 1], by[uu + 1]);

    // continuous: dx / |normalize(det - src)| component (MultiSlicesProjGPU.cu:139,158,181)
    const float srcx = g.x0 * cs.x - g.y0 * cs.y, srcy = g.x0 * cs.y + g.y0 * cs.x;
    const float ddx = (xds[uu] * cs.x - yds[uu] * cs.y) - srcx;
    const float ddy = (xds[uu] * cs.y + yds[uu] * cs.x) - srcy;
    const float nrm = sqrtf(fmaf(ddx, ddx, ddy * ddy));
    const float constVal = g.dx * nrm / fabsf(ray.xslab ? ddx : ddy);

    const int Ns = ray.xslab ? g.XN : g.YN;
    const int Nk = ray.xslab ? g.YN : g.XN;
    const float ocS = ray.xslab ? g.ocxP : g.ocyP;
    const float ocK = ray.xslab ? g.ocyP : g.ocxP;
    const size_t rowStride = (size_t)(Nk + 1) * STR;
    const T* sat = (ray.xslab ? satY : satX) + (size_t)zc * Ns * rowStride;

    float acc[ZT];
#pragma unroll
    for (int q = 0; q < ZT; ++q) acc[q] = 0.0f;
    for (int ii = 0; ii < Ns; ++ii) {
        const float obj = slab_obj(ii, ocS, g.dx);
        const float rL = slab_r(obj, ray.aL, ray.fL, ray.bL);
        const float rR = slab_r(obj, ray.aR, ray.fR, ray.bR);
        const float cL = slab_coord(rL, g.invdx, ocK);
        const float cR = slab_coord(rR, g.invdx, ocK);
        const int iL = texel(cL, Nk), iR = texel(cR, Nk);
        const float tL = (cL - 0.5f) - (float)iL;
        const float tR = (cR - 0.5f) - (float)iR;
        const float invlen = __builtin_amdgcn_rcpf(rR - rL);
        sat_diff_accum<MODE, ZT>(acc, sat + (size_t)ii * rowStride, Nk, iL, tL, iR, tR, invlen);
    }
    if (!act) return;
    const size_t o0 = ((size_t)j * g.DNU + u) * g.SLN + z0;
#pragma unroll
    for (int q = 0; q < ZT; ++q) {
        const int z = z0 + q;
        if (z < g.SLN) {
            const float p = acc[q] * constVal;
            if constexpr (EPI == PROJ_PLAIN) {
                out[o0 + q] = p;
            } else {
                const float w = rowSum[(size_t)j * g.DNU + u];
                const float b = meas[((size_t)vg * g.DNU + u) * g.SLN + z];
                out[o0 + q] = (w != 0.0f) ? (b - p) / w : p;
            }
        }
    }
}

// ------------------------------------------------------------------ backprojector
template <int MODE, int ZT, int EPI>
__global__ __launch_bounds__(256) void k_back(
    const GeomDev g, const int nzc, const int ntx, const int nty, const int nwork,
    const float2* __restrict__ cossin, const uint8_t* __restrict__ mask,
    const typename Sat<MODE>::T* __restrict__ P, const Subset sub, float* __restrict__ vol,
    const float* __restrict__ colSum, float* __restrict__ xprev, const float* __restrict__ momTab,
    const int* __restrict__ iterCounter) {
    using T = typename Sat<MODE>::T;
    constexpr int STR = ZT * Sat<MODE>::W;
    const int lin = xcd_linear(blockIdx.x, gridDim.x);
    if (lin >= nwork) return;
    const int tile = lin % (ntx * nty);
    const int zc = lin / (ntx * nty);
    const int x = (tile % ntx) * 16 + (threadIdx.x & 15);
    const int y = (tile / ntx) * 16 + (threadIdx.x >> 4);
    const bool inb = x < g.XN && y < g.YN;
    const int m = inb ? mask[(size_t)y * g.XN + x] : 0;
    const int z0 = zc * ZT;
    const int zg = min(z0, g.SLN - 1);
    const int DNU = g.DNU;
    const size_t rowStride = (size_t)(DNU + 1) * STR;
    const T* Pc = P + (size_t)zc * sub.count * rowStride;

    float acc[ZT];
#pragma unroll
    for (int q = 0; q < ZT; ++q) acc[q] = 0.0f;
    if (m != 0) {
        const float vx = ((float)x - g.ocxB) * g.dx;
        const float vy = ((float)y - g.ocyB) * g.dx;
        for (int j = 0; j < sub.count; ++j) {
            const int vg = sub.first + j * sub.stride;
            const float2 cs = cossin[(size_t)vg * g.SLN + zg];
            const BackSample bs =
                back_sample(vx, vy, cs.x, cs.y, g.x0, g.y0, g.inv_sid, g.invdb, g.detCntIdx, g.dx);
            const int iL = texel(bs.idxL, DNU), iR = texel(bs.idxR, DNU);
            const float tL = (bs.idxL - 0.5f) - (float)iL;
            const float tR = (bs.idxR - 0.5f) - (float)iR;
            // dx / ddv * sqrt(l2) with sqrt(l2) = 1 / l  (MultiSlicesBackGPU.cu:173)
            const float cosVal = g.dx * __builtin_amdgcn_rcpf(bs.ddv * bs.l);
            sat_diff_accum<MODE, ZT>(acc, Pc + (size_t)j * rowStride, DNU, iL, tL, iR, tR, cosVal);
        }
    }
    if (!inb) return;
    const size_t pix = (size_t)y * g.XN + x;
    const size_t o0 = pix * g.SLN + z0;
    float w = 1.0f, mom = 0.0f;
    if constexpr (EPI != BACK_PLAIN) w = colSum[pix];
    if constexpr (EPI == BACK_UPDATE_FISTA) mom = momTab[*iterCounter];
#pragma unroll
    for (int q = 0; q < ZT; ++q) {
        if (z0 + q < g.SLN) {
            if constexpr (EPI == BACK_PLAIN) {
                vol[o0 + q] = acc[q];
            } else {
                const float upd = (w != 0.0f) ? acc[q] / w : acc[q];
                const float xn = vol[o0 + q] + upd;
                if constexpr (EPI == BACK_UPDATE) {
                    vol[o0 + q] = xn;
                } else {
                    const float xo = xn + mom * (xn - xprev[o0 + q]);
                    vol[o0 + q] = xo;
                    xprev[o0 + q] = xo;
                }
            }
        }
    }
}

// ------------------------------------------------------------------ index traces
__global__ __launch_bounds__(256) void k_proj_trace(const GeomDev g, const float2* __restrict__ cossin,
                                                    const float* __restrict__ bx,
                                                    const float* __restrict__ by, const Subset sub,
                                                    int32_t* __restrict__ trace) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)sub.count * g.DNU * g.SLN;
    if (idx >= n) return;
    const int z = (int)(idx % g.SLN);
    const int u = (int)((idx / g.SLN) % g.DNU);
    const int j = (int)(idx / ((long long)g.SLN * g.DNU));
    const int vg = sub.first + j * sub.stride;
    const float2 cs = cossin[(size_t)vg * g.SLN + z];
    const ProjRay ray = proj_ray(g.x0, g.y0, cs.x, cs.y, bx[u], by[u], bx[u + 1], by[u + 1]);
    const int Ns = ray.xslab ? g.XN : g.YN;
    const int Nk = ray.xslab ? g.YN : g.XN;
    const float ocS = ray.xslab ? g.ocxP : g.ocyP;
    const float ocK = ray.xslab ? g.ocyP : g.ocxP;
    const int Smax = g.XN > g.YN ? g.XN : g.YN;
    int32_t* tr = trace + (size_t)idx * Smax * 2;
    for (int ii = 0; ii < Ns; ++ii) {
        const float obj = slab_obj(ii, ocS, g.dx);
        tr[2 * ii] = texel(slab_coord(slab_r(obj, ray.aL, ray.fL, ray.bL), g.invdx, ocK), Nk);
        tr[2 * ii + 1] = texel(slab_coord(slab_r(obj, ray.aR, ray.fR, ray.bR), g.invdx, ocK), Nk);
    }
}

__global__ __launch_bounds__(256) void k_back_trace(const GeomDev g, const float2* __restrict__ cossin,
                                                    const Subset sub, int32_t* __restrict__ trace) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)g.XN * g.YN * g.SLN;
    if (idx >= n) return;
    const int z = (int)(idx % g.SLN);
    const int x = (int)((idx / g.SLN) % g.XN);
    const int y = (int)(idx / ((long long)g.SLN * g.XN));
    const float vx = ((float)x - g.ocxB) * g.dx;
    const float vy = ((float)y - g.ocyB) * g.dx;
    int32_t* tr = trace + (size_t)idx * sub.count * 2;
    for (int j = 0; j < sub.count; ++j) {
        const int vg = sub.first + j * sub.stride;
        const float2 cs = cossin[(size_t)vg * g.SLN + z];
        const BackSample bs =
            back_sample(vx, vy, cs.x, cs.y, g.x0, g.y0, g.inv_sid, g.invdb, g.detCntIdx, g.dx);
        tr[2 * j] = texel(bs.idxL, g.DNU);
        tr[2 * j + 1] = texel(bs.idxR, g.DNU);
    }
}

// ------------------------------------------------------------------ small helpers
__global__ void k_counter_inc(int* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = c[0] + 1;
}

__global__ __launch_bounds__(256) void k_fill(float* __restrict__ p, float v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ rebin
template <int RAW, int LAYOUT>
__global__ __launch_bounds__(256) void k_rebin(const RebinDev p, const void* __restrict__ raw,
                                               const float countScaleLog,
                                               const float* __restrict__ zPos,
                                               const int* __restrict__ start,
                                               const float* __restrict__ SM,
                                               float* __restrict__ Proj, float* __restrict__ Views,
                                               uint8_t* __restrict__ cut,
                                               int32_t* __restrict__ idvTrace) {
    constexpr int MAXC = 8;  // channels per thread: W <= 2048
    __shared__ int anyCut;
    const int s = blockIdx.x / p.DefTimes;
    const int a = blockIdx.x - s * p.DefTimes;
    if (threadIdx.x == 0) anyCut = 0;
    __syncthreads();
    const int st = start[s];
    const float z = zPos[s];
    const long long i = (long long)st + a;  // 1-based raw view of rebinned view a
    const bool valid = st >= 0 && i >= 1;
    float zz = 0.0f, angle = 0.0f;
    if (valid) {
        zz = (float)(i - 1) * p.deltaZ;
        const float lambda = (float)(i - 1) * p.delta + p.BVAngle;
        angle = mod_f(lambda, p.twopi);
    }
    const size_t viewBase = valid ? (size_t)(i - 1 - p.viewOffset) * p.W * p.H : 0;
    float vals[MAXC];
    int myCut = 0;
#pragma unroll
    for (int c = 0; c < MAXC; ++c) {
        const int jj = threadIdx.x + c * 256;
        vals[c] = 0.0f;
        if (jj < p.W) {
            int idv = 0, mk = 0;
            if (valid) {
                float v;
                const float dv = rebin_dv(z, zz, SM[jj], p.SD, p.DetCenterHreal, p.PerDetH, &v);
                idv = (int)floorf(dv);
                if (idv >= 1 && idv <= p.H - 1) {
                    const float t = dv - (float)idv;
                    const float cosA = p.SD / sqrtf(p.SD * p.SD + v * v);
                    const size_t o = viewBase + (size_t)jj * p.H + (idv - 1);
                    float d0, d1;
                    if constexpr (RAW == 0) {
                        const float* r = reinterpret_cast<const float*>(raw);
                        d0 = r[o];
                        d1 = r[o + 1];
                    } else {
                        const uint16_t* r = reinterpret_cast<const uint16_t*>(raw);
                        d0 = countScaleLog - logf(fmaxf((float)r[o], 1.0f));
                        d1 = countScaleLog - logf(fmaxf((float)r[o + 1], 1.0f));
                    }
                    vals[c] = cosA * (d0 * (1.0f - t) + d1 * t);
                    mk = 1;
                }
            }
            if (!mk) myCut = 1;
            if (idvTrace) idvTrace[((size_t)s * p.DefTimes + a) * p.W + jj] = idv;
        }
    }
    if (myCut) atomicOr(&anyCut, 1);
    __syncthreads();
    const bool isCut = anyCut != 0;
#pragma unroll
    for (int c = 0; c < MAXC; ++c) {
        const int jj = threadIdx.x + c * 256;
        if (jj < p.W) {
            const float val = isCut ? 0.0f : vals[c];
            if constexpr (LAYOUT == 0)
                Proj[((size_t)s * p.DefTimes + a) * p.W + jj] = val;
            else
                Proj[((size_t)a * p.W + jj) * p.SLN + s] = val;
        }
    }
    if (threadIdx.x == 0) {
        if constexpr (LAYOUT == 0)
            Views[(size_t)s * p.DefTimes + a] = angle;
        else
            Views[(size_t)a * p.SLN + s] = angle;
        if (cut) cut[(size_t)s * p.DefTimes + a] = isCut ? 1 : 0;
    }
}

// ================================================================== launchers
#define DDF_ZT_SWITCH(ZT, ...)                                            \
    switch (ZT) {                                                         \
        case 1: { constexpr int ZTC = 1; __VA_ARGS__; } break;           \
        case 2: { constexpr int ZTC = 2; __VA_ARGS__; } break;           \
        case 4: { constexpr int ZTC = 4; __VA_ARGS__; } break;           \
        case 8: { constexpr int ZTC = 8; __VA_ARGS__; } break;           \
        case 16: { constexpr int ZTC = 16; __VA_ARGS__; } break;         \
        default: return hipErrorInvalidValue;                             \
    }

#define DDF_MODE_SWITCH(MODE, ...)                                        \
    switch (MODE) {                                                       \
        case SAT_F32: { constexpr int MC = SAT_F32; __VA_ARGS__; } break; \
        case SAT_TWOFLOAT: { constexpr int MC = SAT_TWOFLOAT; __VA_ARGS__; } break; \
        case SAT_F64: { constexpr int MC = SAT_F64; __VA_ARGS__; } break; \
        default: return hipErrorInvalidValue;                             \
    }

hipError_t launch_sat_vol(int mode, int ZT, const GeomDev& g, int nzc, const float* vol,
                          const uint8_t* mask, void* satX, void* satY, hipStream_t st) {
    const long long n = (long long)(g.XN > g.YN ? g.XN : g.YN) * nzc * ZT;
    dim3 grid((unsigned)((n + 255) / 256), 2, 1);
    DDF_MODE_SWITCH(mode, DDF_ZT_SWITCH(ZT, {
        using T = typename Sat<MC>::T;
        k_sat_vol<MC, ZTC><<<grid, 256, 0, st>>>(g, nzc, vol, mask, (T*)satX, (T*)satY);
    }));
    return hipGetLastError();
}

hipError_t launch_sat_sino(int mode, int ZT, int DNU, int SLN, int nzc, int PNs, const float* r,
                           void* P, hipStream_t st) {
    const long long n = (long long)PNs * nzc * ZT;
    if (n == 0) return hipSuccess;
    dim3 grid((unsigned)((n + 255) / 256));
    DDF_MODE_SWITCH(mode, DDF_ZT_SWITCH(ZT, {
        using T = typename Sat<MC>::T;
        k_sat_sino<MC, ZTC><<<grid, 256, 0, st>>>(DNU, SLN, nzc, PNs, r, (T*)P);
    }));
    return hipGetLastError();
}

hipError_t launch_proj(int mode, int ZT, int epi, const GeomDev& g, int nzc, const float2* cossin,
                       const float* bx, const float* by, const float* xds, const float* yds,
                       const void* satX, const void* satY, Subset sub, float* out,
                       const float* meas, const float* rowSum, hipStream_t st) {
    if (sub.count <= 0) return hipSuccess;
    const int nub = (g.DNU + 63) / 64;
    const long long nw = (long long)nub * sub.count * nzc;
    if (nw > 0x7fffff00LL) return hipErrorInvalidValue;
    const int nwork = (int)nw;
    dim3 grid((unsigned)xcd_grid(nwork));
    DDF_MODE_SWITCH(mode, DDF_ZT_SWITCH(ZT, {
        using T = typename Sat<MC>::T;
        if (epi == PROJ_PLAIN)
            k_proj<MC, ZTC, PROJ_PLAIN><<<grid, 64, 0, st>>>(g, nzc, nub, nwork, cossin, bx, by,
                                                             xds, yds, (const T*)satX,
                                                             (const T*)satY, sub, out, meas, rowSum);
        else
            k_proj<MC, ZTC, PROJ_RATIO><<<grid, 64, 0, st>>>(g, nzc, nub, nwork, cossin, bx, by,
                                                             xds, yds, (const T*)satX,
                                                             (const T*)satY, sub, out, meas, rowSum);
    }));
    return hipGetLastError();
}

hipError_t launch_back(int mode, int ZT, int epi, const GeomDev& g, int nzc, const float2* cossin,
                       const uint8_t* mask, const void* P, Subset sub, float* vol,
                       const float* colSum, float* xprev, const float* momTab,
                       const int* iterCounter, hipStream_t st) {
    const int ntx = (g.XN + 15) / 16, nty = (g.YN + 15) / 16;
    const int nwork = ntx * nty * nzc;
    dim3 grid((unsigned)xcd_grid(nwork));
    DDF_MODE_SWITCH(mode, DDF_ZT_SWITCH(ZT, {
        using T = typename Sat<MC>::T;
        const T* Pt = (const T*)P;
        if (epi == BACK_PLAIN)
            k_back<MC, ZTC, BACK_PLAIN><<<grid, 256, 0, st>>>(g, nzc, ntx, nty, nwork, cossin, mask,
                                                              Pt, sub, vol, colSum, xprev, momTab,
                                                              iterCounter);
        else if (epi == BACK_UPDATE)
            k_back<MC, ZTC, BACK_UPDATE><<<grid, 256, 0, st>>>(g, nzc, ntx, nty, nwork, cossin,
                                                               mask, Pt, sub, vol, colSum, xprev,
                                                               momTab, iterCounter);
        else
            k_back<MC, ZTC, BACK_UPDATE_FISTA><<<grid, 256, 0, st>>>(
                g, nzc, ntx, nty, nwork, cossin, mask, Pt, sub, vol, colSum, xprev, momTab,
                iterCounter);
    }));
    return hipGetLastError();
}

hipError_t launch_proj_trace(const GeomDev& g, const float2* cossin, const float* bx,
                             const float* by, Subset sub, int32_t* trace, hipStream_t st) {
    const long long n = (long long)sub.count * g.DNU * g.SLN;
    if (n == 0) return hipSuccess;
    k_proj_trace<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(g, cossin, bx, by, sub, trace);
    return hipGetLastError();
}

hipError_t launch_back_trace(const GeomDev& g, const float2* cossin, Subset sub, int32_t* trace,
                             hipStream_t st) {
    const long long n = (long long)g.XN * g.YN * g.SLN;
    if (n == 0) return hipSuccess;
    k_back_trace<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(g, cossin, sub, trace);
    return hipGetLastError();
}

hipError_t launch_counter_inc(int* counter, hipStream_t st) {
    k_counter_inc<<<1, 64, 0, st>>>(counter);
    return hipGetLastError();
}

hipError_t launch_fill(float* p, float v, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_fill<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, v, n);
    return hipGetLastError();
}

hipError_t launch_rebin(const RebinDev& p, int rawKind, int outLayout, const void* raw,
                        float countScale, const float* zPos, const int* start, const float* SM,
                        float* Proj, float* Views, uint8_t* cut, int32_t* idvTrace,
                        hipStream_t st) {
    if (p.W > 8 * 256) return hipErrorInvalidValue;
    const long long nb = (long long)p.SLN * p.DefTimes;
    if (nb == 0) return hipSuccess;
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    const float csl = rawKind == 1 ? logf(countScale) : 0.0f;
#define DDF_REBIN(RK, LY)                                                                     \
    k_rebin<RK, LY><<<(unsigned)nb, 256, 0, st>>>(p, raw, csl, zPos, start, SM, Proj, Views, cut, \
                                                  idvTrace)
    if (rawKind == 0 && outLayout == 0) DDF_REBIN(0, 0);
    else if (rawKind == 0 && outLayout == 1) DDF_REBIN(0, 1);
    else if (rawKind == 1 && outLayout == 0) DDF_REBIN(1, 0);
    else if (rawKind == 1 && outLayout == 1) DDF_REBIN(1, 1);
    else return hipErrorInvalidValue;
#undef DDF_REBIN
    return hipGetLastError();
}

}  // namespace ddf
