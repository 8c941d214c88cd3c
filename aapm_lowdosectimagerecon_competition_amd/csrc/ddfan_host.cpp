// ddfan_host.cpp -- host geometry tables (see ddfan_host.h).  Compiled with
// -ffp-contract=off so every expression rounds like the reference's host code.
#include "ddfan_host.h"

#include <math.h>
#include <string.h>

namespace ddf {

static const double kTwoPi = 6.283185307179586;  // utilities.cuh:43

void detector_arc(int DNU, float detCntIdx, float detCellWidth, float sdd, float sid, float* xds,
                  float* yds) {
    // MATLAB single arithmetic: stepTheta = atan(w/2/sdd)*2 (DD2MutiSlices.m:61)
    const float stepTheta = atanf(detCellWidth / 2.0f / sdd) * 2.0f;
    for (int ii = 1; ii <= DNU; ++ii) {
        const float beta = ((float)ii - detCntIdx) * stepTheta;
        xds[ii - 1] = sinf(beta) * sdd;
        yds[ii - 1] = sid - cosf(beta) * sdd;
    }
}

float regularize_angle(float c) {
    while ((double)c >= kTwoPi) c = (float)((double)c - kTwoPi);
    while (c < 0.0f) c = (float)((double)c + kTwoPi);
    return c;
}

void cossin_table(const float* hangs, size_t n, float* cs) {
    for (size_t i = 0; i < n; ++i) {
        const float a = regularize_angle(hangs[i]);
        cs[2 * i] = cosf(a);
        cs[2 * i + 1] = sinf(a);
    }
}

void source_table(const float* cs, size_t n, float x0, float y0, float* src) {
    for (size_t i = 0; i < n; ++i) {
        const float c = cs[2 * i], s = cs[2 * i + 1];
        src[2 * i] = x0 * c - y0 * s;
        src[2 * i + 1] = x0 * s + y0 * c;
    }
}

int back_mid_range(int XN, int YN, float dx, float ocxB, float ocyB, float x0, float y0) {
    // voxel centres (i - oc) * dx, i = 0..N-1: largest magnitude per axis
    const double ax = fmax(fabs((0.0 - ocxB) * dx), fabs((XN - 1.0 - ocxB) * dx));
    const double ay = fmax(fabs((0.0 - ocyB) * dx), fabs((YN - 1.0 - ocyB) * dx));
    const double rv = sqrt(ax * ax + ay * ay);      // bounding radius of the voxel centres
    const double rs = sqrt((double)x0 * x0 + (double)y0 * y0);  // source radius
    if (!(rs > 0.0) || !isfinite(rs) || !isfinite(rv)) return 0;
    const double dmin = rs - rv, dmax = rs + rv;  // voxel-source distance
    // ddv is the offset along the axis where the source's coordinate is the
    // larger in magnitude (>= rs / sqrt(2)), so |ddv| >= rs / sqrt(2) - max(ax, ay).
    const double vmin = rs * 0.7071067811865476 - fmax(ax, ay);
    const double lo2 = 0x1p-18, hi2 = 0x1p38, lo1 = 0x1p-38, hi1 = 0x1p18;
    // 1e-3 relative slack covers the FP32 rounding of the positions and sums
    return dmin > 0.0 && 0.999 * dmin * dmin >= lo2 && 1.001 * dmax * dmax <= hi2 &&
           0.999 * vmin >= lo1 && 1.001 * dmax <= hi1;
}

void dd_boundaries(int nb, const float* c, float* b) {
    if (nb >= 3) {
        for (int i = 0; i < nb; ++i) {
            if (i == 0)
                b[0] = (float)(1.5 * (double)c[0] - 0.5 * (double)c[1]);
            else if (i == nb - 1)
                b[i] = (float)(1.5 * (double)c[i - 1] - 0.5 * (double)c[i - 2]);
            else
                b[i] = (float)(0.5 * (double)c[i - 1] + 0.5 * (double)c[i]);
        }
    } else {
        b[0] = (float)((double)c[0] - 0.5);
        b[1] = (float)((double)c[0] + 0.5);
    }
}

void back_scalars(float x0, float y0, const float* xds, const float* yds, float* S2D,
                  float* dbeta) {
    const float s2d = hypotf(xds[0] - x0, yds[0] - y0);
    // sqrt of a float sum: correctly rounded in double then to float equals
    // the correctly rounded float square root.
    const float dts = (float)sqrt((double)(powf(xds[1] - xds[0], 2) + powf(yds[1] - yds[0], 2)));
    *S2D = s2d;
    *dbeta = atanf((float)((double)(dts / s2d) * 0.5)) * 2.0f;
}

void recon_mask(int XN, int YN, uint8_t* mask) {
    for (int jj = 1; jj <= YN; ++jj)
        for (int ii = 1; ii <= XN; ++ii) {
            const double a = ((double)ii - 0.5 - (double)XN / 2.0) / ((double)XN / 2.0);
            const double b = ((double)jj - 0.5 - (double)YN / 2.0) / ((double)YN / 2.0);
            mask[(size_t)(jj - 1) * XN + (ii - 1)] = sqrt(a * a + b * b) < 1.3 ? 1 : 0;
        }
}

int subset_views(int numViews, int numOS, int k, int* out) {
    int n = 0;
    for (int v = k; v < numViews; v += numOS) {
        if (out) out[n] = v;
        ++n;
    }
    return n;
}

RebinParams rebin_params(float SD, float SO, int DetHeight, float PerDetW, float PerDetH,
                         int DefTimes, float pitch) {
    RebinParams p;
    p.PerDetA = 2.0f * atanf(PerDetW / SD * 0.5f);
    p.delta = (float)(2.0 * 3.141592653589793 / DefTimes);
    p.DetCenterHreal = (float)((DetHeight - 1.0) * 0.5 / PerDetH);
    p.h = pitch * PerDetH * (float)DetHeight * SO / SD;
    p.deltaZ = p.h / (float)DefTimes;
    return p;
}

int rebin_start(float z, float h, float deltaZ) { return (int)floorf((z - h * 0.5f) / deltaZ); }

void rebin_sm_table(int W, float SO, float DetCenterW, float PerDetA, float* SM) {
    for (int j = 1; j <= W; ++j) SM[j - 1] = SO * cosf(((float)j - DetCenterW) * PerDetA);
}

int choose_slice_chunk(const float* hangs, int PN, int SLN, int limit) {
    static const int cands[] = {16, 8, 4, 2};
    for (int ZT : cands) {
        if (limit > 0 && ZT > limit) continue;
        if (ZT >= 2 * SLN) continue;  // at most half of a chunk is padding
        bool ok = true;
        for (int v = 0; v < PN && ok; ++v) {
            const float* h = hangs + (size_t)v * SLN;
            for (int z = 0; z < SLN && ok; ++z) {
                const int z0 = (z / ZT) * ZT;
                if (memcmp(&h[z], &h[z0], sizeof(float)) != 0) ok = false;
            }
        }
        if (ok) return ZT;
    }
    return 1;
}

int default_proj_chunk(int ZTmax, int SLN) {
    for (int ZT = ZTmax < 8 ? ZTmax : 8; ZT > 1; ZT >>= 1) {
        const int nzc = (SLN + ZT - 1) / ZT;
        if ((long long)(nzc * ZT - SLN) * 8 <= SLN) return ZT;
    }
    return 1;
}

}  // namespace ddf
