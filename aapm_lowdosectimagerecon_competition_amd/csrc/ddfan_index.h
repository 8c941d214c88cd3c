// ddfan_index.h -- FP32 index path shared by the GPU kernels and the host.
//
// Every expression here decides an integer (slab branch, texel index, rebin
// detector row) or feeds one, so it is written as a sequence of IEEE-754
// single operations that round identically on gfx950 and on the host:
// + - * / sqrtf fmaf floorf only, no FMA contraction (the library is compiled
// with -ffp-contract=off; fmaf appears only where it is written out), and
// correctly rounded division and square root (on the device, where operands
// allow, as short sequences from the hardware estimates that give the same
// bits).  Where the reference relied on CUDA fast-math intrinsics (rsqrtf,
// asinf, approximate divide) the correctly rounded operation or the Cephes
// arcsine below is used instead.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DDF_HD __host__ __device__ __forceinline__
#else
#define DDF_HD static inline
#endif

namespace ddf {

// Single-precision arcsine from the Cephes method (S. L. Moshier, Cephes Math
// Library asinf): odd minimax polynomial on [0, 0.5] and
// asin(x) = pi/2 - 2 asin(sqrt((1-x)/2)) above.  Horner steps in fmaf.
DDF_HD float asinf_cephes(float xx) {
    float a = fabsf(xx);
    if (a > 1.0f) return __builtin_nanf("");
    if (a < 1.0e-4f) return xx;
    float x, z;
    bool flag = a > 0.5f;
    if (flag) {
        z = 0.5f * (1.0f - a);
        x = sqrtf(z);
    } else {
        x = a;
        z = x * x;
    }
    float p = fmaf(4.2163199048e-2f, z, 2.4181311049e-2f);
    p = fmaf(p, z, 4.5470025998e-2f);
    p = fmaf(p, z, 7.4953002686e-2f);
    p = fmaf(p, z, 1.6666752422e-1f);
    z = fmaf(p * z, x, x);
    if (flag) {
        z = z + z;
        z = 1.5707963267948966f - z;
    }
    return xx < 0.0f ? -z : z;
}

#if defined(__HIPCC__)
// asinf_cephes for a wave: when no lane's |x| exceeds 0.5 (a NaN fails the
// test as well), the polynomial branch alone, with the sign copied onto the
// positive result; the same bits as asinf_cephes, without the select and
// checks of the other branches.  Else every lane takes asinf_cephes.
__device__ __forceinline__ float asinf_cephes_wave(float xx) {
    const float a = fabsf(xx);
    if (__builtin_amdgcn_ballot_w64(!(a <= 0.5f)) == 0) {
        const float z = a * a;
        float p = fmaf(4.2163199048e-2f, z, 2.4181311049e-2f);
        p = fmaf(p, z, 4.5470025998e-2f);
        p = fmaf(p, z, 7.4953002686e-2f);
        p = fmaf(p, z, 1.6666752422e-1f);
        const float r = fmaf(p * z, a, a);  // > 0 for a >= 1e-4
        return a < 1.0e-4f ? xx : copysignf(r, xx);
    }
    return asinf_cephes(xx);
}

// Correctly rounded square root, reciprocal and quotient for operands in a
// middle range.  The library is compiled with
// -fhip-fp32-correctly-rounded-divide-sqrt, whose expansions of sqrtf and '/'
// rescale operands near the ends of the exponent range and fix up zeros,
// infinities and NaNs around the same core: the hardware estimate (v_sqrt_f32,
// v_rcp_f32) and its residual corrections.  For a positive normal x in
// [2^-20, 2^40], and for a quotient whose divisor lies there and whose dividend
// is 0 or of magnitude in [2^-40, 2^20], neither the rescaling nor the fix-up
// changes anything, so a short sequence from the hardware estimate gives the
// same bits (P. Markstein, IA-64 and Elementary Functions, 2000, residual
// corrections with a fused multiply-add): the square root as
// g = x y, g + (x - g g)(y / 2) from y = v_rsq_f32(x), the reciprocal as one
// Newton step from v_rcp_f32 (both exhaustively checked: the correctly rounded
// sqrt(x) and 1/x for every float x of that range), and the quotient as one
// residual correction of q = a * (1/b) (with the correctly rounded reciprocal
// and q within an ulp, q + (a - b q)(1/b) rounds to the correctly rounded
// quotient).
__device__ __forceinline__ float sqrt_rn_mid(float x) {
    const float y = __builtin_amdgcn_rsqf(x);
    const float g = x * y, h = 0.5f * y;
    return fmaf(fmaf(-g, g, x), h, g);
}
__device__ __forceinline__ float rcp_rn_mid(float b) {
    const float r = __builtin_amdgcn_rcpf(b);
    return fmaf(fmaf(-b, r, 1.0f), r, r);
}
__device__ __forceinline__ float div_rn_mid(float a, float b) {
    const float r = rcp_rn_mid(b);
    const float q = a * r;
    return fmaf(fmaf(-b, q, a), r, q);
}
__device__ __forceinline__ bool rn_mid_divisor(float b) { return b >= 0x1p-20f && b <= 0x1p40f; }
__device__ __forceinline__ bool rn_mid_dividend(float a) {
    const float m = fabsf(a);
    return a == 0.0f || (m >= 0x1p-40f && m <= 0x1p20f);
}
#endif

// Texel index of an unnormalised, texel-centred linear fetch at coordinate c:
// floor(c - 0.5), limited to [-1, N] (anything outside reads the clamped edge).
DDF_HD int texel(float c, int N) {
    float f = floorf(c - 0.5f);
    f = (f >= -1.0f) ? f : -1.0f;  // NaN -> -1
    f = (f > (float)N) ? (float)N : f;
    return (int)f;
}

// Projector ray set-up for one (view, slice, detector)
// (MultiSlicesProjGPU.cu:122-158,176-181).
struct ProjRay {
    float aL, bL, fL;  // boundary ray L: r(obj) = (obj - aL) * fL + bL
    float aR, bR, fR;
    bool xslab;        // |cos| <= |sin|: slabs are image columns, SAT along y
};

// One detector boundary ray: detector u's left boundary is boundary u and its
// right boundary is boundary u+1, so neighbouring detectors share a boundary
// ray bit for bit (MultiSlicesProjGPU.cu:131-137,155-156,178-179).
struct ProjBoundary {
    float a, b, f;  // r(obj) = (obj - a) * f + b
    bool xslab;
};

DDF_HD bool proj_xslab(float c, float s) { return fabsf(c) <= fabsf(s); }

DDF_HD ProjBoundary proj_boundary(float x0, float y0, float c, float s, float bx, float by) {
    const float srcx = x0 * c - y0 * s;
    const float srcy = x0 * s + y0 * c;
    const float dx = bx * c - by * s, dy = bx * s + by * c;
    ProjBoundary r;
    r.xslab = proj_xslab(c, s);
    if (r.xslab) {
        r.f = (dy - srcy) / (dx - srcx);
        r.a = dx;
        r.b = dy;
    } else {
        r.f = (dx - srcx) / (dy - srcy);
        r.a = dy;
        r.b = dx;
    }
    return r;
}

DDF_HD ProjRay proj_ray(float x0, float y0, float c, float s, float bxL, float byL, float bxR,
                        float byR) {
    const ProjBoundary L = proj_boundary(x0, y0, c, s, bxL, byL);
    const ProjBoundary R = proj_boundary(x0, y0, c, s, bxR, byR);
    ProjRay r;
    r.xslab = L.xslab;
    r.aL = L.a; r.bL = L.b; r.fL = L.f;
    r.aR = R.a; r.bR = R.b; r.fR = R.f;
    return r;
}

// Boundary position on slab ii and its texture coordinate
// (MultiSlicesProjGPU.cu:161-168): obj = (ii - ocSlab)*dx,
// r = (obj - a)*f + b, coord = r*invdx + ocK + 1.
DDF_HD float slab_obj(int ii, float ocSlab, float dx) { return ((float)ii - ocSlab) * dx; }
DDF_HD float slab_r(float obj, float a, float f, float b) { return (obj - a) * f + b; }
DDF_HD float slab_coord(float r, float invdx, float ocK) { return (r * invdx + ocK) + 1.0f; }

// Backprojector per (voxel, view) detector coordinates
// (MultiSlicesBackGPU.cu:159-176).  vx, vy: voxel centre in mm; srcx, srcy:
// the view's source position (MultiSlicesBackGPU.cu:155-157, source_table).
// mid: the geometry keeps every operand of the footprint's square root and
// quotients in the short sequences' range (back_mid_range), so no lane needs
// checking.
struct BackSample {
    float idxL, idxR;  // texture coordinates along the detector SAT
    float ddv, l2, l;  // for the continuous weight dx/ddv*sqrt(l2)
};

DDF_HD BackSample back_sample(float vx, float vy, float srcx, float srcy, float inv_sid,
                              float invdb, float detCntIdx, float dx, int mid) {
    const float ddx = vx - srcx, ddy = vy - srcy;
    BackSample b;
    b.l2 = ddx * ddx + ddy * ddy;
#if defined(__HIP_DEVICE_COMPILE__)
    // |srcx| > |srcy| on the magnitude bits (the same order for finite
    // values), so a view-uniform source stays on the scalar unit
    const bool xmaj = (__float_as_uint(srcx) & 0x7fffffffu) > (__float_as_uint(srcy) & 0x7fffffffu);
#else
    const bool xmaj = fabsf(srcx) > fabsf(srcy);
#endif
    b.ddv = xmaj ? ddx : ddy;
    float q;  // ddv / l2
#if defined(__HIP_DEVICE_COMPILE__)
    // the short correctly rounded sequences when the geometry, or else every
    // lane of the wave (a wave-uniform branch), is in their range; else the
    // compiler's full ones
    if (mid || __builtin_amdgcn_ballot_w64(!(rn_mid_divisor(b.l2) && rn_mid_dividend(b.ddv))) == 0) {
        b.l = rcp_rn_mid(sqrt_rn_mid(b.l2));
        q = div_rn_mid(b.ddv, b.l2);
    } else {
        b.l = 1.0f / sqrtf(b.l2);
        q = b.ddv / b.l2;
    }
#else
    (void)mid;
    b.l = 1.0f / sqrtf(b.l2);
    q = b.ddv / b.l2;
#endif
#if defined(__HIP_DEVICE_COMPILE__)
    const float alpha = asinf_cephes_wave(((srcy * ddx - srcx * ddy) * inv_sid) * b.l);
#else
    const float alpha = asinf_cephes(((srcy * ddx - srcx * ddy) * inv_sid) * b.l);
#endif
    const float dA = (q * dx) * 0.5f;
    b.idxL = ((alpha - dA) * invdb + detCntIdx) + 1.0f;
    b.idxR = ((alpha + dA) * invdb + detCntIdx) + 1.0f;
    return b;
}

// Helical-to-fan detector row for channel table value SM
// (HelicalToFanFunc_mex.cpp:85-90).  Returns dv; the row is floor(dv).
DDF_HD float rebin_dv(float z, float zz, float SM, float SD, float DetCenterHreal, float PerDetH,
                      float* v_out) {
    const float tanAngV = (z - zz) / SM;
    const float v = SD * tanAngV;
    *v_out = v;
    // (float)((double)q + 1.0) equals the float sum q + 1.0f: q + 1 is exact
    // in double, so both round once from the same exact value.
    return (DetCenterHreal + v) / PerDetH + 1.0f;
}

// AAPM::mod<float> (HelicalToFanFunc_mex.cpp:27-34).
DDF_HD float mod_f(float lambda, float regV) {
    float v = lambda;
    while (v > regV) v -= regV;
    return v;
}

}  // namespace ddf
