/*
 * ddfan_mex.c -- MATLAB MEX gateways over libddfan.so with the reference's
 * argument lists, so DD2MutiSlices.m and HelicalToFan_routine.m run unchanged.
 *
 * One source, three MEX files (choose the entry with DDFAN_MEX_ENTRY):
 *   1  MultiSlicesProjGPU     prhs as MultiSlicesProjGPU.cu:552-567
 *   2  MultiSlicesBackGPU     prhs as MultiSlicesBackGPU.cu:411-427
 *   3  HelicalToFanFunc_mex   prhs as HelicalToFanFunc_mex.cpp:186-232
 * e.g.  mex -DDDFAN_MEX_ENTRY=1 -I<repo>/include ddfan_mex.c -L<pkg> -lddfan
 *           -output MultiSlicesProjGPU
 *
 * Unlike the reference gateways these check argument count, class and size
 * before touching the data, and report library failures through
 * mexErrMsgIdAndTxt instead of exiting MATLAB.  startIdx (one slice offset
 * per GPU) is passed through with its length; the library validates it and
 * runs every shard on the current device.
 */
#include <stddef.h>
#include <stdio.h>

#include "ddfan.h"
#include "mex.h"

#ifndef DDFAN_MEX_ENTRY
#define DDFAN_MEX_ENTRY 1
#endif

static void fail_arg(const char* id, int k, const char* what) {
    char msg[256];
    snprintf(msg, sizeof msg, "argument %d: %s", k + 1, what);
    mexErrMsgIdAndTxt(id, "%s", msg);
}

static float get_single(const mxArray* a, const char* id, int k) {
    if (!mxIsSingle(a) || mxGetNumberOfElements(a) < 1) fail_arg(id, k, "expected a single scalar");
    return *(const float*)mxGetData(a);
}

static int get_int32(const mxArray* a, const char* id, int k) {
    if (!mxIsInt32(a) || mxGetNumberOfElements(a) < 1) fail_arg(id, k, "expected an int32 scalar");
    return *(const int*)mxGetData(a);
}

static const float* get_singles(const mxArray* a, size_t need, const char* id, int k) {
    if (!mxIsSingle(a)) fail_arg(id, k, "expected a single array");
    if (mxGetNumberOfElements(a) < need) fail_arg(id, k, "array has too few elements");
    return (const float*)mxGetData(a);
}

#if DDFAN_MEX_ENTRY != 3
/* uint8 or int8 mask of at least XN*YN bytes (the reference passes [mask, mask']). */
static const unsigned char* get_mask(const mxArray* a, size_t need, const char* id, int k) {
    if (!(mxIsUint8(a) || mxIsInt8(a))) fail_arg(id, k, "expected a uint8 or int8 mask");
    if (mxGetNumberOfElements(a) < need) fail_arg(id, k, "mask has fewer than XN*YN elements");
    return (const unsigned char*)mxGetData(a);
}

static const int* get_start(const mxArray* a, int* n, const char* id, int k) {
    if (!mxIsInt32(a)) fail_arg(id, k, "expected int32 startIdx");
    *n = (int)mxGetNumberOfElements(a);
    return *n > 0 ? (const int*)mxGetData(a) : NULL;
}
#endif

static void check_status(int st, const char* id) {
    if (st != DDFAN_OK) mexErrMsgIdAndTxt(id, "%s (status %d)", ddfan_last_error(), st);
}

#if DDFAN_MEX_ENTRY == 1
/* proj = MultiSlicesProjGPU(x0, y0, DNU, xds, yds, imgXCenter, imgYCenter,
 *                           hangs, PN, XN, YN, SLN, vol, dx, mask, startIdx) */
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    static const char* id = "ddfan:MultiSlicesProjGPU";
    if (nrhs != 16) mexErrMsgIdAndTxt(id, "expected 16 inputs, got %d", nrhs);
    if (nlhs > 1) mexErrMsgIdAndTxt(id, "one output");
    const float x0 = get_single(prhs[0], id, 0), y0 = get_single(prhs[1], id, 1);
    const int DNU = get_int32(prhs[2], id, 2);
    const float imgXCenter = get_single(prhs[5], id, 5), imgYCenter = get_single(prhs[6], id, 6);
    const int PN = get_int32(prhs[8], id, 8), XN = get_int32(prhs[9], id, 9);
    const int YN = get_int32(prhs[10], id, 10), SLN = get_int32(prhs[11], id, 11);
    if (DNU < 2 || PN < 0 || XN < 1 || YN < 1 || SLN < 1) mexErrMsgIdAndTxt(id, "bad sizes");
    const float* xds = get_singles(prhs[3], (size_t)DNU, id, 3);
    const float* yds = get_singles(prhs[4], (size_t)DNU, id, 4);
    const float* hangs = get_singles(prhs[7], (size_t)SLN * PN, id, 7);
    const float* vol = get_singles(prhs[12], (size_t)XN * YN * SLN, id, 12);
    const float dx = get_single(prhs[13], id, 13);
    const unsigned char* mask = get_mask(prhs[14], (size_t)XN * YN, id, 14);
    int nStart = 0;
    const int* startIdx = get_start(prhs[15], &nStart, id, 15);
    const mwSize dims[3] = {(mwSize)SLN, (mwSize)DNU, (mwSize)PN};
    plhs[0] = mxCreateNumericArray(3, dims, mxSINGLE_CLASS, mxREAL);
    float* prj = (float*)mxGetData(plhs[0]);
    check_status(ddfan_proj(x0, y0, DNU, xds, yds, imgXCenter, imgYCenter, hangs, PN, XN, YN, SLN,
                            vol, prj, dx, mask, startIdx, nStart),
                 id);
}
#elif DDFAN_MEX_ENTRY == 2
/* vol = MultiSlicesBackGPU(x0, y0, DNU, xds, yds, detCntIdx, imgXCenter, imgYCenter,
 *                          hangs, PN, XN, YN, SLN, proj, dx, mask, startIdx) */
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    static const char* id = "ddfan:MultiSlicesBackGPU";
    if (nrhs != 17) mexErrMsgIdAndTxt(id, "expected 17 inputs, got %d", nrhs);
    if (nlhs > 1) mexErrMsgIdAndTxt(id, "one output");
    const float x0 = get_single(prhs[0], id, 0), y0 = get_single(prhs[1], id, 1);
    const int DNU = get_int32(prhs[2], id, 2);
    const float detCntIdx = get_single(prhs[5], id, 5);
    const float imgXCenter = get_single(prhs[6], id, 6), imgYCenter = get_single(prhs[7], id, 7);
    const int PN = get_int32(prhs[9], id, 9), XN = get_int32(prhs[10], id, 10);
    const int YN = get_int32(prhs[11], id, 11), SLN = get_int32(prhs[12], id, 12);
    if (DNU < 2 || PN < 0 || XN < 1 || YN < 1 || SLN < 1) mexErrMsgIdAndTxt(id, "bad sizes");
    const float* xds = get_singles(prhs[3], (size_t)DNU, id, 3);
    const float* yds = get_singles(prhs[4], (size_t)DNU, id, 4);
    const float* hangs = get_singles(prhs[8], (size_t)SLN * PN, id, 8);
    const float* prj = get_singles(prhs[13], (size_t)SLN * DNU * PN, id, 13);
    const float dx = get_single(prhs[14], id, 14);
    const unsigned char* mask = get_mask(prhs[15], (size_t)XN * YN, id, 15);
    int nStart = 0;
    const int* startIdx = get_start(prhs[16], &nStart, id, 16);
    const mwSize dims[3] = {(mwSize)SLN, (mwSize)XN, (mwSize)YN};
    plhs[0] = mxCreateNumericArray(3, dims, mxSINGLE_CLASS, mxREAL);
    float* vol = (float*)mxGetData(plhs[0]);
    check_status(ddfan_back(x0, y0, DNU, xds, yds, detCntIdx, imgXCenter, imgYCenter, hangs, PN, XN,
                            YN, SLN, vol, prj, dx, mask, startIdx, nStart),
                 id);
}
#elif DDFAN_MEX_ENTRY == 3
/* [Proj, Views] = HelicalToFanFunc_mex(proj, zPos, SLN, SD, SO, BVAngle, DetWidth,
 *      DetHeight, PerDetW, PerDetH, DefTimes, DetCenterW, SpiralPitchFactor);
 * proj is single (DetHeight, DetWidth, TotalViews); TotalViews comes from its
 * size, so slices whose rotation window runs past the data are an error. */
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    static const char* id = "ddfan:HelicalToFanFunc_mex";
    if (nrhs != 13) mexErrMsgIdAndTxt(id, "expected 13 inputs, got %d", nrhs);
    if (nlhs > 2) mexErrMsgIdAndTxt(id, "two outputs");
    const int SLN = get_int32(prhs[2], id, 2);
    const float SD = get_single(prhs[3], id, 3), SO = get_single(prhs[4], id, 4);
    const float BVAngle = get_single(prhs[5], id, 5);
    const int W = get_int32(prhs[6], id, 6), H = get_int32(prhs[7], id, 7);
    const float PerDetW = get_single(prhs[8], id, 8), PerDetH = get_single(prhs[9], id, 9);
    const int DefTimes = get_int32(prhs[10], id, 10);
    const float DetCenterW = get_single(prhs[11], id, 11);
    const float pitch = get_single(prhs[12], id, 12);
    if (SLN < 0 || W < 1 || H < 1 || DefTimes < 1) mexErrMsgIdAndTxt(id, "bad sizes");
    if (!mxIsSingle(prhs[0])) fail_arg(id, 0, "expected single raw projections");
    const long long total = (long long)(mxGetNumberOfElements(prhs[0]) / ((size_t)W * H));
    const float* proj = (const float*)mxGetData(prhs[0]);
    const float* zPos = get_singles(prhs[1], (size_t)SLN, id, 1);
    const mwSize dp[3] = {(mwSize)W, (mwSize)DefTimes, (mwSize)SLN};
    const mwSize dv[2] = {(mwSize)DefTimes, (mwSize)SLN};
    plhs[0] = mxCreateNumericArray(3, dp, mxSINGLE_CLASS, mxREAL);
    plhs[1] = mxCreateNumericArray(2, dv, mxSINGLE_CLASS, mxREAL);
    check_status(ddfan_helical_to_fan((float*)mxGetData(plhs[0]), (float*)mxGetData(plhs[1]), proj,
                                      total, zPos, SLN, SD, SO, BVAngle, W, H, PerDetW, PerDetH,
                                      DefTimes, DetCenterW, pitch),
                 id);
}
#else
#error "DDFAN_MEX_ENTRY must be 1, 2 or 3"
#endif
