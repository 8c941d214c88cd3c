// ddfan_host.h -- host-side geometry tables of the DD fan-beam path.
//
// Everything here runs on the host in FP32 (or the reference's own FP64
// intermediate steps) with the platform libm, exactly as the reference
// computes it on the host, so the tables handed to the kernels are the ones
// the reference would hand to its kernels.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace ddf {

// DD2MutiSlices.m:61-76: equiangular arc detector centres.
void detector_arc(int DNU, float detCntIdx, float detCellWidth, float sdd, float sid, float* xds,
                  float* yds);

// utilities.cuh:143-148 regularizeAngle<float>.
float regularize_angle(float c);

// utilities.cuh:218-227: (cosf, sinf) of the regularised angle, interleaved.
void cossin_table(const float* hangs, size_t n, float* cs);

// MultiSlicesBackGPU.cu:154-157: the source position of each (view, slice),
// (x0, y0) rotated by its (cos, sin) pair from cossin_table, as interleaved
// (x, y) pairs.  The backprojectors read it instead of the angle.
void source_table(const float* cs, size_t n, float x0, float y0, float* src);

// 1 when, for every voxel centre of the XN x YN grid and every source on the
// circle through (x0, y0), the backprojector's squared voxel-source distance
// lies in [2^-18, 2^38] and its view-axis offset ddv (MultiSlicesBackGPU.cu:
// 166-170) in [2^-38, 2^18] in magnitude: inside the range where the short
// correctly rounded square root and quotients need no per-lane check, with a
// factor 4 to spare.  It holds when the source is farther from the centre
// than sqrt(2) times the grid's largest half-extent (any CT geometry whose
// source lies outside the grid) at millimetre-like scales; 0 otherwise.
int back_mid_range(int XN, int YN, float dx, float ocxB, float ocyB, float x0, float y0);

// utilities.cuh:268-287 DD3Boundaries<float>.
void dd_boundaries(int nb, const float* centres, float* bnd);

// MultiSlicesBackGPU.cu:211,226-227.
void back_scalars(float x0, float y0, const float* xds, const float* yds, float* S2D,
                  float* dbeta);

// reconscript.m:28-38.
void recon_mask(int XN, int YN, uint8_t* mask);

// OSSART_AAPM.m:60 (0-based): views k, k+numOS, ... < numViews.
int subset_views(int numViews, int numOS, int k, int* out);

// HelicalToFanFunc_mex.cpp:160-164.
struct RebinParams {
    float PerDetA, delta, DetCenterHreal, h, deltaZ;
};
RebinParams rebin_params(float SD, float SO, int DetHeight, float PerDetW, float PerDetH,
                         int DefTimes, float pitch);
// HelicalToFanFunc_mex.cpp:71.
int rebin_start(float z, float h, float deltaZ);
// HelicalToFanFunc_mex.cpp:84-85: SM[j-1] = SO*cosf((j - DetCenterW)*PerDetA).
void rebin_sm_table(int W, float SO, float DetCenterW, float PerDetA, float* SM);

// Largest slice chunk ZT in {16,8,4,2,1} such that every chunk of ZT
// consecutive slices has bit-identical angles in every view and ZT < 2*SLN
// (less than half of a chunk is padding).  The last chunk may be partly
// padding: the backprojector computes each voxel's detector footprint (about
// 100 VALU operations) once per view and chunk and then about 6 per slice, so
// 13 slices in one chunk of 16 take about 196 operations per voxel and view
// against 7 x 112 in chunks of 2, and 18 slices 2 x 196 in chunks of 16
// against 5 x 124 in chunks of 4.  limit caps ZT (0: none).
int choose_slice_chunk(const float* hangs, int PN, int SLN, int limit);
// The projector's default chunk: the widest ZT <= min(ZTmax, 8) whose padding
// (nzc*ZT - SLN) stays within an eighth of SLN (the projector's per-slice
// sums are most of its work, so padding costs it more), else 1.
int default_proj_chunk(int ZTmax, int SLN);

}  // namespace ddf
