"""Helical-to-fan rebinning: mirrors of ``HelicalToFan_routine.m`` and of the
``HelicalToFanFunc_mex`` call it makes (HelicalToFanFunc_mex.cpp:190-232).

Raw data follow the reference: proj (DetHeight, DetWidth, TotalViews) in
MATLAB order (proj[(view*W + ch)*H + row] in memory).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native as N


@dataclass
class HelicalCfg:
    """The cfg fields HelicalToFan_routine.m:20-33 reads."""
    ConstantRadialDistance: float             # SD
    DetectorFocalCenterRadialDistance: float  # SO
    DetectorFocalCenterAngularPosition: float # BVAngle, degrees in cfg
    NumberofDetectorColumns: int              # DetWidth
    NumberofDetectorRows: int                 # DetHeight
    DetectorElementTransverseSpacing: float   # PerDetW
    DetectorElementAxialSpacing: float        # PerDetH
    NumberofSourceAngularSteps: int           # DefTimes
    DetectorCentralElementX: float            # DetCenterW (1-based)
    SpiralPitchFactor: float


def helical_to_fan_func(proj, zPos, SLN, SD, SO, BVAngle, DetWidth, DetHeight, PerDetW, PerDetH,
                        DefTimes, DetCenterW, SpiralPitchFactor):
    """``[Proj, Views] = HelicalToFanFunc_mex(...)``: Proj (W, DefTimes, SLN)
    and Views (DefTimes, SLN), Fortran-ordered, before the routine's permute."""
    raw = np.asfortranarray(proj, np.float32)
    H, W, TV = raw.shape if raw.ndim == 3 else (raw.shape[0], raw.shape[1], 1)
    if H != DetHeight or W != DetWidth:
        raise ValueError("proj must be (DetHeight, DetWidth, TotalViews)")
    z = np.ascontiguousarray(zPos, np.float32)
    out = np.zeros((W, DefTimes, SLN), np.float32, order="F")
    views = np.zeros((DefTimes, SLN), np.float32, order="F")
    N.call("ddfan_helical_to_fan", N.ptr(out), N.ptr(views), N.ptr(raw), int(TV), N.ptr(z),
           int(SLN), float(SD), float(SO), float(BVAngle), int(DetWidth), int(DetHeight),
           float(PerDetW), float(PerDetH), int(DefTimes), float(DetCenterW),
           float(SpiralPitchFactor))
    return out, views


def helical_to_fan_routine(proj, cfg: HelicalCfg, zPos):
    """HelicalToFan_routine.m: returns Proj (SLN, W, DefTimes) and Views
    (SLN, DefTimes), the sinogram and angle layouts the operators take."""
    SD = np.float32(cfg.ConstantRadialDistance)
    SO = np.float32(cfg.DetectorFocalCenterRadialDistance)
    BVAngle = np.float32(cfg.DetectorFocalCenterAngularPosition / 180.0 * np.pi)
    zPos = np.asarray(zPos)
    Proj, Views = helical_to_fan_func(
        proj, zPos, zPos.size, SD, SO, BVAngle, cfg.NumberofDetectorColumns,
        cfg.NumberofDetectorRows, cfg.DetectorElementTransverseSpacing,
        cfg.DetectorElementAxialSpacing, cfg.NumberofSourceAngularSteps,
        cfg.DetectorCentralElementX, cfg.SpiralPitchFactor)
    return np.asfortranarray(np.transpose(Proj, (2, 0, 1))), np.asfortranarray(Views.T)


def slice_positions(cfg: HelicalCfg, total_views: int, SLN: int):
    """zPos = linspace(h, deltaZ*TotalView - h, SLN) (reconscript.m:18-22)."""
    h = (cfg.SpiralPitchFactor * cfg.DetectorElementAxialSpacing * cfg.NumberofDetectorRows *
         cfg.DetectorFocalCenterRadialDistance / cfg.ConstantRadialDistance)
    deltaZ = h / cfg.NumberofSourceAngularSteps
    return np.linspace(h, deltaZ * total_views - h, SLN)


def rebin_starts(zPos, SD, SO, DetHeight, PerDetW, PerDetH, DefTimes, pitch) -> np.ndarray:
    """First raw view (1-based) of every slice (HelicalToFanFunc_mex.cpp:71)."""
    z = np.ascontiguousarray(zPos, np.float32)
    out = np.zeros(z.size, np.int32)
    N.call("ddfan_rebin_starts", N.ptr(z), int(z.size), float(SD), float(SO), int(DetHeight),
           float(PerDetW), float(PerDetH), int(DefTimes), float(pitch), N.ptr(out))
    return out


RAW_FLOAT, RAW_COUNTS_U16, RAW_LINE_INTEGRALS_U16 = 0, 1, 2
ORDER_REBIN, ORDER_FRAMES_REVERSED = 0, 1


def helical_to_fan_device(d_raw, raw_kind, count_scale, layout, total_views, zPos, SD, SO,
                          BVAngle, DetWidth, DetHeight, PerDetW, PerDetH, DefTimes, DetCenterW,
                          pitch, d_out, d_views, d_cut=None, d_idv=None, raw_order=ORDER_REBIN):
    """Device-resident rebin (raw and outputs are device buffers).  raw_kind:
    RAW_FLOAT, RAW_COUNTS_U16 (log(count_scale/max(c,1))) or
    RAW_LINE_INTEGRALS_U16 (value*count_scale); raw_order: ORDER_REBIN
    (view, channel, row) or ORDER_FRAMES_REVERSED (the projection files'
    frames in file order, (channel, row) each, last view first: readProj.m
    flips them)."""
    z = np.ascontiguousarray(zPos, np.float32)
    N.call("ddfan_helical_to_fan_dev", N.ptr(d_out), N.ptr(d_views), N.ptr(d_raw), int(raw_kind),
           int(raw_order), float(count_scale), int(layout), int(total_views), N.ptr(z),
           int(z.size), float(SD),
           float(SO), float(BVAngle), int(DetWidth), int(DetHeight), float(PerDetW),
           float(PerDetH), int(DefTimes), float(DetCenterW), float(pitch), N.ptr(d_cut),
           N.ptr(d_idv))
