"""Reconstruction configuration and detector geometry.

``Conf`` carries the fields of the reference's ``conf`` struct
(ConvertToReconConf.m:23-33): ``recon.XN/YN/dx``, ``SLN`` and
``acq.sid/DNU/PN/sdd/detCellWidth/detCntIdx``.  ``detector_geometry`` is the
part of DD2MutiSlices.m:49-76 that turns it into projector arguments; the
transcendental arithmetic runs in the native library (platform libm, MATLAB
single semantics) so the tables are bit-identical to the host restatement.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _native as N


@dataclass
class Recon:
    XN: int
    YN: int
    dx: float


@dataclass
class Acq:
    sid: float           # source to isocentre (DetectorFocalCenterRadialDistance)
    sdd: float           # source to detector (ConstantRadialDistance)
    DNU: int             # detector channels
    detCellWidth: float  # channel spacing on the detector arc (mm)
    detCntIdx: float     # 1-based central channel (DetectorCentralElement.X)
    PN: int = 0          # views per rotation (informational)


@dataclass
class Conf:
    recon: Recon
    acq: Acq
    SLN: int
    imgXCenter: float = 0.0
    imgYCenter: float = 0.0

    @classmethod
    def make(cls, XN, YN, SLN, dx, sid, sdd, DNU, detCellWidth, detCntIdx, PN=0):
        return cls(Recon(int(XN), int(YN), float(dx)),
                   Acq(float(sid), float(sdd), int(DNU), float(detCellWidth), float(detCntIdx),
                       int(PN)), int(SLN))


@dataclass
class DetectorGeometry:
    """Projector arguments derived from a ``Conf`` (DD2MutiSlices.m:49-76)."""
    x0: float
    y0: float
    xds: np.ndarray
    yds: np.ndarray
    detCntIdx: float
    imgXCenter: float
    imgYCenter: float
    XN: int
    YN: int
    SLN: int
    dx: float
    _keep: list = field(default_factory=list, repr=False)

    @property
    def DNU(self) -> int:
        return int(self.xds.size)

    def as_struct(self) -> N.GeomStruct:
        xds = np.ascontiguousarray(self.xds, np.float32)
        yds = np.ascontiguousarray(self.yds, np.float32)
        self._keep[:] = [xds, yds]
        fp = ctypes.POINTER(ctypes.c_float)
        return N.GeomStruct(self.x0, self.y0, self.DNU, xds.ctypes.data_as(fp),
                            yds.ctypes.data_as(fp), self.detCntIdx, self.imgXCenter,
                            self.imgYCenter, self.XN, self.YN, self.SLN, self.dx)

    def as_dict(self) -> dict:
        return dict(x0=self.x0, y0=self.y0, xds=self.xds, yds=self.yds, dx=self.dx,
                    detCntIdx=self.detCntIdx, imgXCenter=self.imgXCenter,
                    imgYCenter=self.imgYCenter)

    def matched_backprojector(self) -> "DetectorGeometry":
        """The same geometry with the backprojector's channels aligned to the
        projector's.  The reference backprojector reads detector coordinate
        beta/dbeta + detCntIdx + 1 (MultiSlicesBackGPU.cu:175-176), one channel
        above the projector's channel of beta (DD2MutiSlices.m:71; the "should
        minus 1" note of CollectReconCfg.m:56-60).  Only the backprojector reads
        detCntIdx (the projector's channels come from xds/yds), so lowering it by
        one aligns the pair; OS-SART with small subsets then converges where the
        reference convention diverges.  Not the reference's arithmetic."""
        import dataclasses
        return dataclasses.replace(self, detCntIdx=float(np.float32(self.detCntIdx - 1.0)),
                                   _keep=[])


def detector_arc(DNU, detCntIdx, detCellWidth, sdd, sid):
    """xds, yds of DD2MutiSlices.m:61-76 (source at (0, sid), arc detector)."""
    xds = np.zeros(int(DNU), np.float32)
    yds = np.zeros(int(DNU), np.float32)
    N.call("ddfan_detector_arc", int(DNU), float(detCntIdx), float(detCellWidth), float(sdd),
           float(sid), N.ptr(xds), N.ptr(yds))
    return xds, yds


def detector_geometry(conf: Conf, SLN: int | None = None) -> DetectorGeometry:
    sid = np.float32(conf.acq.sid)
    xds, yds = detector_arc(conf.acq.DNU, conf.acq.detCntIdx, conf.acq.detCellWidth,
                            conf.acq.sdd, conf.acq.sid)
    return DetectorGeometry(0.0, float(sid), xds, yds, float(np.float32(conf.acq.detCntIdx)),
                            float(conf.imgXCenter), float(conf.imgYCenter), conf.recon.XN,
                            conf.recon.YN, conf.SLN if SLN is None else SLN,
                            float(np.float32(conf.recon.dx)))


def recon_mask(XN, YN) -> np.ndarray:
    """Circular mask of reconscript.m:28-38 as mask[y, x] (uint8)."""
    m = np.zeros((int(YN), int(XN)), np.uint8)
    N.call("ddfan_recon_mask", int(XN), int(YN), N.ptr(m))
    return m


def subset_views(numViews, numOS, k) -> np.ndarray:
    """0-based views of OS-SART subset k (OSSART_AAPM.m:60)."""
    cnt = ctypes.c_int(0)
    N.call("ddfan_subset_views", int(numViews), int(numOS), int(k), None, ctypes.byref(cnt))
    out = np.zeros(cnt.value, np.int32)
    N.call("ddfan_subset_views", int(numViews), int(numOS), int(k), N.ptr(out), ctypes.byref(cnt))
    return out


def arc_cell_width(step_theta: float, sdd: float) -> float:
    """Detector cell width on the arc for an angular channel pitch."""
    return float(np.float32(2.0 * sdd * np.tan(step_theta / 2.0)))
