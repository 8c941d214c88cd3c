"""DICOM-CT-PD projection data and CT image headers: mirrors of readProj.m,
CollectReconCfg.m, CollectImageCfg.m and ConvertToReconConf.m over the native
reader in libddfan.so (``ddfan_ctpd_*``, ``ddfan_ct_image_read_header``).

The projection files of a scan are read in file-name order (MATLAB's ``dir``
sorts names); each holds one frame of NumberofDetectorColumns rows x
NumberofDetectorRows columns of unscaled 16-bit values, which is the rebin's
(channel, row) order.  readProj.m then reverses the view order
(readProj.m:52-53); ``read_proj`` returns that array, ``read_proj_frames``
the frames as stored for the device rebin to read in place
(``rebin.ORDER_FRAMES_REVERSED``).
"""
from __future__ import annotations

import ctypes
import glob
import math
import os
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from . import _native as N
from .geometry import Conf
from .rebin import HelicalCfg


class _CtpdHeader(ctypes.Structure):
    _fields_ = [
        ("HUCalibrationFactor", ctypes.c_double),
        ("DataCollectionDiameter", ctypes.c_double),
        ("SpiralPitchFactor", ctypes.c_double),
        ("NumberofDetectorRows", ctypes.c_int),
        ("NumberofDetectorColumns", ctypes.c_int),
        ("DetectorElementTransverseSpacing", ctypes.c_float),
        ("DetectorElementAxialSpacing", ctypes.c_float),
        ("DetectorShape", ctypes.c_char * 17),
        ("DetectorFocalCenterAngularPosition", ctypes.c_float),
        ("DetectorFocalCenterAxialPosition", ctypes.c_float),
        ("DetectorFocalCenterRadialDistance", ctypes.c_float),
        ("DetectorCentralElementX", ctypes.c_float),
        ("DetectorCentralElementY", ctypes.c_float),
        ("ConstantRadialDistance", ctypes.c_float),
        ("SourceAngularPositionShift", ctypes.c_float),
        ("SourceAxialPositionShift", ctypes.c_float),
        ("SourceRadialDistanceShift", ctypes.c_float),
        ("NumberofSourceAngularSteps", ctypes.c_int),
        ("Rows", ctypes.c_int),
        ("Columns", ctypes.c_int),
        ("PhotonStatisticsCount", ctypes.c_int),
    ]


class _ImageHeader(ctypes.Structure):
    _fields_ = [
        ("Width", ctypes.c_int), ("Height", ctypes.c_int), ("BitDepth", ctypes.c_int),
        ("SliceThickness", ctypes.c_double), ("KVP", ctypes.c_double),
        ("DataCollectionDiameter", ctypes.c_double), ("ReconstructionDiameter", ctypes.c_double),
        ("SingleCollimationWidth", ctypes.c_double), ("TotalCollimationWidth", ctypes.c_double),
        ("TableSpeed", ctypes.c_double), ("TableFeedPerRotation", ctypes.c_double),
        ("SpiralPitchFactor", ctypes.c_double),
        ("DataCollectionCenterPatient", ctypes.c_double * 3),
        ("ReconstructionTargetCenterPatient", ctypes.c_double * 3),
        ("PixelSpacing", ctypes.c_double * 2),
        ("SmallestImagePixelValue", ctypes.c_int), ("LargestImagePixelValue", ctypes.c_int),
        ("WindowCenter", ctypes.c_double * 2), ("WindowWidth", ctypes.c_double * 2),
        ("RescaleIntercept", ctypes.c_double), ("RescaleSlope", ctypes.c_double),
        ("PixelSize", ctypes.c_double),
    ]


@dataclass
class DetectorCentralElement:
    X: float
    Y: float


@dataclass
class ReconCfg:
    """The cfg of CollectReconCfg.m:5-92 (plus readProj.m's NumOfDataViews).
    Angles and distances as stored (DetectorFocalCenterAngularPosition is
    converted from degrees by HelicalToFan_routine.m:22, as the reference
    does)."""
    HUCalibrationFactor: float
    DataCollectionDiameter: float
    SpiralPitchFactor: float
    NumberofDetectorRows: int
    NumberofDetectorColumns: int
    DetectorElementTransverseSpacing: float
    DetectorElementAxialSpacing: float
    DetectorType: str
    DetectorFocalCenterAngularPosition: float
    DetectorFocalCenterAxialPosition: float
    DetectorFocalCenterRadialDistance: float
    DetectorCentralElement: DetectorCentralElement
    ConstantRadialDistance: float
    SourceAngularPositionShift: float
    SourceAxialPositionShift: float
    SourceRadialDistanceShift: float
    NumberofSourceAngularSteps: int
    PhotonStatistics: np.ndarray
    NumOfDataViews: int = 0
    frame_shape: Tuple[int, int] = (0, 0)  # (Rows, Columns) of the stored frames

    @property
    def SpiralPitchGE(self) -> float:
        """CollectReconCfg.m:88."""
        return self.SpiralPitchFactor * self.NumberofDetectorRows

    def helical(self) -> HelicalCfg:
        """The fields HelicalToFan_routine.m:20-33 reads."""
        return HelicalCfg(
            ConstantRadialDistance=self.ConstantRadialDistance,
            DetectorFocalCenterRadialDistance=self.DetectorFocalCenterRadialDistance,
            DetectorFocalCenterAngularPosition=self.DetectorFocalCenterAngularPosition,
            NumberofDetectorColumns=self.NumberofDetectorColumns,
            NumberofDetectorRows=self.NumberofDetectorRows,
            DetectorElementTransverseSpacing=self.DetectorElementTransverseSpacing,
            DetectorElementAxialSpacing=self.DetectorElementAxialSpacing,
            NumberofSourceAngularSteps=self.NumberofSourceAngularSteps,
            DetectorCentralElementX=self.DetectorCentralElement.X,
            SpiralPitchFactor=self.SpiralPitchFactor)


@dataclass
class ImageCfg:
    """The cfg of CollectImageCfg.m:3-31 (dicominfo field names)."""
    Width: int
    Height: int
    BitDepth: int
    SliceThickness: float
    KVP: float
    DataCollectionDiameter: float
    ReconstructionDiameter: float
    SingleCollimationWidth: float
    TotalCollimationWidth: float
    TableSpeed: float
    TableFeedPerRotation: float
    SpiralPitchFactor: float
    DataCollectionCenterPatient: Tuple[float, float, float]
    ReconstructionTargetCenterPatient: Tuple[float, float, float]
    Rows: int
    Columns: int
    PixelSpacing: Tuple[float, float]
    SmallestImagePixelValue: int
    LargestImagePixelValue: int
    WindowCenter: Tuple[float, ...]
    WindowWidth: Tuple[float, ...]
    RescaleIntercept: float
    RescaleSlope: float
    PixelSize: float


def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


def collect_recon_cfg(path) -> ReconCfg:
    """CollectReconCfg.m: the scan geometry of one projection file."""
    h = _CtpdHeader()
    N.call("ddfan_ctpd_read_header", _path(path), ctypes.byref(h), None, 0)
    stats = np.zeros(h.PhotonStatisticsCount, np.float32)
    if stats.size:
        N.call("ddfan_ctpd_read_header", _path(path), ctypes.byref(h), N.ptr(stats),
               int(stats.size))
    return ReconCfg(
        HUCalibrationFactor=h.HUCalibrationFactor,
        DataCollectionDiameter=h.DataCollectionDiameter,
        SpiralPitchFactor=h.SpiralPitchFactor,
        NumberofDetectorRows=h.NumberofDetectorRows,
        NumberofDetectorColumns=h.NumberofDetectorColumns,
        DetectorElementTransverseSpacing=h.DetectorElementTransverseSpacing,
        DetectorElementAxialSpacing=h.DetectorElementAxialSpacing,
        DetectorType=h.DetectorShape.decode(errors="replace"),
        DetectorFocalCenterAngularPosition=h.DetectorFocalCenterAngularPosition,
        DetectorFocalCenterAxialPosition=h.DetectorFocalCenterAxialPosition,
        DetectorFocalCenterRadialDistance=h.DetectorFocalCenterRadialDistance,
        DetectorCentralElement=DetectorCentralElement(h.DetectorCentralElementX,
                                                      h.DetectorCentralElementY),
        ConstantRadialDistance=h.ConstantRadialDistance,
        SourceAngularPositionShift=h.SourceAngularPositionShift,
        SourceAxialPositionShift=h.SourceAxialPositionShift,
        SourceRadialDistanceShift=h.SourceRadialDistanceShift,
        NumberofSourceAngularSteps=h.NumberofSourceAngularSteps,
        PhotonStatistics=stats.astype(np.float64),
        frame_shape=(h.Rows, h.Columns))


def projection_files(path) -> List[str]:
    """The *.dcm files of a scan directory in MATLAB ``dir`` order (names
    sorted by code point), readProj.m:33-34."""
    files = sorted(glob.glob(os.path.join(os.fspath(path), "*.dcm")), key=os.path.basename)
    if not files:
        raise N.DDFanError(N.DDFAN_EIO, "projection_files", f"no *.dcm files in {path}")
    return files


def read_proj_frames(path, threads: int = 0) -> Tuple[np.ndarray, ReconCfg]:
    """The projection frames of a scan directory as stored: uint16
    (views, NumberofDetectorColumns, NumberofDetectorRows) in file order, and
    the cfg of the first file.  The device rebin reads this array in place
    with ``raw_order=ORDER_FRAMES_REVERSED`` (readProj.m's flip included)."""
    files = projection_files(path)
    cfg = collect_recon_cfg(files[0])
    cfg.NumOfDataViews = len(files)
    W, H = cfg.NumberofDetectorColumns, cfg.NumberofDetectorRows
    frames = np.empty((len(files), W, H), np.uint16)
    arr = (ctypes.c_char_p * len(files))(*[_path(f) for f in files])
    N.call("ddfan_ctpd_read_frames", arr, len(files), W, H, N.ptr(frames), int(threads))
    return frames, cfg


def read_proj(path, threads: int = 0) -> Tuple[np.ndarray, ReconCfg]:
    """readProj.m: proj (NumberofDetectorRows, NumberofDetectorColumns, views)
    in MATLAB order, unscaled, views reversed (float32; the reference holds the
    same integers in double and passes single(proj) to the rebin), and cfg."""
    frames, cfg = read_proj_frames(path, threads)
    proj = frames[::-1].astype(np.float32)  # (views, W, H) C order, views reversed
    return proj.T, cfg  # (H, W, views), Fortran order: proj[(v*W + ch)*H + row]


def collect_image_cfg(path) -> ImageCfg:
    """CollectImageCfg.m: the reconstruction geometry of a CT image file."""
    h = _ImageHeader()
    N.call("ddfan_ct_image_read_header", _path(path), ctypes.byref(h))

    def pair(a):
        return tuple(x for x in (a[0], a[1]) if not math.isnan(x))

    return ImageCfg(
        Width=h.Width, Height=h.Height, BitDepth=h.BitDepth, SliceThickness=h.SliceThickness,
        KVP=h.KVP, DataCollectionDiameter=h.DataCollectionDiameter,
        ReconstructionDiameter=h.ReconstructionDiameter,
        SingleCollimationWidth=h.SingleCollimationWidth,
        TotalCollimationWidth=h.TotalCollimationWidth, TableSpeed=h.TableSpeed,
        TableFeedPerRotation=h.TableFeedPerRotation, SpiralPitchFactor=h.SpiralPitchFactor,
        DataCollectionCenterPatient=tuple(h.DataCollectionCenterPatient),
        ReconstructionTargetCenterPatient=tuple(h.ReconstructionTargetCenterPatient),
        Rows=h.Height, Columns=h.Width, PixelSpacing=tuple(h.PixelSpacing),
        SmallestImagePixelValue=h.SmallestImagePixelValue,
        LargestImagePixelValue=h.LargestImagePixelValue,
        WindowCenter=pair(h.WindowCenter), WindowWidth=pair(h.WindowWidth),
        RescaleIntercept=h.RescaleIntercept, RescaleSlope=h.RescaleSlope, PixelSize=h.PixelSize)


def convert_to_recon_conf(cfg: ReconCfg, cfgRecon: ImageCfg, SLN: int) -> Conf:
    """ConvertToReconConf.m:21-33."""
    return Conf.make(XN=cfgRecon.Width, YN=cfgRecon.Height, SLN=SLN, dx=cfgRecon.PixelSize,
                     sid=cfg.DetectorFocalCenterRadialDistance, sdd=cfg.ConstantRadialDistance,
                     DNU=cfg.NumberofDetectorColumns,
                     detCellWidth=cfg.DetectorElementTransverseSpacing,
                     detCntIdx=cfg.DetectorCentralElement.X,
                     PN=cfg.NumberofSourceAngularSteps)
