"""MI355X-native distance-driven fan-beam reconstruction path.

Drop-in replacement for the native hot path of
Angus1996/AAPM_LowDoseCTImageRecon_competition: the multi-slice DD projector
and backprojector (MultiSlicesProjGPU.cu, MultiSlicesBackGPU.cu), the
helical-to-fan rebinning (HelicalToFanFunc_mex.cpp) and the OS-SART driver
(OSSART_AAPM.m), implemented as hand-written HIP kernels for gfx950 behind
the C ABI in include/ddfan.h (libddfan.so).
"""
from . import _native
from ._native import DDFanError, DeviceBuffer, device_count
from .batch import Batch
from .dd2 import dd2_multi_slices
from .dicom import (ImageCfg, ReconCfg, collect_image_cfg, collect_recon_cfg,
                    convert_to_recon_conf, read_proj, read_proj_frames)
from .geometry import Conf, detector_arc, detector_geometry, recon_mask, subset_views
from .ossart import ossart_aapm
from .plan import Plan, plan_footprint, tune_clear, tune_load, tune_save
from .rebin import HelicalCfg, helical_to_fan_func, helical_to_fan_routine
from .serve import Result, Server

__all__ = [
    "Batch", "Conf", "DDFanError", "DeviceBuffer", "HelicalCfg", "ImageCfg", "Plan", "ReconCfg",
    "collect_image_cfg", "collect_recon_cfg", "convert_to_recon_conf", "dd2_multi_slices",
    "read_proj", "read_proj_frames",
    "detector_arc", "detector_geometry", "device_count", "helical_to_fan_func",
    "helical_to_fan_routine", "ossart_aapm", "plan_footprint", "recon_mask", "Result", "Server",
    "subset_views", "tune_clear", "tune_load", "tune_save",
]


def native_library_path() -> str:
    return _native.LIB_PATH
