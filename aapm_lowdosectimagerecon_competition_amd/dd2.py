"""Mirror of the reference's operator wrapper ``DD2MutiSlices.m``.

``dd2_multi_slices('Proj'|'Back', conf, input, hangs, view_ind, mask,
startIdx)`` takes MATLAB-shaped arrays (any numpy order; they are laid out
column-major like MATLAB): volume (SLN, XN, YN), sinogram (SLN, DNU, PN),
angles (SLN, PN), mask (XN, YN).  ``view_ind`` is 1-based, as in MATLAB.
Results come back as Fortran-ordered arrays of the MATLAB shape.  The calls
go to the drop-in C entry points ``ddfan_proj`` / ``ddfan_back``.
"""
from __future__ import annotations

import math

import numpy as np

from . import _native as N
from .geometry import Conf, detector_geometry


def _colmajor(a, dtype):
    return np.asfortranarray(a, dtype=dtype)


def default_start_idx(SLN: int, num_gpu: int = 1) -> np.ndarray:
    """startIdx(k) = ceil((k-1)/numgpu*SLN) (DD2MutiSlices.m:95-101)."""
    return np.array([math.ceil((k - 1) / num_gpu * SLN) for k in range(1, num_gpu + 1)], np.int32)


def dd2_multi_slices(flag_fw: str, conf: Conf, inp, hangs, view_ind=None, mask=None,
                     startIdx=None):
    g = detector_geometry(conf)
    XN, YN, SLN, DNU = g.XN, g.YN, g.SLN, g.DNU
    h = _colmajor(hangs, np.float32)
    if h.ndim == 1:
        h = h.reshape(1, -1, order="F")
    slnc, viewNum = h.shape
    if slnc != SLN:
        raise ValueError(f"hangs has {slnc} slices, conf.SLN is {SLN}")
    if view_ind is None or len(view_ind) == 0:
        view_ind = np.arange(1, viewNum + 1)
    vi = np.asarray(view_ind, np.int64) - 1
    if vi.min() < 0 or vi.max() >= viewNum:
        raise IndexError("view_ind out of range (1-based)")
    hs = np.asfortranarray(h[:, vi])
    PN = int(vi.size)
    # the reference packs [mask, mask'] and reads the first XN*YN bytes
    m = None if mask is None else _colmajor(mask, np.uint8)
    if m is not None and m.shape != (XN, YN):
        raise ValueError(f"mask must be (XN, YN) = {(XN, YN)}")
    if startIdx is None:
        startIdx = default_start_idx(SLN, max(1, N.device_count()))
    st = np.ascontiguousarray(startIdx, np.int32)
    if flag_fw == "Proj":
        vol = _colmajor(inp, np.float32)
        if vol.shape != (SLN, XN, YN):
            raise ValueError(f"volume must be (SLN, XN, YN) = {(SLN, XN, YN)}")
        out = np.zeros((SLN, DNU, PN), np.float32, order="F")
        N.call("ddfan_proj", g.x0, g.y0, DNU, N.ptr(g.xds), N.ptr(g.yds), g.imgXCenter,
               g.imgYCenter, N.ptr(hs), PN, XN, YN, SLN, N.ptr(vol), N.ptr(out), g.dx, N.ptr(m),
               N.ptr(st), int(st.size))
        return out
    if flag_fw == "Back":
        prj = _colmajor(inp, np.float32)
        if prj.ndim == 2:
            prj = prj.reshape(prj.shape + (1,), order="F")
        prj = np.asfortranarray(prj[:, :, vi])
        if prj.shape != (SLN, DNU, PN):
            raise ValueError(f"sinogram must be (SLN, DNU, views) with SLN={SLN}, DNU={DNU}")
        out = np.zeros((SLN, XN, YN), np.float32, order="F")
        N.call("ddfan_back", g.x0, g.y0, DNU, N.ptr(g.xds), N.ptr(g.yds), g.detCntIdx,
               g.imgXCenter, g.imgYCenter, N.ptr(hs), PN, XN, YN, SLN, N.ptr(out), N.ptr(prj),
               g.dx, N.ptr(m), N.ptr(st), int(st.size))
        return out
    raise ValueError("flag_fw must be 'Proj' or 'Back'")
