"""Mirror of the reference's ``OSSART_AAPM.m`` (OS-SART with FISTA).

``ossart_aapm(Proj, Views, initImg, conf, numOfIter, numOfOS, mask,
useFISTA)`` takes MATLAB-shaped arrays: Proj (SLN, DNU, numViews), Views
(SLN, numViews), initImg (SLN, XN, YN), mask (XN, YN); returns the
reconstruction (SLN, XN, YN), Fortran-ordered.

* ``mode="device"`` (default): the whole iteration runs on the GPU with the
  sinogram and iterate resident, ratio/normalise/update/FISTA fused into the
  operators, and one iteration replayed as a hipGraph (``ddfan_ossart``).
* ``mode="plain"``: the reference's form, one host-buffer projector and
  backprojector call per subset with the elementwise steps on the host
  (``ddfan_ossart_host_plain``).
"""
from __future__ import annotations

import numpy as np

from . import _native as N
from .geometry import Conf, detector_geometry
from .plan import Plan, _sat_mode


def ossart_aapm(Proj, Views, initImg, conf: Conf, numOfIter: int, numOfOS: int = 1, mask=None,
                useFISTA: int = 1, mode: str = "device", sat_mode="twofloat", graph: bool = True):
    g = detector_geometry(conf)
    P = np.asfortranarray(Proj, np.float32)
    SLN, DNU, numViews = P.shape
    if SLN != g.SLN or DNU != g.DNU:
        raise ValueError("Proj must be (SLN, DNU, numViews) of conf")
    V = np.asfortranarray(Views, np.float32).reshape(SLN, numViews, order="F")
    x = np.array(initImg, np.float32, order="F").reshape(SLN, g.XN, g.YN, order="F")
    m = None if mask is None else np.asfortranarray(mask, np.uint8)
    if numOfOS < 1 or numOfOS > numViews:
        raise ValueError("numOfOS must be in [1, numViews]")
    if mode == "plain":
        N.call("ddfan_ossart_host_plain", g.x0, g.y0, DNU, N.ptr(g.xds), N.ptr(g.yds), g.detCntIdx,
               g.imgXCenter, g.imgYCenter, N.ptr(V), numViews, g.XN, g.YN, SLN, N.ptr(P), N.ptr(x),
               g.dx, N.ptr(m), int(numOfIter), int(numOfOS), int(useFISTA), _sat_mode(sat_mode))
        return x
    if mode != "device":
        raise ValueError("mode must be 'device' or 'plain'")
    # column-major (SLN, ...) buffers are the C-order (..., SLN) arrays of the plan
    hangs_c = V.T  # (numViews, SLN) C-contiguous view of the same memory
    mask_c = None if m is None else m.T
    with Plan(g, hangs_c, mask_c, sat_mode=sat_mode) as plan:
        xc = plan.ossart_host(P.T, x.T, numOfIter, numOfOS, useFISTA, graph)
    return np.asfortranarray(xc.T)
