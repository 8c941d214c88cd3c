"""ctypes binding of the CPU restatement in ``ddfan_oracle.c``.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product package
never imports this module.

Array conventions (C order, matching the reference's MATLAB column-major
buffers read as flat memory):

* volume   ``vol[y, x, z]``   shape (YN, XN, SLN)  == vol[(y*XN+x)*SLN+z]
* sinogram ``prj[v, u, z]``   shape (PN, DNU, SLN) == prj[(v*DNU+u)*SLN+z]
* angles   ``hangs[v, z]``    shape (PN, SLN)      == hangs[v*SLN+z]
* mask     ``mask[y, x]``     shape (YN, XN) uint8 == mask[y*XN+x]
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libddfan_oracle.so")
_lib = None

_f = ctypes.c_float
_i = ctypes.c_int
_ll = ctypes.c_longlong
_p = ctypes.c_void_p


def build() -> str:
    """Compile the restatement with its Makefile (CPU only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def use_host_tuned_build() -> str:
    """Builds the restatement for the CPU this process runs on (gcc -O3
    -march=native; contraction and fast-math stay off, so every result is the
    same as the prebuilt x86-64-v3 library's) and makes lib() load it.  Call
    before the first lib().  Returns the -march used: "native", or
    "x86-64-v3" when the compiler is unavailable and the prebuilt library
    stays."""
    global _LIB_PATH
    if _lib is not None:
        raise RuntimeError("the restatement is already loaded")
    out = os.path.join(_HERE, "_build", "libddfan_oracle_native.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    cmd = ["gcc", "-O3", "-march=native", "-fPIC", "-std=c11", "-ffp-contract=off",
           "-fno-fast-math", "-shared", "-o", out, os.path.join(_HERE, "ddfan_oracle.c"),
           "-lm", "-lpthread"]
    try:
        subprocess.run(cmd, check=True, capture_output=True)
    except (OSError, subprocess.CalledProcessError):
        return "x86-64-v3"
    _LIB_PATH = out
    return "native"


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_asinf32.restype = _f
        L.orc_asinf32.argtypes = [_f]
        L.orc_regularize.restype = _f
        L.orc_regularize.argtypes = [_f]
        L.orc_mod_f.restype = _f
        L.orc_mod_f.argtypes = [_f, _f]
        L.orc_rebin_start.restype = _i
        L.orc_rebin_start.argtypes = [_f, _f, _f]
        for name in ("orc_proj_f64", "orc_proj_f32", "orc_back_f64", "orc_back_f32",
                     "orc_helical_to_fan", "orc_ossart_f64", "orc_ossart_f32",
                     "orc_subset_views"):
            getattr(L, name).restype = _i
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def nthreads_default() -> int:
    """Host threads for the restatement: the CPUs this process may run on,
    capped by OMP_NUM_THREADS when that is set (a container's CPU share)."""
    try:
        n = max(1, len(os.sched_getaffinity(0)))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    try:
        cap = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        cap = 0
    return min(n, cap) if cap > 0 else n


# ----------------------------------------------------------------- tables
def detector_arc(DNU, detCntIdx, detCellWidth, sdd, sid):
    xds = np.zeros(DNU, np.float32)
    yds = np.zeros(DNU, np.float32)
    lib().orc_detector_arc(_i(DNU), _f(detCntIdx), _f(detCellWidth), _f(sdd), _f(sid),
                           _ptr(xds), _ptr(yds))
    return xds, yds


def cossin(hangs):
    h = _c(hangs, np.float32)
    out = np.zeros(h.shape + (2,), np.float32)
    lib().orc_cossin(_ptr(h), ctypes.c_long(h.size), _ptr(out))
    return out


def boundaries(centers):
    c = _c(centers, np.float32)
    out = np.zeros(c.size + 1, np.float32)
    lib().orc_boundaries(_i(c.size + 1), _ptr(c), _ptr(out))
    return out


def regularize(a: float) -> float:
    return lib().orc_regularize(_f(a))


def asinf32(x: float) -> float:
    return lib().orc_asinf32(_f(x))


def recon_mask(XN, YN):
    m = np.zeros((YN, XN), np.uint8)
    lib().orc_recon_mask(_i(XN), _i(YN), _ptr(m))
    return m


def back_scalars(x0, y0, xds, yds):
    s2d = _f()
    db = _f()
    xs = _c(xds, np.float32)
    ys = _c(yds, np.float32)
    lib().orc_back_scalars(_f(x0), _f(y0), _ptr(xs), _ptr(ys), ctypes.byref(s2d), ctypes.byref(db))
    return s2d.value, db.value


def vol_sat(vol, mask, fp64=True):
    """(SAT_X[z, y, k], SAT_Y[z, x, k]) of the masked volume."""
    YN, XN, SLN = vol.shape
    m = _c(mask, np.uint8)
    if fp64:
        v = _c(vol, np.float64)
        SX = np.zeros((SLN, YN, XN + 1), np.float64)
        SY = np.zeros((SLN, XN, YN + 1), np.float64)
        lib().orc_vol_sat_f64(_ptr(v), _ptr(m), _i(XN), _i(YN), _i(SLN), _ptr(SX), _ptr(SY))
    else:
        v = _c(vol, np.float32)
        SX = np.zeros((SLN, YN, XN + 1), np.float32)
        SY = np.zeros((SLN, XN, YN + 1), np.float32)
        lib().orc_vol_sat_f32(_ptr(v), _ptr(m), _i(XN), _i(YN), _i(SLN), _ptr(SX), _ptr(SY))
    return SX, SY


# ----------------------------------------------------------------- operators
def proj(g, hangs, vol, mask, fp64=True, trace=False, nthreads=None):
    """DD projector.  ``g`` is a mapping with x0, y0, xds, yds, imgXCenter,
    imgYCenter, dx.  Returns prj[v, u, z] (and int32 texel indices
    [v, u, z, slab, {L,R}] when ``trace``)."""
    nthreads = nthreads or nthreads_default()
    h = _c(hangs, np.float32)
    PN, SLN = h.shape
    YN, XN, SLN2 = vol.shape
    assert SLN2 == SLN
    xds = _c(g["xds"], np.float32)
    yds = _c(g["yds"], np.float32)
    DNU = xds.size
    m = _c(mask, np.uint8)
    assert m.shape == (YN, XN)
    tr = np.zeros((PN, DNU, SLN, max(XN, YN), 2), np.int32) if trace else None
    if fp64:
        v = _c(vol, np.float64)
        out = np.zeros((PN, DNU, SLN), np.float64)
        rc = lib().orc_proj_f64(_f(g["x0"]), _f(g["y0"]), _i(DNU), _ptr(xds), _ptr(yds),
                                _f(g.get("imgXCenter", 0.0)), _f(g.get("imgYCenter", 0.0)),
                                _ptr(h), _i(PN), _i(XN), _i(YN), _i(SLN), _ptr(v), _ptr(out),
                                _f(g["dx"]), _ptr(m), _ptr(tr), _i(nthreads))
    else:
        assert not trace
        v = _c(vol, np.float32)
        out = np.zeros((PN, DNU, SLN), np.float32)
        rc = lib().orc_proj_f32(_f(g["x0"]), _f(g["y0"]), _i(DNU), _ptr(xds), _ptr(yds),
                                _f(g.get("imgXCenter", 0.0)), _f(g.get("imgYCenter", 0.0)),
                                _ptr(h), _i(PN), _i(XN), _i(YN), _i(SLN), _ptr(v), _ptr(out),
                                _f(g["dx"]), _ptr(m), _i(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle projector failed: {rc}")
    return (out, tr) if trace else out


def back(g, hangs, prj, mask, XN, YN, fp64=True, trace=False, nthreads=None):
    """DD backprojector.  Returns vol[y, x, z] (and int32 detector texel
    indices [y, x, z, v, {L,R}] when ``trace``)."""
    nthreads = nthreads or nthreads_default()
    h = _c(hangs, np.float32)
    PN, SLN = h.shape
    xds = _c(g["xds"], np.float32)
    yds = _c(g["yds"], np.float32)
    DNU = xds.size
    assert prj.shape == (PN, DNU, SLN), prj.shape
    m = _c(mask, np.uint8)
    tr = np.zeros((YN, XN, SLN, PN, 2), np.int32) if trace else None
    if fp64:
        p = _c(prj, np.float64)
        out = np.zeros((YN, XN, SLN), np.float64)
        rc = lib().orc_back_f64(_f(g["x0"]), _f(g["y0"]), _i(DNU), _ptr(xds), _ptr(yds),
                                _f(g["detCntIdx"]), _f(g.get("imgXCenter", 0.0)),
                                _f(g.get("imgYCenter", 0.0)), _ptr(h), _i(PN), _i(XN), _i(YN),
                                _i(SLN), _ptr(out), _ptr(p), _f(g["dx"]), _ptr(m), _ptr(tr),
                                _i(nthreads))
    else:
        assert not trace
        p = _c(prj, np.float32)
        out = np.zeros((YN, XN, SLN), np.float32)
        rc = lib().orc_back_f32(_f(g["x0"]), _f(g["y0"]), _i(DNU), _ptr(xds), _ptr(yds),
                                _f(g["detCntIdx"]), _f(g.get("imgXCenter", 0.0)),
                                _f(g.get("imgYCenter", 0.0)), _ptr(h), _i(PN), _i(XN), _i(YN),
                                _i(SLN), _ptr(out), _ptr(p), _f(g["dx"]), _ptr(m), _i(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle backprojector failed: {rc}")
    return (out, tr) if trace else out


def rebin_params(SD, SO, DetHeight, PerDetW, PerDetH, DefTimes, pitch):
    out = np.zeros(5, np.float32)
    lib().orc_rebin_params(_f(SD), _f(SO), _i(DetHeight), _f(PerDetW), _f(PerDetH), _i(DefTimes),
                           _f(pitch), _ptr(out))
    return dict(PerDetA=out[0], delta=out[1], DetCenterHreal=out[2], h=out[3], deltaZ=out[4])


def helical_to_fan(raw, zPos, SD, SO, BVAngle, PerDetW, PerDetH, DefTimes, DetCenterW, pitch,
                   fp64=True, nthreads=None):
    """raw[view, ch, row] (TotalViews, W, H) float32 -> dict with Proj[s, v, j],
    Views[s, v], cut[s, v] uint8, idv[s, v, j] int32, start[s] int32."""
    nthreads = nthreads or nthreads_default()
    r = _c(raw, np.float32)
    TV, W, H = r.shape
    z = _c(zPos, np.float32)
    SLN = z.size
    Proj = np.zeros((SLN, DefTimes, W), np.float32)
    Views = np.zeros((SLN, DefTimes), np.float32)
    cut = np.zeros((SLN, DefTimes), np.uint8)
    idv = np.zeros((SLN, DefTimes, W), np.int32)
    start = np.zeros(SLN, np.int32)
    rc = lib().orc_helical_to_fan(_ptr(Proj), _ptr(Views), _ptr(cut), _ptr(idv), _ptr(start),
                                  _ptr(r), _ll(TV), _ptr(z), _i(SLN), _f(SD), _f(SO), _f(BVAngle),
                                  _i(W), _i(H), _f(PerDetW), _f(PerDetH), _i(DefTimes),
                                  _f(DetCenterW), _f(pitch), _i(1 if fp64 else 0), _i(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle rebin failed: {rc}")
    return dict(Proj=Proj, Views=Views, cut=cut, idv=idv, start=start)


def subset_views(numViews, numOS, k):
    out = np.zeros(numViews, np.int32)
    n = lib().orc_subset_views(_i(numViews), _i(numOS), _i(k), _ptr(out))
    return out[:n].copy()


def ossart(g, hangs, meas, mask, init, numIter, numOS, useFISTA=1, fp64=True, nthreads=None):
    """OS-SART(+FISTA) with the restated operators; returns x[y, x, z]."""
    nthreads = nthreads or nthreads_default()
    h = _c(hangs, np.float32)
    PN, SLN = h.shape
    xds = _c(g["xds"], np.float32)
    yds = _c(g["yds"], np.float32)
    DNU = xds.size
    YN, XN, _ = init.shape
    b = _c(meas, np.float32)
    m = _c(mask, np.uint8)
    args_geo = (_f(g["x0"]), _f(g["y0"]), _i(DNU), _ptr(xds), _ptr(yds), _f(g["detCntIdx"]),
                _f(g.get("imgXCenter", 0.0)), _f(g.get("imgYCenter", 0.0)), _ptr(h), _i(PN),
                _i(XN), _i(YN), _i(SLN), _ptr(b))
    if fp64:
        x = np.array(init, dtype=np.float64, order="C")
        rc = lib().orc_ossart_f64(*args_geo, _ptr(x), _f(g["dx"]), _ptr(m), _i(numIter), _i(numOS),
                                  _i(useFISTA), _i(nthreads))
    else:
        x = np.array(init, dtype=np.float32, order="C")
        rc = lib().orc_ossart_f32(*args_geo, _ptr(x), _f(g["dx"]), _ptr(m), _i(numIter), _i(numOS),
                                  _i(useFISTA), _i(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle OS-SART failed: {rc}")
    return x
