"""Device-resident plan: geometry tables, summed-area-table buffers and the
OS-SART state of one reconstruction, kept in HBM across calls."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N
from .geometry import DetectorGeometry


def _sat_mode(m) -> int:
    if isinstance(m, str):
        return N.SAT_MODES[m]
    return int(m)


class Plan:
    """Wraps ``ddfan_plan``.  ``hangs`` is (PN, SLN) float32 (hangs[v, z])."""

    def __init__(self, geom: DetectorGeometry, hangs: np.ndarray, mask=None, sat_mode="f32",
                 slice_chunk: int = 0):
        self.geom = geom
        self.hangs = np.ascontiguousarray(hangs, np.float32)
        assert self.hangs.ndim == 2 and self.hangs.shape[1] == geom.SLN, self.hangs.shape
        self.PN = self.hangs.shape[0]
        self.mask = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        if self.mask is not None:
            assert self.mask.shape == (geom.YN, geom.XN), self.mask.shape
        self._gs = geom.as_struct()
        h = ctypes.c_void_p()
        N.call("ddfan_plan_create", ctypes.byref(self._gs), N.ptr(self.hangs), self.PN,
               N.ptr(self.mask), _sat_mode(sat_mode), int(slice_chunk), ctypes.byref(h))
        self.handle = h.value
        zt, nzc, mode = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        N.call("ddfan_plan_query", self.handle, ctypes.byref(zt), ctypes.byref(nzc),
               ctypes.byref(mode))
        self.slice_chunk, self.n_chunks, self.sat_mode = zt.value, nzc.value, mode.value

    # shapes
    @property
    def vol_shape(self):
        g = self.geom
        return (g.YN, g.XN, g.SLN)

    def sino_shape(self, count=None):
        g = self.geom
        return (self.PN if count is None else count, g.DNU, g.SLN)

    # operators on device buffers
    def proj(self, d_vol, d_prj, first=0, stride=1, count=None):
        count = self.PN if count is None else count
        N.call("ddfan_plan_proj", self.handle, N.ptr(d_vol), N.ptr(d_prj), first, stride, count)

    def back(self, d_prj, d_vol, first=0, stride=1, count=None):
        count = self.PN if count is None else count
        N.call("ddfan_plan_back", self.handle, N.ptr(d_prj), N.ptr(d_vol), first, stride, count)

    def sync(self):
        N.call("ddfan_plan_sync", self.handle)

    def ossart(self, d_meas, d_x, num_iter, num_os, use_fista=1, graph=True):
        N.call("ddfan_ossart", self.handle, N.ptr(d_meas), N.ptr(d_x), int(num_iter), int(num_os),
               int(use_fista), N.OSSART_GRAPH if graph else 0)

    # host-array conveniences
    def proj_host(self, vol, first=0, stride=1, count=None):
        count = self.PN if count is None else count
        v = np.ascontiguousarray(vol, np.float32)
        assert v.shape == self.vol_shape
        out = np.empty(self.sino_shape(count), np.float32)
        with N.DeviceBuffer.from_array(v) as dv, N.DeviceBuffer(out.nbytes) as dp:
            self.proj(dv, dp, first, stride, count)
            self.sync()
            dp.download(out)
        return out

    def back_host(self, prj, first=0, stride=1, count=None):
        count = self.PN if count is None else count
        p = np.ascontiguousarray(prj, np.float32)
        assert p.shape == self.sino_shape(count), (p.shape, self.sino_shape(count))
        out = np.empty(self.vol_shape, np.float32)
        with N.DeviceBuffer.from_array(p) as dp, N.DeviceBuffer(out.nbytes) as dv:
            self.back(dp, dv, first, stride, count)
            self.sync()
            dv.download(out)
        return out

    def ossart_host(self, meas, x0, num_iter, num_os, use_fista=1, graph=True):
        b = np.ascontiguousarray(meas, np.float32)
        assert b.shape == self.sino_shape()
        x = np.array(x0, np.float32, order="C")
        assert x.shape == self.vol_shape
        with N.DeviceBuffer.from_array(b) as db, N.DeviceBuffer.from_array(x) as dx:
            self.ossart(db, dx, num_iter, num_os, use_fista, graph)
            dx.download(x)
        return x

    # summed-area tables (tests, diagnostics)
    _ENTRY = {0: (np.float32, 2), 1: (np.float32, 4), 2: (np.float64, 2)}

    def vol_tables(self, vol):
        """The projector's volume tables of `vol` (YN, XN, SLN) as arrays
        satX[chunk, y, zt, k, e] (k = 0..XN) and satY[chunk, x, zt, k, e]
        (k = 0..YN), e the entry's fields (f32: S, v; two-float: S_hi, S_lo,
        v, 0; f64: S, v)."""
        g = self.geom
        bx, by = ctypes.c_size_t(), ctypes.c_size_t()
        N.call("ddfan_plan_vol_tables", self.handle, None, None, None, ctypes.byref(bx),
               ctypes.byref(by))
        dt, ne = self._ENTRY[self.sat_mode]
        zt = self.proj_chunk
        nzc = (g.SLN + zt - 1) // zt
        v = np.ascontiguousarray(vol, np.float32)
        with N.DeviceBuffer.from_array(v) as dv, N.DeviceBuffer(bx.value) as dX, \
                N.DeviceBuffer(by.value) as dY:
            N.call("ddfan_plan_vol_tables", self.handle, N.ptr(dv), N.ptr(dX), N.ptr(dY),
                   ctypes.byref(bx), ctypes.byref(by))
            X = dX.to_numpy((nzc, g.YN, zt, g.XN + 1, ne), dt)
            Y = dY.to_numpy((nzc, g.XN, zt, g.YN + 1, ne), dt)
        return X, Y

    def sino_table(self, prj):
        """The backprojector's table of a sinogram prj (count, DNU, SLN) as an
        array [chunk, view, zt, k, e] (k = 0..DNU)."""
        g = self.geom
        p = np.ascontiguousarray(prj, np.float32)
        count = p.shape[0]
        b = ctypes.c_size_t()
        dt, ne = self._ENTRY[self.sat_mode]
        zt = self.back_chunk
        nzc = (g.SLN + zt - 1) // zt
        with N.DeviceBuffer.from_array(p) as dp:
            N.call("ddfan_plan_sino_table", self.handle, N.ptr(dp), count, None, ctypes.byref(b))
            with N.DeviceBuffer(b.value) as dT:
                N.call("ddfan_plan_sino_table", self.handle, N.ptr(dp), count, N.ptr(dT),
                       ctypes.byref(b))
                return dT.to_numpy((nzc, count, zt, g.DNU + 1, ne), dt)

    # tunables
    def set_param(self, param, value):
        N.call("ddfan_plan_set_param", self.handle, int(param), int(value))

    def get_param(self, param):
        v = ctypes.c_int()
        N.call("ddfan_plan_get_param", self.handle, int(param), ctypes.byref(v))
        return v.value

    @property
    def slab_tile(self):
        return self.get_param(N.PARAM_SLAB_TILE)

    @slab_tile.setter
    def slab_tile(self, v):
        self.set_param(N.PARAM_SLAB_TILE, v)

    @property
    def views_per_wave(self):
        return self.get_param(N.PARAM_VIEWS_PER_WAVE)

    @views_per_wave.setter
    def views_per_wave(self, v):
        self.set_param(N.PARAM_VIEWS_PER_WAVE, v)

    @property
    def proj_chunk(self):
        return self.get_param(N.PARAM_PROJ_CHUNK)

    @proj_chunk.setter
    def proj_chunk(self, v):
        self.set_param(N.PARAM_PROJ_CHUNK, v)

    @property
    def back_chunk(self):
        return self.get_param(N.PARAM_BACK_CHUNK)

    @back_chunk.setter
    def back_chunk(self, v):
        self.set_param(N.PARAM_BACK_CHUNK, v)

    @property
    def proj_stage(self):
        """Projector kernel: 0 per-lane table reads, 1 table runs staged through
        LDS, 2 / 3 whole table rows in LDS with 4 / 2 slices per block (for
        spaced views; views_per_wave then counts tasks per wave: 9 or 12 in
        blocks of 4 waves, 5 or 6 in blocks of 8, 3 or 4 in blocks of 12)."""
        return self.get_param(N.PARAM_PROJ_STAGE)

    @proj_stage.setter
    def proj_stage(self, v):
        self.set_param(N.PARAM_PROJ_STAGE, v)

    @property
    def back_stage(self):
        """Views per LDS-staged backprojector group (0: per-lane table reads)."""
        return self.get_param(N.PARAM_BACK_STAGE)

    @back_stage.setter
    def back_stage(self, v):
        self.set_param(N.PARAM_BACK_STAGE, v)

    @property
    def back_pass(self):
        """Backprojector views per launch of a plain call (0: one launch)."""
        return self.get_param(N.PARAM_BACK_PASS)

    @back_pass.setter
    def back_pass(self, v):
        self.set_param(N.PARAM_BACK_PASS, v)

    @property
    def back_mid_range(self):
        """1 when the geometry lets the backprojector's footprint skip its
        per-lane operand-range check (read-only)."""
        return self.get_param(N.PARAM_BACK_MID_RANGE)

    def autotune(self, views=None, reps=3, use_cache=True) -> dict:
        """Times launch shapes for calls of `views` views (default PN), sets
        the fastest and records it (see tune_save / tune_load)."""
        prm = (ctypes.c_int * 6)()
        ms = (ctypes.c_double * 2)()
        N.call("ddfan_plan_autotune", self.handle, int(views or 0), int(reps),
               N.TUNE_USE_CACHE if use_cache else 0, prm, ms)
        return {"proj_chunk": prm[0], "views_per_wave": prm[1], "slab_tile": prm[2],
                "back_chunk": prm[3], "back_stage": prm[4], "proj_stage": prm[5],
                "proj_ms": ms[0], "back_ms": ms[1]}

    def apply_tuned(self, views=None) -> bool:
        found = ctypes.c_int()
        N.call("ddfan_plan_apply_tuned", self.handle, int(views or 0), ctypes.byref(found))
        return bool(found.value)

    # timing and traces
    def timing(self, enable=True):
        N.call("ddfan_plan_timing", self.handle, 1 if enable else 0)

    def timing_read(self, kind):
        """(total ms, launches) of kernel kind 0 volume SAT, 1 projector,
        2 sinogram SAT, 3 backprojector since timing was (re)enabled."""
        t, n = ctypes.c_double(), ctypes.c_int()
        N.call("ddfan_plan_timing_read", self.handle, int(kind), ctypes.byref(t), ctypes.byref(n))
        return t.value, n.value

    def trace_proj(self, first=0, stride=1, count=None):
        g = self.geom
        count = self.PN if count is None else count
        shape = (count, g.DNU, g.SLN, max(g.XN, g.YN), 2)
        out = np.zeros(shape, np.int32)
        with N.DeviceBuffer(out.nbytes) as d:
            d.upload(out)
            N.call("ddfan_plan_trace_proj", self.handle, first, stride, count, N.ptr(d))
            d.download(out)
        return out

    def trace_back(self, first=0, stride=1, count=None):
        g = self.geom
        count = self.PN if count is None else count
        out = np.zeros((g.YN, g.XN, g.SLN, count, 2), np.int32)
        with N.DeviceBuffer(out.nbytes) as d:
            d.upload(out)
            N.call("ddfan_plan_trace_back", self.handle, first, stride, count, N.ptr(d))
            d.download(out)
        return out

    def close(self):
        if getattr(self, "handle", None):
            N.call("ddfan_plan_destroy", self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            if getattr(self, "handle", None) and N._lib is not None:
                N._lib.ddfan_plan_destroy(self.handle)
                self.handle = None
        except Exception:  # pragma: no cover
            pass


def plan_footprint(geom: DetectorGeometry, PN, sat_mode="f32", slice_chunk=0, num_os=0) -> int:
    gs = geom.as_struct()
    b = ctypes.c_size_t()
    N.call("ddfan_plan_footprint", ctypes.byref(gs), int(PN), _sat_mode(sat_mode),
           int(slice_chunk), int(num_os), ctypes.byref(b))
    return b.value


def tune_save(path: str) -> None:
    """Writes the process-wide launch-shape table."""
    N.call("ddfan_tune_save", str(path).encode())


def tune_load(path: str) -> int:
    """Merges a saved launch-shape table; returns its entry count."""
    n = ctypes.c_int()
    N.call("ddfan_tune_load", str(path).encode(), ctypes.byref(n))
    return n.value


def tune_clear() -> None:
    N.call("ddfan_tune_clear")
