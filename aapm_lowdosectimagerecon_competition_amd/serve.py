"""Reconstruction request loop (``ddfan_server_*``): requests are admitted in
arrival order by HBM budget and slot count, keep their sinogram and iterate
in device memory, and run together as a continuously re-formed batch that
requests join and leave at OS-SART iteration boundaries."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .geometry import DetectorGeometry
from .plan import _sat_mode


@dataclass
class Result:
    id: int
    x: np.ndarray       # (YN, XN, SLN) reconstruction
    queue_ms: float     # submission -> admission
    latency_ms: float   # submission -> result on the host


class Server:
    """Wraps ``ddfan_server``.  ``hbm_budget`` 0 takes 80% of the free device
    memory."""

    def __init__(self, hbm_budget: int = 0, max_slots: int = 16, sat_mode="twofloat",
                 graphs: bool = False, policy: str = "fifo", max_wait_ms=None):
        """``graphs``: replay each request set's batch iteration as a captured
        hipGraph instead of launching its kernels directly (same results).
        ``policy``: "fifo" (arrival order) or "shortest" (least work first,
        unless a request has waited ``max_wait_ms``, default 10 s)."""
        h = ctypes.c_void_p()
        N.call("ddfan_server_create", ctypes.c_size_t(int(hbm_budget)), int(max_slots),
               _sat_mode(sat_mode), ctypes.byref(h))
        self.handle = h.value
        self._shapes = {}
        if graphs:
            N.call("ddfan_server_set_param", self.handle, N.SERVER_GRAPH, 1)
        pol = {"fifo": N.POLICY_FIFO, "shortest": N.POLICY_SHORTEST}[policy]
        if max_wait_ms is not None:
            N.call("ddfan_server_set_param", self.handle, N.SERVER_MAX_WAIT_MS, int(max_wait_ms))
        N.call("ddfan_server_set_param", self.handle, N.SERVER_POLICY, pol)

    def submit(self, geom: DetectorGeometry, hangs, meas, num_iter, num_os, use_fista=1,
               mask=None, x0=None) -> int:
        """Queues one reconstruction; the inputs are copied (to the device when
        it is admitted at once, else to host memory) before return.
        hangs (PN, SLN), meas (PN, DNU, SLN), x0 (YN, XN, SLN) or None."""
        h = np.ascontiguousarray(hangs, np.float32)
        PN = h.shape[0]
        assert h.shape == (PN, geom.SLN), h.shape
        b = np.ascontiguousarray(meas, np.float32)
        assert b.shape == (PN, geom.DNU, geom.SLN), b.shape
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        if m is not None:
            assert m.shape == (geom.YN, geom.XN)
        x = None if x0 is None else np.ascontiguousarray(x0, np.float32)
        if x is not None:
            assert x.shape == (geom.YN, geom.XN, geom.SLN)
        gs = geom.as_struct()
        rid = ctypes.c_longlong()
        N.call("ddfan_server_submit", self.handle, ctypes.byref(gs), N.ptr(h), PN, N.ptr(m),
               N.ptr(b), N.ptr(x), int(num_iter), int(num_os), int(use_fista), ctypes.byref(rid))
        self._shapes[rid.value] = (geom.YN, geom.XN, geom.SLN)
        return rid.value

    def done(self, rid: int) -> bool:
        d = ctypes.c_int()
        N.call("ddfan_server_poll", self.handle, ctypes.c_longlong(rid), ctypes.byref(d))
        return bool(d.value)

    def wait(self, rid: int) -> Result:
        shape = self._shapes.pop(rid)
        x = np.empty(shape, np.float32)
        q, lat = ctypes.c_double(), ctypes.c_double()
        N.call("ddfan_server_wait", self.handle, ctypes.c_longlong(rid), N.ptr(x),
               ctypes.byref(q), ctypes.byref(lat))
        return Result(rid, x, q.value, lat.value)

    def stats(self) -> dict:
        c, f, w = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        r, u, vv = ctypes.c_int(), ctypes.c_size_t(), ctypes.c_double()
        N.call("ddfan_server_stats", self.handle, ctypes.byref(c), ctypes.byref(f),
               ctypes.byref(r), ctypes.byref(w), ctypes.byref(u), ctypes.byref(vv))
        return {"completed": c.value, "failed": f.value, "running": r.value,
                "waiting": w.value, "used_bytes": u.value, "voxel_views": vv.value}

    def close(self):
        if getattr(self, "handle", None):
            N.call("ddfan_server_destroy", self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            if getattr(self, "handle", None) and N._lib is not None:
                N._lib.ddfan_server_destroy(self.handle)
                self.handle = None
        except Exception:  # pragma: no cover
            pass
