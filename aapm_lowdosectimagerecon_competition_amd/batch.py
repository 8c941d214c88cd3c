"""Batched OS-SART: many reconstructions of uneven size share every kernel
launch (``ddfan_batch_*``), each batch iteration replayed as a hipGraph.

Per reconstruction the arithmetic is that of :meth:`Plan.ossart`, so a
reconstruction's result does not depend on what else is in the batch."""
from __future__ import annotations

import ctypes

from . import _native as N
from .plan import Plan, _sat_mode


class Batch:
    """Wraps ``ddfan_batch``.  Every plan must use ``sat_mode``; a plan holds
    one reconstruction's state and is added once."""

    def __init__(self, sat_mode="twofloat"):
        h = ctypes.c_void_p()
        N.call("ddfan_batch_create", _sat_mode(sat_mode), ctypes.byref(h))
        self.handle = h.value
        self._keep = []  # plans and buffers referenced by the batch

    def add(self, plan: Plan, d_meas, d_x, num_iter, num_os, use_fista=1) -> int:
        idx = ctypes.c_int()
        N.call("ddfan_batch_add", self.handle, plan.handle, N.ptr(d_meas), N.ptr(d_x),
               int(num_iter), int(num_os), int(use_fista), ctypes.byref(idx))
        self._keep.append((plan, d_meas, d_x))
        return idx.value

    def set_fill(self, proj_waves=None, back_blocks=None, solo_bytes=None):
        """Launch grouping: a reconstruction whose volume tables take at least
        solo_bytes runs its kernels back to back; the others share launches
        while one holds at most proj_waves projector waves / back_blocks
        backprojector blocks."""
        if solo_bytes is not None:
            N.call("ddfan_batch_set_param", self.handle, N.BATCH_SOLO_BYTES, int(solo_bytes))
        if proj_waves is not None:
            N.call("ddfan_batch_set_param", self.handle, N.BATCH_PROJ_FILL, int(proj_waves))
        if back_blocks is not None:
            N.call("ddfan_batch_set_param", self.handle, N.BATCH_BACK_FILL, int(back_blocks))

    def set_streams(self, n: int):
        """Split the reconstructions into `n` groups of about equal work that
        run side by side, each on its own stream (1: one group; 0: two)."""
        N.call("ddfan_batch_set_param", self.handle, N.BATCH_STREAMS, int(n))

    def tune(self, reps=2) -> dict:
        """Throughput launch shapes for the reconstructions that share launches
        (``ddfan_batch_tune``): one projector and one backprojector shape for
        all of them, timed on their shared launches.  Returns the launch time
        per subset round before and after, in ms."""
        ms = (ctypes.c_double * 4)()
        N.call("ddfan_batch_tune", self.handle, int(reps), ms)
        return {"proj_ms_before": ms[0], "back_ms_before": ms[1],
                "proj_ms": ms[2], "back_ms": ms[3]}

    def __len__(self):
        n = ctypes.c_int()
        N.call("ddfan_batch_size", self.handle, ctypes.byref(n))
        return n.value

    def run(self, graph=True):
        N.call("ddfan_batch_run", self.handle, N.OSSART_GRAPH if graph else 0)

    def close(self):
        if getattr(self, "handle", None):
            N.call("ddfan_batch_destroy", self.handle)
            self.handle = None
        self._keep = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            if getattr(self, "handle", None) and N._lib is not None:
                N._lib.ddfan_batch_destroy(self.handle)
                self.handle = None
        except Exception:  # pragma: no cover
            pass
