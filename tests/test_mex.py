"""MEX gateways (aapm_lowdosectimagerecon_competition_amd/mex/ddfan_mex.c) built
against a stand-in mex.h (tests/native/mex_stub) and driven through ctypes
with MATLAB-shaped arrays, in the reference's argument order
(MultiSlicesProjGPU.cu:552-567, MultiSlicesBackGPU.cu:411-427,
HelicalToFanFunc_mex.cpp:186-232).

CPU: argument count, class and size checks raise MATLAB errors; with no GPU
the library's status reaches MATLAB as an error.  GPU: gateway outputs equal
the Python mirror (dd2_multi_slices, helical_to_fan_func) bit for bit."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from helpers import small_conf

SRC = os.path.join(ROOT, "aapm_lowdosectimagerecon_competition_amd", "mex", "ddfan_mex.c")
STUB = os.path.join(ROOT, "tests", "native", "mex_stub")
OUT = os.path.join(ROOT, "tests", "native", "_build")
PKG = os.path.join(ROOT, "aapm_lowdosectimagerecon_competition_amd")
ENTRY = {"proj": 1, "back": 2, "rebin": 3}
SINGLE, INT32, UINT8 = 7, 12, 9


def build_gateway(which: str) -> str:
    """Compiles one gateway with the stub API into tests/native/_build."""
    os.makedirs(OUT, exist_ok=True)
    so = os.path.join(OUT, f"mex_{which}.so")
    srcs = [SRC, os.path.join(STUB, "mex_stub.c"), os.path.join(STUB, "mex.h")]
    if os.path.exists(so) and os.path.getmtime(so) >= max(os.path.getmtime(s) for s in srcs):
        return so
    cmd = ["gcc", "-O2", "-std=c11", "-Wall", "-Werror", "-fPIC", "-shared",
           f"-DDDFAN_MEX_ENTRY={ENTRY[which]}", f"-I{STUB}", f"-I{os.path.join(ROOT, 'include')}",
           SRC, os.path.join(STUB, "mex_stub.c"), f"-L{PKG}", "-lddfan",
           "-Wl,-rpath,$ORIGIN/../../../aapm_lowdosectimagerecon_competition_amd", "-o", so]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return so


class Gateway:
    def __init__(self, which):
        from aapm_lowdosectimagerecon_competition_amd import _native as N
        N.lib()  # the gateway links the same libddfan.so
        self.L = ctypes.CDLL(build_gateway(which))
        self.L.stub_array.restype = ctypes.c_void_p
        self.L.stub_array.argtypes = [ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_size_t), ctypes.c_void_p]
        self.L.stub_free.argtypes = [ctypes.c_void_p]
        self.L.stub_data.restype = ctypes.c_void_p
        self.L.stub_data.argtypes = [ctypes.c_void_p]
        self.L.stub_dims.restype = ctypes.c_size_t
        self.L.stub_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
        self.L.stub_call.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p,
                                     ctypes.c_size_t]
        self.keep = []

    def arr(self, a, cls):
        dt = {SINGLE: np.float32, INT32: np.int32, UINT8: np.uint8}[cls]
        a = np.asfortranarray(np.asarray(a, dt))
        dims = (ctypes.c_size_t * max(a.ndim, 2))(*(a.shape if a.ndim >= 2 else (a.size, 1)))
        p = self.L.stub_array(cls, max(a.ndim, 2), dims, a.ctypes.data_as(ctypes.c_void_p))
        self.keep.append(p)
        return p

    def single(self, v):
        return self.arr(np.array([v], np.float32), SINGLE)

    def int32(self, v):
        return self.arr(np.array([v], np.int32), INT32)

    def call(self, args, nlhs=1):
        prhs = (ctypes.c_void_p * len(args))(*args)
        plhs = (ctypes.c_void_p * 2)()
        msg = ctypes.create_string_buffer(1024)
        rc = self.L.stub_call(nlhs, plhs, len(args), prhs, msg, 1024)
        return rc, msg.value.decode(), [plhs[i] for i in range(nlhs)]

    def read(self, p, dtype=np.float32):
        dims = (ctypes.c_size_t * 8)()
        nd = self.L.stub_dims(p, dims)
        shape = tuple(dims[i] for i in range(nd))
        n = int(np.prod(shape))
        buf = ctypes.cast(self.L.stub_data(p), ctypes.POINTER(ctypes.c_float))
        return np.ctypeslib.as_array(buf, (n,)).astype(dtype).reshape(shape, order="F")

    def close(self):
        for p in self.keep:
            self.L.stub_free(p)
        self.keep = []


def _case(SLN=2, PN=12, XN=24, YN=20):
    import aapm_lowdosectimagerecon_competition_amd as dd
    from aapm_lowdosectimagerecon_competition_amd import phantoms
    conf = small_conf(XN=XN, YN=YN, SLN=SLN)
    g = dd.detector_geometry(conf)
    hangs = phantoms.uniform_angles(PN, SLN)
    vol = phantoms.ellipsoid_phantom(XN, YN, SLN, conf.recon.dx, seed=2)
    mask = dd.recon_mask(XN, YN)
    return dd, conf, g, hangs, vol, mask


def _proj_args(G, g, hangs_m, vol_m, mask_m, start):
    return [G.single(g.x0), G.single(g.y0), G.int32(g.DNU), G.arr(g.xds, SINGLE),
            G.arr(g.yds, SINGLE), G.single(g.imgXCenter), G.single(g.imgYCenter),
            G.arr(hangs_m, SINGLE), G.int32(hangs_m.shape[1]), G.int32(g.XN), G.int32(g.YN),
            G.int32(g.SLN), G.arr(vol_m, SINGLE), G.single(g.dx), G.arr(mask_m, UINT8),
            G.arr(start, INT32)]


def _back_args(G, g, hangs_m, prj_m, mask_m, start):
    return [G.single(g.x0), G.single(g.y0), G.int32(g.DNU), G.arr(g.xds, SINGLE),
            G.arr(g.yds, SINGLE), G.single(g.detCntIdx), G.single(g.imgXCenter),
            G.single(g.imgYCenter), G.arr(hangs_m, SINGLE), G.int32(hangs_m.shape[1]),
            G.int32(g.XN), G.int32(g.YN), G.int32(g.SLN), G.arr(prj_m, SINGLE), G.single(g.dx),
            G.arr(mask_m, UINT8), G.arr(start, INT32)]


def test_gateways_check_arguments():
    dd, conf, g, hangs, vol, mask = _case()
    G = Gateway("proj")
    try:
        args = _proj_args(G, g, hangs.T, vol.T, mask.T, np.array([0], np.int32))
        rc, msg, _ = G.call(args[:15])
        assert rc == 1 and "expected 16 inputs" in msg
        bad = list(args)
        bad[0] = G.arr(np.array([1.0]), INT32)  # x0 must be single
        rc, msg, _ = G.call(bad)
        assert rc == 1 and "argument 1" in msg
        bad = list(args)
        bad[14] = G.arr(np.ones((3, 3), np.uint8), UINT8)  # mask smaller than XN*YN
        rc, msg, _ = G.call(bad)
        assert rc == 1 and "argument 15" in msg
        bad = list(args)
        bad[12] = G.arr(np.zeros(5, np.float32), SINGLE)  # volume too small
        rc, msg, _ = G.call(bad)
        assert rc == 1 and "argument 13" in msg
        bad = list(args)
        bad[15] = G.arr(np.array([1], np.int32), INT32)  # startIdx[0] must be 0
        rc, msg, _ = G.call(bad)
        assert rc == 1 and "startIdx" in msg
        if dd.device_count() == 0:  # the library's status reaches MATLAB
            rc, msg, _ = G.call(args)
            assert rc == 1 and "status -5" in msg
    finally:
        G.close()
    for which, n in (("back", 17), ("rebin", 13)):
        Gw = Gateway(which)
        try:
            rc, msg, _ = Gw.call([Gw.single(0.0)] * (n - 1), nlhs=2 if which == "rebin" else 1)
            assert rc == 1 and f"expected {n} inputs" in msg
        finally:
            Gw.close()


@pytest.mark.gpu
def test_gateways_match_python_mirror():
    dd, conf, g, hangs, vol, mask = _case(SLN=3, PN=16)
    start = np.array([0], np.int32)
    hangs_m, vol_m, mask_m = hangs.T, vol.T, mask.T  # MATLAB (SLN, PN), (SLN, XN, YN), (XN, YN)
    print("\n[mex] projector / backprojector gateways vs dd2_multi_slices", flush=True)
    ref_p = dd.dd2_multi_slices("Proj", conf, vol_m, hangs_m, None, mask_m, start)
    ref_b = dd.dd2_multi_slices("Back", conf, ref_p, hangs_m, None, mask_m, start)
    G = Gateway("proj")
    try:
        rc, msg, out = G.call(_proj_args(G, g, hangs_m, vol_m, mask_m, start))
        assert rc == 0, msg
        assert np.array_equal(G.read(out[0]), ref_p)
    finally:
        G.close()
    G = Gateway("back")
    try:
        rc, msg, out = G.call(_back_args(G, g, hangs_m, ref_p, mask_m, start))
        assert rc == 0, msg
        assert np.array_equal(G.read(out[0]), ref_b)
    finally:
        G.close()
    print("[mex] rebin gateway vs helical_to_fan_func", flush=True)
    from aapm_lowdosectimagerecon_competition_amd import rebin
    H, W, DefTimes, SLN, TV = 8, 16, 32, 4, 160
    rng = np.random.default_rng(3)
    raw = np.asfortranarray(rng.uniform(0, 2, (H, W, TV)).astype(np.float32))
    SD, SO, PerDetW, PerDetH, pitch, cw = 1085.6, 595.0, 1.2858, 1.0947, 0.8, 8.5
    h = pitch * PerDetH * H * SO / SD
    zPos = np.linspace(h, h / DefTimes * TV - h, SLN).astype(np.float32)
    ref_P, ref_V = rebin.helical_to_fan_func(raw, zPos, SLN, SD, SO, 0.1, W, H, PerDetW, PerDetH,
                                             DefTimes, cw, pitch)
    G = Gateway("rebin")
    try:
        args = [G.arr(raw, SINGLE), G.arr(zPos, SINGLE), G.int32(SLN), G.single(SD),
                G.single(SO), G.single(0.1), G.int32(W), G.int32(H), G.single(PerDetW),
                G.single(PerDetH), G.int32(DefTimes), G.single(cw), G.single(pitch)]
        rc, msg, out = G.call(args, nlhs=2)
        assert rc == 0, msg
        assert np.array_equal(G.read(out[0]), ref_P)
        assert np.array_equal(G.read(out[1]), ref_V)
    finally:
        G.close()
