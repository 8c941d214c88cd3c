"""ctypes binding of ``libddfan.so`` (the C ABI declared in include/ddfan.h).

The library is the product: there is no Python or CPU fallback.  Loading
fails loudly when the shared object is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libddfan.so")

_f = ctypes.c_float
_i = ctypes.c_int
_ll = ctypes.c_longlong
_sz = ctypes.c_size_t
_p = ctypes.c_void_p
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_ubyte)

DDFAN_OK = 0
DDFAN_EINVAL = -1
DDFAN_ENOMEM = -2
DDFAN_EBOUNDS = -3
DDFAN_EDEVICE = -4
DDFAN_ENODEV = -5
DDFAN_EBUSY = -6
DDFAN_EIO = -7

SAT_F32 = 0
SAT_TWOFLOAT = 1
SAT_F64 = 2
SAT_MODES = {"f32": SAT_F32, "twofloat": SAT_TWOFLOAT, "f64": SAT_F64}

OSSART_GRAPH = 1
PARAM_SLAB_TILE = 1
PARAM_VIEWS_PER_WAVE = 2
PARAM_PROJ_CHUNK = 3
PARAM_BACK_CHUNK = 4
PARAM_BACK_STAGE = 5
PARAM_PROJ_STAGE = 6
PARAM_BACK_PASS = 7
PARAM_BACK_MID_RANGE = 8
TUNE_USE_CACHE = 1
BATCH_PROJ_FILL = 1
BATCH_BACK_FILL = 2
BATCH_SOLO_BYTES = 3
BATCH_STREAMS = 4
SERVER_GRAPH = 1
SERVER_POLICY = 2
SERVER_MAX_WAIT_MS = 3
POLICY_FIFO = 0
POLICY_SHORTEST = 1


class DDFanError(RuntimeError):
    """A non-zero status from the native library."""

    def __init__(self, status: int, where: str, detail: str):
        self.status = status
        super().__init__(f"{where}: {detail} (status {status})")


class GeomStruct(ctypes.Structure):
    _fields_ = [
        ("x0", _f), ("y0", _f), ("DNU", _i), ("xds", _fp), ("yds", _fp), ("detCntIdx", _f),
        ("imgXCenter", _f), ("imgYCenter", _f), ("XN", _i), ("YN", _i), ("SLN", _i), ("dx", _f),
    ]


# name -> argtypes (restype is int unless listed in _RESTYPES)
_SIGS = {
    "ddfan_status_string": [_i],
    "ddfan_last_error": [],
    "ddfan_version": [],
    "ddfan_device_count": [_ip],
    "ddfan_set_default_sat_mode": [_i],
    "ddfan_proj": [_f, _f, _i, _p, _p, _f, _f, _p, _i, _i, _i, _i, _p, _p, _f, _p, _p, _i],
    "ddfan_back": [_f, _f, _i, _p, _p, _f, _f, _f, _p, _i, _i, _i, _i, _p, _p, _f, _p, _p, _i],
    "ddfan_proj_noalloc": [_f, _f, _i, _p, _p, _f, _f, _p, _i, _i, _i, _i, _p, _p, _f, _p, _p,
                           _i],
    "ddfan_helical_to_fan": [_p, _p, _p, _ll, _p, _i, _f, _f, _f, _i, _i, _f, _f, _i, _f, _f],
    "ddfan_detector_arc": [_i, _f, _f, _f, _f, _p, _p],
    "ddfan_recon_mask": [_i, _i, _p],
    "ddfan_subset_views": [_i, _i, _i, _p, _ip],
    "ddfan_rebin_starts": [_p, _i, _f, _f, _i, _f, _f, _i, _f, _p],
    "ddfan_dev_alloc": [_sz, ctypes.POINTER(_p)],
    "ddfan_dev_free": [_p],
    "ddfan_h2d": [_p, _p, _sz],
    "ddfan_d2h": [_p, _p, _sz],
    "ddfan_d2d": [_p, _p, _sz],
    "ddfan_dev_mem_info": [ctypes.POINTER(_sz), ctypes.POINTER(_sz)],
    "ddfan_plan_create": [ctypes.POINTER(GeomStruct), _p, _i, _p, _i, _i, ctypes.POINTER(_p)],
    "ddfan_plan_destroy": [_p],
    "ddfan_plan_query": [_p, _ip, _ip, _ip],
    "ddfan_plan_set_param": [_p, _i, _i],
    "ddfan_plan_get_param": [_p, _i, _ip],
    "ddfan_plan_footprint": [ctypes.POINTER(GeomStruct), _i, _i, _i, _i, ctypes.POINTER(_sz)],
    "ddfan_plan_proj": [_p, _p, _p, _i, _i, _i],
    "ddfan_plan_back": [_p, _p, _p, _i, _i, _i],
    "ddfan_plan_sync": [_p],
    "ddfan_plan_vol_tables": [_p, _p, _p, _p, ctypes.POINTER(_sz), ctypes.POINTER(_sz)],
    "ddfan_plan_sino_table": [_p, _p, _i, _p, ctypes.POINTER(_sz)],
    "ddfan_ossart": [_p, _p, _p, _i, _i, _i, _i],
    "ddfan_ossart_host_plain": [_f, _f, _i, _p, _p, _f, _f, _f, _p, _i, _i, _i, _i, _p, _p, _f,
                                _p, _i, _i, _i, _i],
    "ddfan_plan_timing": [_p, _i],
    "ddfan_plan_timing_read": [_p, _i, ctypes.POINTER(ctypes.c_double), _ip],
    "ddfan_plan_trace_proj": [_p, _i, _i, _i, _p],
    "ddfan_plan_trace_back": [_p, _i, _i, _i, _p],
    "ddfan_batch_create": [_i, ctypes.POINTER(_p)],
    "ddfan_batch_add": [_p, _p, _p, _p, _i, _i, _i, _ip],
    "ddfan_batch_size": [_p, _ip],
    "ddfan_batch_set_param": [_p, _i, _ll],
    "ddfan_batch_tune": [_p, _i, ctypes.POINTER(ctypes.c_double)],
    "ddfan_batch_run": [_p, _i],
    "ddfan_batch_destroy": [_p],
    "ddfan_server_create": [_sz, _i, _i, ctypes.POINTER(_p)],
    "ddfan_server_submit": [_p, ctypes.POINTER(GeomStruct), _p, _i, _p, _p, _p, _i, _i, _i,
                            ctypes.POINTER(_ll)],
    "ddfan_server_poll": [_p, _ll, _ip],
    "ddfan_server_set_param": [_p, _i, _ll],
    "ddfan_server_wait": [_p, _ll, _p, ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double)],
    "ddfan_server_stats": [_p, ctypes.POINTER(_ll), ctypes.POINTER(_ll), _ip, ctypes.POINTER(_ll),
                           ctypes.POINTER(_sz), ctypes.POINTER(ctypes.c_double)],
    "ddfan_server_destroy": [_p],
    "ddfan_plan_autotune": [_p, _i, _i, _i, _ip, ctypes.POINTER(ctypes.c_double)],
    "ddfan_plan_apply_tuned": [_p, _i, _ip],
    "ddfan_tune_save": [ctypes.c_char_p],
    "ddfan_tune_load": [ctypes.c_char_p, _ip],
    "ddfan_tune_clear": [],
    "ddfan_helical_to_fan_dev": [_p, _p, _p, _i, _i, _f, _i, _ll, _p, _i, _f, _f, _f, _i, _i, _f,
                                 _f, _i, _f, _f, _p, _p],
    "ddfan_ctpd_read_header": [ctypes.c_char_p, _p, _p, _i],
    "ddfan_ct_image_read_header": [ctypes.c_char_p, _p],
    "ddfan_dicom_read_u16": [ctypes.c_char_p, _i, _i, _p],
    "ddfan_ctpd_read_frames": [ctypes.POINTER(ctypes.c_char_p), _i, _i, _i, _p, _i],
}
_RESTYPES = {"ddfan_status_string": ctypes.c_char_p, "ddfan_last_error": ctypes.c_char_p}

_lib = None


def lib():
    """Load libddfan.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` (make -C aapm_lowdosectimagerecon_competition_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, _i)
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def check(status: int, where: str) -> None:
    if status != DDFAN_OK:
        L = lib()
        detail = (L.ddfan_last_error() or b"").decode(errors="replace") or \
            L.ddfan_status_string(status).decode()
        raise DDFanError(status, where, detail)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)


def ptr(a):
    """ctypes pointer of a numpy array (or None)."""
    if a is None:
        return None
    if isinstance(a, int):
        return ctypes.c_void_p(a)
    if hasattr(a, "ptr"):
        return ctypes.c_void_p(a.ptr)
    return a.ctypes.data_as(_p)


def device_count() -> int:
    n = _i(0)
    call("ddfan_device_count", ctypes.byref(n))
    return n.value


def last_error() -> str:
    return (lib().ddfan_last_error() or b"").decode(errors="replace")


class DeviceBuffer:
    """An owning device allocation made by the library (hipMalloc)."""

    def __init__(self, nbytes: int):
        p = _p()
        call("ddfan_dev_alloc", _sz(int(nbytes)), ctypes.byref(p))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        call("ddfan_h2d", _p(self.ptr), a.ctypes.data_as(_p), _sz(a.nbytes))

    def download(self, out: np.ndarray) -> np.ndarray:
        assert out.flags.c_contiguous and out.nbytes <= self.nbytes
        call("ddfan_d2h", out.ctypes.data_as(_p), _p(self.ptr), _sz(out.nbytes))
        return out

    def copy_from(self, other: "DeviceBuffer", nbytes: int | None = None) -> None:
        n = other.nbytes if nbytes is None else nbytes
        call("ddfan_d2d", _p(self.ptr), _p(other.ptr), _sz(n))

    def to_numpy(self, shape, dtype=np.float32) -> np.ndarray:
        out = np.empty(shape, dtype)
        return self.download(out)

    def free(self) -> None:
        if self.ptr:
            call("ddfan_dev_free", _p(self.ptr))
            self.ptr = 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.ddfan_dev_free(_p(self.ptr))
                self.ptr = 0
        except Exception:  # pragma: no cover - interpreter shutdown
            pass
