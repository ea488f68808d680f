"""MI355X (gfx950) implementation of the FARRSIGHT hot path.

Python view of ``libfarrsight_hip.so`` (C ABI in ``include/farrsight_hip.h``).
The library must be built (``python -m multi_species_amd.build``); there is no
CPU fallback, every entry point runs the HIP kernels or raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfarrsight_hip.so")
HEADER = os.path.join(HERE, "..", "include", "farrsight_hip.h")

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_d = ctypes.c_double
_int = ctypes.c_int

_lib = None


class FarrsightError(RuntimeError):
    pass


def lib():
    """Loads libfarrsight_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FarrsightError(
                f"{LIB_PATH} is not built; run `python -m multi_species_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.fs_last_error.restype = ctypes.c_char_p
        L.fs_ctx_create.argtypes = [ctypes.POINTER(_vp), _int]
        L.fs_ctx_destroy.argtypes = [_vp]
        L.fs_ctx_stream.argtypes = [_vp]
        L.fs_ctx_stream.restype = _vp
        L.fs_ctx_synchronize.argtypes = [_vp]
        L.fs_emq_periodic_f64.argtypes = [_vp, _dp, _dp, _i64, _dp, _dp, _i64, _d, _d]
        L.fs_emq_periodic_f64_dev.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _d, _d, _vp]
        L.fs_emq_set_tuning.argtypes = [_vp, _int, _int]
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().fs_last_error().decode(errors="replace")
        raise FarrsightError(f"{what} failed ({rc}): {msg}")


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _P(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class Context:
    """One HIP device with its own stream and device scratch (fs_ctx)."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib().fs_ctx_create(ctypes.byref(h), device), "fs_ctx_create")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            lib().fs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return lib().fs_ctx_stream(self.h) or 0

    def synchronize(self):
        _check(lib().fs_ctx_synchronize(self.h), "fs_ctx_synchronize")

    def set_tuning(self, targets_per_thread: int = 0, source_splits: int = 0):
        _check(lib().fs_emq_set_tuning(self.h, targets_per_thread, source_splits),
               "fs_emq_set_tuning")

    # ---- part A: periodic field -------------------------------------------
    def emq_periodic_f64(self, targets, sources, q_ws, L: float, eps: float) -> np.ndarray:
        """E_MQ_DirectSum::operator() on host arrays (synchronous)."""
        t, s, w = _f64(targets), _f64(sources), _f64(q_ws)
        if s.size != w.size:
            raise ValueError("sources and q_ws differ in length")
        es = np.empty(t.size)
        _check(lib().fs_emq_periodic_f64(self.h, _P(es), _P(t), t.size, _P(s), _P(w),
                                         s.size, L, eps), "fs_emq_periodic_f64")
        return es

    def emq_periodic_f64_dev(self, es_ptr: int, t_ptr: int, nt: int, s_ptr: int, w_ptr: int,
                             ns: int, L: float, eps: float, stream: Optional[int] = None):
        """Same on device pointers, enqueued on `stream` (default: the context's)."""
        _check(lib().fs_emq_periodic_f64_dev(self.h, es_ptr, t_ptr, nt, s_ptr, w_ptr, ns, L,
                                             eps, stream or None),
               "fs_emq_periodic_f64_dev")


def declared_symbols(header: str = HEADER):
    """Function names declared in include/farrsight_hip.h."""
    import re

    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fs_[a-z0-9_]+)\s*\(", text)))
