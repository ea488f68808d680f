"""Loads the ElectricField binding (multi_species_amd/shim) the way a
FARRSIGHT program does: the reference's compiled FieldStructure.cpp (here
oracle/_ref/libfsref.so, the checker's copy) provides the base class, the
binding provides the HIP subclasses, and every call goes through the virtual
ElectricField::operator()."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "multi_species_amd", "shim", "libfarrsight_field_shim.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libfsref.so")

KIND_DIRECT_HIP, KIND_OPEN_HIP, KIND_TREECODE_HIP = 0, 1, 2
KIND_DIRECT_REF, KIND_OPEN_REF = 10, 11

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def available():
    return os.path.exists(SHIM) and os.path.exists(REF)


def lib():
    global _lib
    if _lib is None:
        ctypes.CDLL(REF, mode=ctypes.RTLD_GLOBAL)
        L = ctypes.CDLL(SHIM)
        L.fs_shim_new_field.restype = ctypes.c_void_p
        L.fs_shim_new_field.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.fs_shim_call.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _dp, _dp,
                                   ctypes.c_int]
        L.fs_shim_print.argtypes = [ctypes.c_void_p]
        L.fs_shim_delete.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


class Field:
    """An ElectricField object of the binding or of the reference."""

    def __init__(self, kind, L=0.0, eps=0.0, mac=0.0, degree=0, max_source=0):
        self.h = lib().fs_shim_new_field(kind, L, eps, mac, degree, max_source)
        if not self.h:
            raise RuntimeError(f"ElectricField kind {kind} could not be made")

    def __call__(self, targets, sources, q_ws):
        # the reference's contract: non-const double pointers, es overwritten
        t = np.array(targets, dtype=np.float64)
        s = np.array(sources, dtype=np.float64)
        w = np.array(q_ws, dtype=np.float64)
        # es[0:nt], or es[0:ns] for the treecode (given no targets: it evaluates at
        # the sources)
        es = np.full(t.size if t.size else s.size, np.nan)
        P = lambda a: a.ctypes.data_as(_dp)  # noqa: E731
        if lib().fs_shim_call(self.h, P(es), P(t), t.size, P(s), P(w), s.size) != 0:
            raise RuntimeError("ElectricField::operator() threw")
        return es

    def close(self):
        if self.h:
            lib().fs_shim_delete(self.h)
            self.h = None
