"""MI355X (gfx950) implementation of the FARRSIGHT hot path.

Python view of ``libfarrsight_hip.so`` (C ABI in ``include/farrsight_hip.h``).
The library must be built (``python -m multi_species_amd.build``); there is no
CPU fallback, every entry point runs the HIP kernels or raises.
"""
from __future__ import annotations

import ctypes
import os
import sys
import weakref
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
        L.fs_emq_open_f64.argtypes = [_vp, _dp, _dp, _i64, _dp, _dp, _i64, _d]
        L.fs_emq_open_f64_dev.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _d, _vp]
        L.fs_external_field_add.argtypes = [_vp, _dp, _dp, _i64, _d, ctypes.c_int, _d, _d, _d]
        L.fs_external_field_add_dev.argtypes = [_vp, _vp, _vp, _i64, _d, ctypes.c_int, _d, _d,
                                                _d, _vp]
        L.fs_emq_set_tuning.argtypes = [_vp, _int, _int]
        L.fs_emq_set_symmetric.argtypes = [_vp, _int]
        L.fs_emq_set_pair_once_budget.argtypes = [_vp, _i64]
        L.fs_emq_periodic_f64_noalloc.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _d, _d,
                                                  _vp]
        L.fs_emq_pair_once_plan.argtypes = [_i64, _i64, ctypes.POINTER(_i64),
                                            ctypes.POINTER(_int), ctypes.POINTER(_i64)]
        L.fs_deck_footprint.argtypes = [ctypes.POINTER(Deck), _int, ctypes.POINTER(_i64),
                                        ctypes.POINTER(_i64)]
        L.fs_emq_f64_pair_evals.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64)]
        L.fs_emq_periodic_mixed.argtypes = [_vp, _dp, _dp, _i64, _dp, _dp, _i64, _d, _d]
        L.fs_emq_mixed_shape.argtypes = [_vp, ctypes.POINTER(_int), ctypes.POINTER(_int)]
        L.fs_emq_set_mixed_target.argtypes = [_vp, _d]
        L.fs_emq_set_mixed_min_pairs.argtypes = [_vp, _i64]
        L.fs_emq_periodic_mixed_dev.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _d, _d,
                                                _vp]
        L.fs_sim_set_timing.argtypes = [_vp, _int]
        L.fs_sim_remesh_ms.argtypes = [_vp, ctypes.POINTER(ctypes.c_double)]
        L.fs_sim_set_refinement.argtypes = [_vp, _int]
        L.fs_sim_step_counts.argtypes = [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]
        L.fs_sim_refinement_counts.argtypes = [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]
        L.fs_emq_treecode_f64.argtypes = [_vp, _dp, _dp, _dp, _i64, _d, _d, _d, _int, _int,
                                          ctypes.POINTER(_i64)]
        _ip = ctypes.POINTER(ctypes.c_int32)
        L.fs_treecode_tree.argtypes = [_vp, _dp, _i64, _d, _d, _int, _int, ctypes.POINTER(_i64),
                                       _ip, _ip, _ip, _ip, _ip, _ip]
        L.fs_remesh_interp_grid.argtypes = [_vp, _ip, _i64, _dp, _dp, _dp, _i64, _dp, _dp, _int,
                                            _int, _d, _d, _int, _d, _dp]
        L.fs_remesh_interp_points.argtypes = [_vp, _ip, _i64, _dp, _dp, _dp, _i64, _dp, _dp,
                                              _i64, _d, _d, _int, _d, _dp]
        L.fs_mesh_build.argtypes = [ctypes.POINTER(SimParams), ctypes.POINTER(SpeciesParams),
                                    ctypes.POINTER(_i64), ctypes.POINTER(_i64), _ip, _dp, _dp,
                                    _dp, _dp]
        L.fs_sim_create.argtypes = [_vp, ctypes.POINTER(SimParams), ctypes.POINTER(SpeciesParams),
                                    _int, ctypes.POINTER(_vp)]
        L.fs_sim_destroy.argtypes = [_vp]
        L.fs_sim_step.argtypes = [_vp]
        L.fs_sim_iter.argtypes = [_vp, ctypes.POINTER(_int)]
        L.fs_sim_species_counts.argtypes = [_vp, _int, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]
        L.fs_sim_get_points.argtypes = [_vp, _int, _dp, _dp, _dp, _dp, _dp]
        L.fs_sim_field_governor.argtypes = [_vp, _int, _dp, _dp, _dp]
        L.fs_sim_get_panels.argtypes = [_vp, _int, _ip]
        L.fs_sim_diagnostics.argtypes = [_vp, _int, _dp]
        L.fs_sim_diagnostics_full.argtypes = [_vp, _int, _dp, _int]
        _cpp = ctypes.POINTER(ctypes.c_char_p)
        L.fs_write_species_output.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _int, _i64, _dp,
                                              _dp, _dp, _dp, _dp, _i64, _ip]
        L.fs_read_species_output.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _int,
                                             ctypes.POINTER(_i64), _dp, _dp, _dp,
                                             ctypes.POINTER(_i64), _ip]
        L.fs_batch_create_empty.argtypes = [_vp, _int, ctypes.POINTER(_vp)]
        L.fs_sim_write_output.argtypes = [_vp, _int, ctypes.c_char_p, _cpp]
        L.fs_sim_restart_deck.argtypes = [_vp, ctypes.POINTER(Deck), ctypes.c_char_p, _cpp, _int,
                                          ctypes.POINTER(_int)]
        L.fs_server_write_result.argtypes = [_vp, _i64, ctypes.c_char_p, _cpp]
        L.fs_server_submit_restart.argtypes = [_vp, ctypes.POINTER(Deck), ctypes.c_char_p, _cpp,
                                               _int, _int, ctypes.POINTER(_i64)]
        L.fs_batch_create.argtypes = [_vp, ctypes.POINTER(Deck), _int, _int, ctypes.POINTER(_vp)]
        L.fs_sim_add_deck.argtypes = [_vp, ctypes.POINTER(Deck), ctypes.POINTER(_int)]
        L.fs_sim_remove_deck.argtypes = [_vp, _int]
        L.fs_sim_num_decks.argtypes = [_vp, ctypes.POINTER(_int)]
        L.fs_sim_deck_info.argtypes = [_vp, _int, ctypes.POINTER(_int), ctypes.POINTER(_int),
                                       ctypes.POINTER(_int), ctypes.POINTER(_i64)]
        L.fs_sim_set_options.argtypes = [_vp, _int, _int]
        L.fs_server_create.argtypes = [_vp, ctypes.POINTER(ServerParams), ctypes.POINTER(_vp)]
        L.fs_server_submit.argtypes = [_vp, ctypes.POINTER(Deck), _int, ctypes.POINTER(_i64)]
        L.fs_server_status.argtypes = [_vp, _i64, ctypes.POINTER(_int), ctypes.POINTER(_int)]
        L.fs_server_wait.argtypes = [_vp, _i64, _d]
        L.fs_server_result.argtypes = [_vp, _i64, _int, ctypes.POINTER(_i64), _dp, _dp, _dp,
                                       _dp, _dp, _dp]
        L.fs_server_get_stats.argtypes = [_vp, ctypes.POINTER(ServerStats)]
        L.fs_server_destroy.argtypes = [_vp]
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
        # simulations and servers on this context: closed before it is
        # destroyed (their C objects use the context until they are)
        self._children = weakref.WeakSet()

    def _adopt(self, child):
        self._children.add(child)

    def close(self):
        if getattr(self, "h", None):
            for child in list(getattr(self, "_children", ())):
                child.close()
            lib().fs_ctx_destroy(self.h)
            self.h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the HIP runtime may already be gone (and this
        # module's globals with it: the check is bound at definition)
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return lib().fs_ctx_stream(self.h) or 0

    def synchronize(self):
        _check(lib().fs_ctx_synchronize(self.h), "fs_ctx_synchronize")

    def f64_pair_evals(self) -> int:
        """Kernel evaluations of the last FP64 direct sum on this context."""
        v = ctypes.c_int64()
        _check(lib().fs_emq_f64_pair_evals(self.h, ctypes.byref(v)), "fs_emq_f64_pair_evals")
        return v.value

    def set_symmetric(self, mode: int = 0):
        """Pair-once FP64 sum when targets are the sources: 0 = from 16384
        points up (default), 1 = never, 2 = always."""
        _check(lib().fs_emq_set_symmetric(self.h, mode), "fs_emq_set_symmetric")

    def set_pair_once_budget(self, nbytes: int = 0):
        """Bound on the pair-once form's partials in bytes (0 = 4 GiB); the
        result does not depend on it."""
        _check(lib().fs_emq_set_pair_once_budget(self.h, int(nbytes)),
               "fs_emq_set_pair_once_budget")

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

    def emq_open_f64(self, targets, sources, q_ws, eps: float) -> np.ndarray:
        """E_MQ_DirectSum_openbcs::operator() on host arrays (synchronous)."""
        t, s, w = _f64(targets), _f64(sources), _f64(q_ws)
        if s.size != w.size:
            raise ValueError("sources and q_ws differ in length")
        es = np.empty(t.size)
        _check(lib().fs_emq_open_f64(self.h, _P(es), _P(t), t.size, _P(s), _P(w), s.size, eps),
               "fs_emq_open_f64")
        return es

    def external_field_add(self, es, targets, t: float, type_: int, Am: float, k: float,
                           omega: float) -> np.ndarray:
        """ExternalTanh (type 0) / ExternalSine (1) / ExternalLogistic (2)
        ::operator() on host arrays: returns es + a(t) sin(k x - omega t)."""
        e, x = _f64(es).copy(), _f64(targets)
        if e.size != x.size:
            raise ValueError("es and targets differ in length")
        _check(lib().fs_external_field_add(self.h, _P(e), _P(x), x.size, t, type_, Am, k, omega),
               "fs_external_field_add")
        return e

    def emq_periodic_mixed(self, targets, sources, q_ws, L: float, eps: float,
                           same: bool = False) -> np.ndarray:
        """Mixed-precision direct sum on host arrays (synchronous). same=True
        evaluates at the sources (targets ignored) with one shared sort."""
        s, w = _f64(sources), _f64(q_ws)
        t = s if same else _f64(targets)
        if s.size != w.size:
            raise ValueError("sources and q_ws differ in length")
        es = np.empty(t.size)
        _check(lib().fs_emq_periodic_mixed(self.h, _P(es), _P(t), t.size, _P(s), _P(w),
                                           s.size, L, eps), "fs_emq_periodic_mixed")
        return es

    def set_mixed_target(self, rel_err: float):
        """Error target of the mixed calls on this context: 0 = minimum FP64
        band, no probe; > 0 = each call first measures, at up to 128 targets
        per segment against the exact FP64 field, the error of each candidate
        band and takes the narrowest one within rel_err * max|E|."""
        _check(lib().fs_emq_set_mixed_target(self.h, rel_err), "fs_emq_set_mixed_target")

    def set_mixed_min_pairs(self, min_pairs: int):
        """Mixed calls with nt * ns below min_pairs return the FP64 sum bit
        for bit (faster at such sizes); 0 = every call through the mixed
        kernels. Default MIXED_MIN_PAIRS_DEFAULT."""
        _check(lib().fs_emq_set_mixed_min_pairs(self.h, int(min_pairs)),
               "fs_emq_set_mixed_min_pairs")

    def mixed_shape(self):
        """(targets per thread, workgroups per CU) of the last mixed call."""
        t, g = _int(), _int()
        _check(lib().fs_emq_mixed_shape(self.h, ctypes.byref(t), ctypes.byref(g)),
               "fs_emq_mixed_shape")
        return t.value, g.value

    def emq_periodic_mixed_dev(self, es_ptr: int, t_ptr: int, nt: int, s_ptr: int, w_ptr: int,
                               ns: int, L: float, eps: float, stream: Optional[int] = None):
        """Mixed-precision direct sum on device pointers, enqueued on `stream`."""
        _check(lib().fs_emq_periodic_mixed_dev(self.h, es_ptr, t_ptr, nt, s_ptr, w_ptr, ns, L,
                                               eps, stream or None),
               "fs_emq_periodic_mixed_dev")

    def emq_treecode_f64(self, sources, q_ws, L: float, eps: float, mac: float, degree: int,
                         max_source: int, return_evals: bool = False):
        """E_MQ_Treecode::operator() on host arrays: E at the sources."""
        s, w = _f64(sources), _f64(q_ws)
        if s.size != w.size:
            raise ValueError("sources and q_ws differ in length")
        es = np.empty(s.size)
        ev = (_i64 * 2)()
        _check(lib().fs_emq_treecode_f64(self.h, _P(es), _P(s), _P(w), s.size, L, eps, mac,
                                         degree, max_source, ev), "fs_emq_treecode_f64")
        return (es, (ev[0], ev[1])) if return_evals else es

    def treecode_tree(self, sources, L: float, mac: float, degree: int, max_source: int):
        """The tree and interaction lists emq_treecode_f64 builds on the device
        (fs_treecode_tree): stats (nodes, leaves, far entries, near entries),
        perm, leaf nodes and the far/near lists (CSR), a node as
        (first, last, depth) in tree order."""
        s = _f64(sources)
        st = (_i64 * 4)()
        _check(lib().fs_treecode_tree(self.h, _P(s), s.size, L, mac, degree, max_source, st,
                                      None, None, None, None, None, None), "fs_treecode_tree")
        nn, nl, nf, nr = (int(v) for v in st)
        out = dict(stats=(nn, nl, nf, nr), perm=np.empty(s.size, np.int32),
                   leaves=np.empty((nl, 3), np.int32), far_off=np.empty(nl + 1, np.int32),
                   far=np.empty((max(nf, 1), 3), np.int32), near_off=np.empty(nl + 1, np.int32),
                   near=np.empty((max(nr, 1), 3), np.int32))
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        _check(lib().fs_treecode_tree(self.h, _P(s), s.size, L, mac, degree, max_source, st,
                                      ip(out["perm"]), ip(out["leaves"]), ip(out["far_off"]),
                                      ip(out["far"]), ip(out["near_off"]), ip(out["near"])),
               "fs_treecode_tree")
        out["far"], out["near"] = out["far"][:nf], out["near"][:nr]
        return out

    def emq_periodic_f64_dev(self, es_ptr: int, t_ptr: int, nt: int, s_ptr: int, w_ptr: int,
                             ns: int, L: float, eps: float, stream: Optional[int] = None):
        """Same on device pointers, enqueued on `stream` (default: the context's)."""
        _check(lib().fs_emq_periodic_f64_dev(self.h, es_ptr, t_ptr, nt, s_ptr, w_ptr, ns, L,
                                             eps, stream or None),
               "fs_emq_periodic_f64_dev")

    def emq_periodic_f64_noalloc(self, es_ptr: int, t_ptr: int, nt: int, s_ptr: int,
                                 w_ptr: int, ns: int, L: float, eps: float,
                                 stream: Optional[int] = None):
        """fs_emq_periodic_f64_dev without any device allocation: the work
        space is the library's fixed 256 MiB per device (at most about 1.6e6
        points with targets = sources). Calls overlapping on one device race."""
        _check(lib().fs_emq_periodic_f64_noalloc(self.h, es_ptr, t_ptr, nt, s_ptr, w_ptr, ns, L,
                                                 eps, stream or None),
               "fs_emq_periodic_f64_noalloc")

    def emq_open_f64_dev(self, es_ptr: int, t_ptr: int, nt: int, s_ptr: int, w_ptr: int,
                         ns: int, eps: float, stream: Optional[int] = None):
        """fs_emq_open_f64 on device pointers, enqueued on `stream`."""
        _check(lib().fs_emq_open_f64_dev(self.h, es_ptr, t_ptr, nt, s_ptr, w_ptr, ns, eps,
                                         stream or None),
               "fs_emq_open_f64_dev")


PANEL_INTS = 20


class SpeciesParams(ctypes.Structure):
    """fs_species_params (the per-species deck keys)."""
    _fields_ = [
        ("q", _d), ("m", _d), ("p_min", _d), ("p_max", _d),
        ("initial_height", _int), ("p_height", _int), ("max_height", _int),
        ("do_adaptively_refine", _int), ("n_amr_epsilons", _int), ("amr_epsilons", _d * 5),
        ("ics_type", _int), ("normalized_wavenumber", _d), ("amp", _d), ("pth", _d),
        ("pstr", _d), ("beam_frac", _d), ("beam_v", _d), ("beam_vth", _d),
    ]


class SimParams(ctypes.Structure):
    """fs_sim_params (the deck's top-level keys)."""
    _fields_ = [
        ("x_min", _d), ("x_max", _d), ("bcs", _int), ("quadrature", _int),
        ("greens_epsilon", _d), ("use_treecode", _int), ("mac", _d), ("degree", _int),
        ("max_source", _int), ("dt", _d), ("remesh_period", _int),
        ("use_external_field", _int), ("external_type", _int), ("Am", _d), ("k", _d),
        ("omega", _d),
    ]


class Deck(ctypes.Structure):
    """fs_deck: one deck of a batch."""
    _fields_ = [("sim", ctypes.POINTER(SimParams)), ("species", ctypes.POINTER(SpeciesParams)),
                ("n_species", _int)]


PRECISION_F64 = 0
PRECISION_MIXED = 1
MIXED_MIN_PAIRS_DEFAULT = 1 << 29

DIAG_COUNT = 14
DIAG_NAMES = ("charge", "momentum", "kinetic", "electric", "entropy", "l1", "l2", "min_f",
              "max_f", "negative_area", "num_negative", "frac_negative", "negative_mass",
              "num_points")

REQ_QUEUED, REQ_BUILDING, REQ_READY, REQ_ACTIVE, REQ_DONE, REQ_FAILED = 0, 1, 2, 3, 4, -1


class ServerParams(ctypes.Structure):
    """fs_server_params."""
    _fields_ = [("hbm_budget_bytes", _d), ("max_active", _int), ("precision", _int),
                ("use_graph", _int)]


class ServerStats(ctypes.Structure):
    """fs_server_stats."""
    _fields_ = [("submitted", _i64), ("completed", _i64), ("failed", _i64), ("active", _i64),
                ("queued", _i64), ("steps", _i64), ("point_steps", _i64), ("busy_s", _d),
                ("span_s", _d), ("point_steps_per_s", _d), ("step_p50_ms", _d),
                ("step_p99_ms", _d), ("hbm_in_use", _i64), ("steps_replayed", _i64)]


def _deck(sim: SimParams, species):
    arr = (SpeciesParams * len(species))(*species)
    d = Deck(ctypes.pointer(sim), ctypes.cast(arr, ctypes.POINTER(SpeciesParams)), len(species))
    return d, arr


def species_params(q=-1.0, m=1.0, p_min=-1.0, p_max=1.0, initial_height=6, p_height=0,
                   max_height=None, do_adaptively_refine=False, amr_epsilons=(), ics_type=1,
                   normalized_wavenumber=1.0, amp=0.0, pth=1.0, pstr=0.0, beam_frac=0.0,
                   beam_v=0.0, beam_vth=1.0, **_ignored) -> SpeciesParams:
    """A species entry with the reference deck's defaults (AMRSimulation_load.cpp:226-254)."""
    sp = SpeciesParams()
    sp.q, sp.m, sp.p_min, sp.p_max = q, m, p_min, p_max
    sp.initial_height, sp.p_height = initial_height, p_height
    sp.max_height = initial_height if max_height is None else max_height
    sp.do_adaptively_refine = int(bool(do_adaptively_refine))
    eps = list(amr_epsilons)[:5]
    sp.n_amr_epsilons = len(eps)
    for i, e in enumerate(eps):
        sp.amr_epsilons[i] = e
    sp.ics_type = ics_type
    sp.normalized_wavenumber, sp.amp, sp.pth, sp.pstr = normalized_wavenumber, amp, pth, pstr
    sp.beam_frac, sp.beam_v, sp.beam_vth = beam_frac, beam_v, beam_vth
    return sp


def sim_params(x_min=0.0, x_max=0.0, bcs=0, quadrature=1, greens_epsilon=0.2, use_treecode=0,
               mac=-1.0, degree=-1, max_source=2000, dt=0.5, remesh_period=1,
               use_external_field=0, external_type=0, Am=0.0, k=0.0, omega=0.0,
               **_ignored) -> SimParams:
    """Top-level deck keys with the reference's defaults (AMRSimulation_load.cpp:14-59),
    and the external drive block (:99-136)."""
    p = SimParams()
    p.x_min, p.x_max, p.bcs, p.quadrature = x_min, x_max, bcs, quadrature
    p.greens_epsilon, p.use_treecode, p.mac = greens_epsilon, use_treecode, mac
    p.degree, p.max_source, p.dt, p.remesh_period = degree, max_source, dt, remesh_period
    p.use_external_field, p.external_type = use_external_field, external_type
    p.Am, p.k, p.omega = Am, k, omega
    return p


def pair_once_plan(n: int, budget_bytes: int = 0):
    """(scratch bytes, passes, kernel evaluations) of the pair-once FP64 sum
    of n points under a budget (host-only)."""
    b, p, e = _i64(), _int(), _i64()
    _check(lib().fs_emq_pair_once_plan(int(n), int(budget_bytes), ctypes.byref(b),
                                       ctypes.byref(p), ctypes.byref(e)), "fs_emq_pair_once_plan")
    return b.value, p.value, e.value


def deck_footprint(sim: "SimParams", species, precision: int = 0):
    """(estimated device bytes, points) of a deck in a batch or server
    (host-only: builds the deck's meshes on the host)."""
    d, arr = _deck(sim, species)
    b, n = _i64(), _i64()
    _check(lib().fs_deck_footprint(ctypes.byref(d), int(precision), ctypes.byref(b),
                                   ctypes.byref(n)), "fs_deck_footprint")
    return b.value, n.value


def _names(names):
    arr = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
    return ctypes.cast(arr, ctypes.POINTER(ctypes.c_char_p)), arr


def write_species_output(sim_dir: str, name: str, it: int, xs, ps, fs, panels, q_ws=None,
                         es=None):
    """One species' files in the reference's formats (host arrays)."""
    x, p, f = _f64(xs), _f64(ps), _f64(fs)
    rec = _i32(panels).reshape(-1, PANEL_INTS)
    w = None if q_ws is None else _f64(q_ws)
    e = None if es is None else _f64(es)
    _check(lib().fs_write_species_output(sim_dir.encode(), name.encode(), it, x.size, _P(x),
                                         _P(p), _P(f), None if w is None else _P(w),
                                         None if e is None else _P(e), rec.shape[0], _PI(rec)),
           "fs_write_species_output")


def read_species_output(sim_dir: str, name: str, it: int):
    """(xs, ps, fs, panels) read back from one species' files."""
    n, P = _i64(), _i64()
    _check(lib().fs_read_species_output(sim_dir.encode(), name.encode(), it, ctypes.byref(n),
                                        None, None, None, ctypes.byref(P), None),
           "fs_read_species_output")
    xs, ps, fs = (np.empty(n.value) for _ in range(3))
    rec = np.empty((P.value, PANEL_INTS), dtype=np.int32)
    _check(lib().fs_read_species_output(sim_dir.encode(), name.encode(), it, None, _P(xs),
                                        _P(ps), _P(fs), None, _PI(rec)),
           "fs_read_species_output")
    return xs, ps, fs, rec


def mesh_build(sim: SimParams, sp: SpeciesParams):
    """The initial mesh of one species, built by the library's host code."""
    n, P = _i64(), _i64()
    _check(lib().fs_mesh_build(ctypes.byref(sim), ctypes.byref(sp), ctypes.byref(n),
                               ctypes.byref(P), None, None, None, None, None), "fs_mesh_build")
    rec = np.empty((P.value, PANEL_INTS), dtype=np.int32)
    xs, ps, fs, qw = (np.empty(n.value) for _ in range(4))
    _check(lib().fs_mesh_build(ctypes.byref(sim), ctypes.byref(sp), None, None, _PI(rec),
                               _P(xs), _P(ps), _P(fs), _P(qw)), "fs_mesh_build")
    return rec, xs, ps, fs, qw


class Simulation:
    """AMRSimulation on the device (fs_sim): one deck, or a batch of decks.

    Simulation(ctx, sim, species) is one deck (fs_sim_create);
    Simulation.batch(ctx, [(sim, species), ...], precision) a batch."""

    def __init__(self, ctx: "Context", sim: SimParams = None, species=None, *, _handle=None,
                 _keep=None):
        self.ctx = ctx
        ctx._adopt(self)
        if _handle is not None:
            self.h = _handle
            self._keep = _keep
            return
        self._sim = sim
        arr = (SpeciesParams * len(species))(*species)
        self._keep = [sim, arr]
        h = _vp()
        _check(lib().fs_sim_create(ctx.h, ctypes.byref(sim), arr, len(species), ctypes.byref(h)),
               "fs_sim_create")
        self.h = h

    @classmethod
    def batch(cls, ctx: "Context", decks, precision: int = PRECISION_F64) -> "Simulation":
        keep, ds = [], []
        for sim, species in decks:
            d, arr = _deck(sim, species)
            keep += [sim, arr]
            ds.append(d)
        darr = (Deck * len(ds))(*ds)
        h = _vp()
        _check(lib().fs_batch_create(ctx.h, darr, len(ds), precision, ctypes.byref(h)),
               "fs_batch_create")
        return cls(ctx, _handle=h, _keep=keep)

    @classmethod
    def empty(cls, ctx: "Context", precision: int = PRECISION_F64) -> "Simulation":
        """A batch with no decks (they join with add_deck or restart_deck)."""
        h = _vp()
        _check(lib().fs_batch_create_empty(ctx.h, precision, ctypes.byref(h)),
               "fs_batch_create_empty")
        return cls(ctx, _handle=h, _keep=[])

    def write_output(self, deck: int, sim_dir: str, names):
        """The deck's state in the reference's files at the step it reached."""
        p, keep = _names(names)
        _check(lib().fs_sim_write_output(self.h, deck, sim_dir.encode(), p),
               "fs_sim_write_output")

    def restart_deck(self, sim: SimParams, species, sim_dir: str, names, it: int) -> int:
        """A deck joins from the files of step `it` (written after a remesh)."""
        d, arr = _deck(sim, species)
        p, keep = _names(names)
        idx = _int()
        _check(lib().fs_sim_restart_deck(self.h, ctypes.byref(d), sim_dir.encode(), p, it,
                                         ctypes.byref(idx)), "fs_sim_restart_deck")
        return idx.value

    def add_deck(self, sim: SimParams, species) -> int:
        d, arr = _deck(sim, species)
        idx = _int()
        _check(lib().fs_sim_add_deck(self.h, ctypes.byref(d), ctypes.byref(idx)),
               "fs_sim_add_deck")
        return idx.value

    def remove_deck(self, deck: int):
        _check(lib().fs_sim_remove_deck(self.h, deck), "fs_sim_remove_deck")

    @property
    def n_decks(self) -> int:
        n = _int()
        _check(lib().fs_sim_num_decks(self.h, ctypes.byref(n)), "fs_sim_num_decks")
        return n.value

    def deck_info(self, deck: int):
        """(first_species, n_species, iter, n_points) of a deck."""
        a, b, c, n = _int(), _int(), _int(), _i64()
        _check(lib().fs_sim_deck_info(self.h, deck, ctypes.byref(a), ctypes.byref(b),
                                      ctypes.byref(c), ctypes.byref(n)), "fs_sim_deck_info")
        return a.value, b.value, c.value, n.value

    def set_timing(self, on: bool = True):
        """Eager steps with HIP events around the remesh's device work."""
        _check(lib().fs_sim_set_timing(self.h, int(bool(on))), "fs_sim_set_timing")

    def remesh_ms(self) -> float:
        """Device time of the last step's remesh (set_timing first)."""
        out = ctypes.c_double()
        _check(lib().fs_sim_remesh_ms(self.h, ctypes.byref(out)), "fs_sim_remesh_ms")
        return out.value

    def set_refinement(self, on_device: bool = True):
        """Adaptive remeshes refine on the device (default) or on the host."""
        _check(lib().fs_sim_set_refinement(self.h, 0 if on_device else 1),
               "fs_sim_set_refinement")

    def refinement_counts(self) -> tuple:
        """(remeshes refined on the device, on the host)."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().fs_sim_refinement_counts(self.h, ctypes.byref(a), ctypes.byref(b)),
               "fs_sim_refinement_counts")
        return a.value, b.value

    def step_counts(self) -> tuple:
        """(steps replayed from a captured graph, steps run eagerly)."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().fs_sim_step_counts(self.h, ctypes.byref(a), ctypes.byref(b)),
               "fs_sim_step_counts")
        return a.value, b.value

    def set_options(self, precision: int = PRECISION_F64, use_graph: bool = True):
        _check(lib().fs_sim_set_options(self.h, precision, int(bool(use_graph))),
               "fs_sim_set_options")

    @property
    def n_species(self) -> int:
        return sum(self.deck_info(d)[1] for d in range(self.n_decks))

    def close(self):
        if getattr(self, "h", None):
            lib().fs_sim_destroy(self.h)
            self.h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the HIP runtime may already be gone (and this
        # module's globals with it: the check is bound at definition)
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        for _ in range(n):
            _check(lib().fs_sim_step(self.h), "fs_sim_step")

    @property
    def iter(self) -> int:
        it = _int()
        _check(lib().fs_sim_iter(self.h, ctypes.byref(it)), "fs_sim_iter")
        return it.value

    def counts(self, k: int):
        n, P = _i64(), _i64()
        _check(lib().fs_sim_species_counts(self.h, k, ctypes.byref(n), ctypes.byref(P)),
               "fs_sim_species_counts")
        return n.value, P.value

    def field_governor(self, deck: int):
        """(max|E|, errors[8], bands[8]) of the mixed-precision accuracy governor
        for one deck (its last full-field evaluation)."""
        em = np.zeros(1)
        errs, bands = np.zeros(8), np.zeros(8)
        _check(lib().fs_sim_field_governor(self.h, deck, _P(em), _P(errs), _P(bands)),
               "fs_sim_field_governor")
        return float(em[0]), errs, bands

    def points(self, k: int):
        """(xs, ps, fs, q_ws, es) of species k."""
        n, _ = self.counts(k)
        arrs = [np.empty(n) for _ in range(5)]
        _check(lib().fs_sim_get_points(self.h, k, *[_P(a) for a in arrs]), "fs_sim_get_points")
        return tuple(arrs)

    def panels(self, k: int) -> np.ndarray:
        _, P = self.counts(k)
        rec = np.empty((P, PANEL_INTS), dtype=np.int32)
        _check(lib().fs_sim_get_panels(self.h, k, _PI(rec)), "fs_sim_get_panels")
        return rec

    def diagnostics(self, k: int) -> np.ndarray:
        out = np.empty(4)
        _check(lib().fs_sim_diagnostics(self.h, k, _P(out)), "fs_sim_diagnostics")
        return out

    def diagnostics_full(self, k: int) -> np.ndarray:
        """The DIAG_COUNT per-species diagnostics (DIAG_NAMES), reduced on the device."""
        out = np.empty(DIAG_COUNT)
        _check(lib().fs_sim_diagnostics_full(self.h, k, _P(out), DIAG_COUNT),
               "fs_sim_diagnostics_full")
        return out


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _PI(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _remesh_args(panels, old_xs, old_ps, old_fs):
    rec = _i32(panels).reshape(-1, PANEL_INTS)
    ox, op, of = _f64(old_xs), _f64(old_ps), _f64(old_fs)
    if not (ox.size == op.size == of.size):
        raise ValueError("old point arrays differ in length")
    return rec, ox, op, of


def remesh_interp_grid(ctx: "Context", panels, old_xs, old_ps, old_fs, new_xs, new_ps,
                       nx: int, np_: int, x_min: float, x_max: float, periodic: bool,
                       f_beyond: float) -> np.ndarray:
    """AMRStructure::interpolate_to_initial_xps on the GPU (host arrays)."""
    rec, ox, op, of = _remesh_args(panels, old_xs, old_ps, old_fs)
    nxs, nps = _f64(new_xs), _f64(new_ps)
    out = np.empty(nxs.size)
    _check(lib().fs_remesh_interp_grid(ctx.h, _PI(rec), rec.shape[0], _P(ox), _P(op), _P(of),
                                       ox.size, _P(nxs), _P(nps), nx, np_, x_min, x_max,
                                       int(periodic), f_beyond, _P(out)),
           "fs_remesh_interp_grid")
    return out


def remesh_interp_points(ctx: "Context", panels, old_xs, old_ps, old_fs, x, p,
                         x_min: float, x_max: float, periodic: bool,
                         f_beyond: float) -> np.ndarray:
    """AMRStructure::interpolate_from_mesh at many points on the GPU."""
    rec, ox, op, of = _remesh_args(panels, old_xs, old_ps, old_fs)
    xx, pp = _f64(x), _f64(p)
    out = np.empty(xx.size)
    _check(lib().fs_remesh_interp_points(ctx.h, _PI(rec), rec.shape[0], _P(ox), _P(op), _P(of),
                                         ox.size, _P(xx), _P(pp), xx.size, x_min, x_max,
                                         int(periodic), f_beyond, _P(out)),
           "fs_remesh_interp_points")
    return out


def declared_symbols(header: str = HEADER):
    """Function names declared in include/farrsight_hip.h."""
    import re

    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fs_[a-z0-9_]+)\s*\(", text)))


class Server:
    """Serving loop with continuous batching (fs_server). The context must not
    be used by other calls while the server runs."""

    def __init__(self, ctx: "Context", hbm_budget_bytes: float = 200e9, max_active: int = 0,
                 precision: int = PRECISION_MIXED, use_graph: bool = True):
        self.ctx = ctx
        p = ServerParams(float(hbm_budget_bytes), int(max_active), int(precision),
                         int(bool(use_graph)))
        h = _vp()
        _check(lib().fs_server_create(ctx.h, ctypes.byref(p), ctypes.byref(h)),
               "fs_server_create")
        self.h = h
        self.n_species = {}
        ctx._adopt(self)

    def submit(self, sim: "SimParams", species, n_steps: int) -> int:
        d, arr = _deck(sim, species)
        rid = _i64()
        _check(lib().fs_server_submit(self.h, ctypes.byref(d), int(n_steps), ctypes.byref(rid)),
               "fs_server_submit")
        self.n_species[rid.value] = len(species)
        return rid.value

    def status(self, rid: int):
        st, done = _int(), _int()
        _check(lib().fs_server_status(self.h, rid, ctypes.byref(st), ctypes.byref(done)),
               "fs_server_status")
        return st.value, done.value

    def wait(self, rid: int, timeout_s: float = 0.0):
        _check(lib().fs_server_wait(self.h, rid, float(timeout_s)), "fs_server_wait")

    def result(self, rid: int, k: int):
        """(xs, ps, fs, q_ws, es, diagnostics) of species k of a finished request."""
        n = _i64()
        _check(lib().fs_server_result(self.h, rid, k, ctypes.byref(n), None, None, None, None,
                                      None, None), "fs_server_result")
        arrs = [np.empty(n.value) for _ in range(5)]
        diag = np.empty(4)
        _check(lib().fs_server_result(self.h, rid, k, None, *[_P(a) for a in arrs], _P(diag)),
               "fs_server_result")
        return (*arrs, diag)

    def write_result(self, rid: int, sim_dir: str, names):
        """A finished request's final state in the reference's files."""
        p, keep = _names(names)
        _check(lib().fs_server_write_result(self.h, rid, sim_dir.encode(), p),
               "fs_server_write_result")

    def submit_restart(self, sim: "SimParams", species, sim_dir: str, names, it: int,
                       n_steps: int) -> int:
        """A request resuming a deck from the files of step `it`."""
        d, arr = _deck(sim, species)
        p, keep = _names(names)
        rid = _i64()
        _check(lib().fs_server_submit_restart(self.h, ctypes.byref(d), sim_dir.encode(), p, it,
                                              int(n_steps), ctypes.byref(rid)),
               "fs_server_submit_restart")
        self.n_species[rid.value] = len(species)
        return rid.value

    def stats(self) -> dict:
        st = ServerStats()
        _check(lib().fs_server_get_stats(self.h, ctypes.byref(st)), "fs_server_get_stats")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            lib().fs_server_destroy(self.h)
            self.h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass
