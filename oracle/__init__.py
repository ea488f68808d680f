"""Test infrastructure only: ctypes bindings of the CPU oracle.

``liboracle.so`` is this repository's FP64 restatement of the FARRSIGHT hot
path (``oracle/fs_oracle_*.cpp``); ``_ref/libfsref.so`` is the reference's own
``src/FieldStructure.cpp`` compiled behind ``oracle/ref_harness.cpp``. Only the
tests, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import this
module, as the checker; the product never does.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.c_int64
_d = ctypes.c_double

_lib = None
_ref = None


def _P(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _PI(a: np.ndarray):
    return a.ctypes.data_as(_ip)


class SpeciesParams(ctypes.Structure):
    _fields_ = [
        ("q", _d), ("m", _d), ("p_min", _d), ("p_max", _d),
        ("initial_height", ctypes.c_int), ("p_height", ctypes.c_int),
        ("max_height", ctypes.c_int), ("do_adaptively_refine", ctypes.c_int),
        ("n_amr_epsilons", ctypes.c_int), ("amr_epsilons", _d * 5),
        ("ics_type", ctypes.c_int), ("normalized_wavenumber", _d), ("amp", _d),
        ("pth", _d), ("pstr", _d), ("beam_frac", _d), ("beam_v", _d),
        ("beam_vth", _d),
    ]


class SimParams(ctypes.Structure):
    _fields_ = [
        ("x_min", _d), ("x_max", _d), ("bcs", ctypes.c_int),
        ("quadrature", ctypes.c_int), ("greens_epsilon", _d),
        ("use_treecode", ctypes.c_int), ("mac", _d), ("degree", ctypes.c_int),
        ("max_source", ctypes.c_int), ("dt", _d), ("remesh_period", ctypes.c_int),
        ("use_external_field", ctypes.c_int), ("external_type", ctypes.c_int), ("Am", _d),
        ("k", _d), ("omega", _d),
    ]


def lib():
    """Loads (building it first if needed) oracle/liboracle.so."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle oracle`")
        L = ctypes.CDLL(path)
        L.orc_direct_sum.argtypes = [_dp, _dp, _i64, _dp, _dp, _i64, _d, _d]
        L.orc_direct_sum_ext.argtypes = [_dp, _dp, _i64, _dp, _dp, _i64, _d, _d]
        L.orc_treecode.argtypes = [_dp, _dp, _dp, _i64, _d, _d, _d, ctypes.c_int, ctypes.c_int]
        L.orc_direct_sum_open.argtypes = [_dp, _dp, _i64, _dp, _dp, _i64, _d]
        L.orc_direct_sum_open_ext.argtypes = [_dp, _dp, _i64, _dp, _dp, _i64, _d]
        L.orc_external_field_add.argtypes = [_dp, _dp, _i64, _d, ctypes.c_int, _d, _d, _d]
        L.orc_external_amplitude.argtypes = [ctypes.c_int, _d, _d, _d]
        L.orc_external_amplitude.restype = _d
        L.orc_treecode_last_stats.argtypes = [ctypes.POINTER(ctypes.c_int64)]
        _ip = ctypes.POINTER(ctypes.c_int32)
        L.orc_treecode_tree.argtypes = [_dp, _i64, _d, _d, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int64), _ip, _ip, _ip, _ip, _ip,
                                        _ip]
        L.orc_species_create.restype = ctypes.c_void_p
        L.orc_species_create.argtypes = [ctypes.POINTER(SimParams), ctypes.POINTER(SpeciesParams)]
        for n in ("destroy", "copy_to_old", "remesh"):
            getattr(L, f"orc_species_{n}").argtypes = [ctypes.c_void_p]
        for n in ("num_points", "num_panels", "num_leaves"):
            f = getattr(L, f"orc_species_{n}")
            f.argtypes = [ctypes.c_void_p]
            f.restype = ctypes.c_int64
        L.orc_species_get_points.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_species_set_points.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_species_get_panels.argtypes = [ctypes.c_void_p, _ip]
        L.orc_species_get_leaves.argtypes = [ctypes.c_void_p, _ip]
        L.orc_species_f_beyond_boundary.argtypes = [ctypes.c_void_p]
        L.orc_species_f_beyond_boundary.restype = _d
        L.orc_species_interp_from_mesh.argtypes = [ctypes.c_void_p, _dp, _dp, _i64, _dp]
        L.orc_species_interp_to_grid.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int,
                                                 ctypes.c_int, _dp, _ip, _ip]
        L.orc_sim_create.restype = ctypes.c_void_p
        L.orc_sim_create.argtypes = [ctypes.POINTER(SimParams), ctypes.POINTER(SpeciesParams),
                                     ctypes.c_int]
        L.orc_sim_destroy.argtypes = [ctypes.c_void_p]
        L.orc_sim_step.argtypes = [ctypes.c_void_p]
        L.orc_sim_num_species.argtypes = [ctypes.c_void_p]
        L.orc_sim_species.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_sim_species.restype = ctypes.c_void_p
        L.orc_sim_get_es.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_sim_diagnostics.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_sim_diagnostics_full.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_diagnostics.argtypes = [_i64, _dp, _dp, _dp, _dp, _dp, _d, _d, _d, _dp]
        L.orc_species_set_fs.argtypes = [ctypes.c_void_p, _dp]
        L.orc_species_descend_points.argtypes = [ctypes.c_void_p, _dp, _dp, _i64, _ip]
        L.orc_sim_iter.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def ref_lib(required: bool = True):
    """The reference's own compiled field solver, or None when absent."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libfsref.so")
        if not os.path.exists(path):
            if required:
                raise FileNotFoundError(f"{path} missing: run `make -C oracle ref`")
            return None
        R = ctypes.CDLL(path)
        R.fsref_direct_sum.argtypes = [_dp, _dp, ctypes.c_int, _dp, _dp, ctypes.c_int, _d, _d]
        R.fsref_direct_sum_open.argtypes = [_dp, _dp, ctypes.c_int, _dp, _dp, ctypes.c_int, _d]
        R.fsref_external_field_add.argtypes = [_dp, _dp, ctypes.c_int, _d, ctypes.c_int, _d, _d,
                                               _d]
        R.fsref_treecode.argtypes = [_dp, _dp, _dp, ctypes.c_int, _d, _d, _d, ctypes.c_int,
                                     ctypes.c_int]
        _ref = R
    return _ref


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- field

def direct_sum(targets, sources, q_ws, L: float, eps: float) -> np.ndarray:
    t, s, w = _f64(targets), _f64(sources), _f64(q_ws)
    es = np.empty(t.size)
    rc = lib().orc_direct_sum(_P(es), _P(t), t.size, _P(s), _P(w), s.size, L, eps)
    assert rc == 0
    return es


def direct_sum_ext(targets, sources, q_ws, L: float, eps: float) -> np.ndarray:
    t, s, w = _f64(targets), _f64(sources), _f64(q_ws)
    es = np.empty(t.size)
    rc = lib().orc_direct_sum_ext(_P(es), _P(t), t.size, _P(s), _P(w), s.size, L, eps)
    assert rc == 0
    return es


def treecode(sources, q_ws, L: float, eps: float, mac: float, degree: int,
             max_source: int) -> np.ndarray:
    s, w = _f64(sources), _f64(q_ws)
    es = np.empty(s.size)
    rc = lib().orc_treecode(_P(es), _P(s), _P(w), s.size, L, eps, mac, degree, max_source)
    assert rc == 0
    return es


def treecode_tree(sources, L: float, mac: float, max_source: int):
    """E_MQ_Treecode's tree and interaction lists (orc_treecode_tree), in the
    layout of the product's treecode_tree; None when more than max_source
    points share a coordinate."""
    s = _f64(sources)
    st = (ctypes.c_int64 * 4)()
    rc = lib().orc_treecode_tree(_P(s), s.size, L, mac, max_source, st, None, None, None, None,
                                 None, None)
    if rc == -2:
        return None
    assert rc == 0
    nn, nl, nf, nr = (int(v) for v in st)
    out = dict(stats=(nn, nl, nf, nr), perm=np.empty(s.size, np.int32),
               leaves=np.empty((nl, 3), np.int32), far_off=np.empty(nl + 1, np.int32),
               far=np.empty((max(nf, 1), 3), np.int32), near_off=np.empty(nl + 1, np.int32),
               near=np.empty((max(nr, 1), 3), np.int32))
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    rc = lib().orc_treecode_tree(_P(s), s.size, L, mac, max_source, st, ip(out["perm"]),
                                 ip(out["leaves"]), ip(out["far_off"]), ip(out["far"]),
                                 ip(out["near_off"]), ip(out["near"]))
    assert rc == 0
    out["far"], out["near"] = out["far"][:nf], out["near"][:nr]
    return out


def treecode_stats():
    out = (ctypes.c_int64 * 4)()
    lib().orc_treecode_last_stats(out)
    return dict(nodes=out[0], leaves=out[1], far_evals=out[2], near_evals=out[3])


def ref_direct_sum(targets, sources, q_ws, L: float, eps: float) -> np.ndarray:
    t, s, w = _f64(targets), _f64(sources), _f64(q_ws)
    es = np.empty(t.size)
    rc = ref_lib().fsref_direct_sum(_P(es), _P(t), t.size, _P(s), _P(w), s.size, L, eps)
    assert rc == 0
    return es


def direct_sum_open(targets, sources, q_ws, eps: float) -> np.ndarray:
    """E_MQ_DirectSum_openbcs::operator() (src/FieldStructure.cpp:166-191)."""
    t, s, w = _f64(targets), _f64(sources), _f64(q_ws)
    es = np.empty(t.size)
    assert lib().orc_direct_sum_open(_P(es), _P(t), t.size, _P(s), _P(w), s.size, eps) == 0
    return es


def direct_sum_open_ext(targets, sources, q_ws, eps: float) -> np.ndarray:
    t, s, w = _f64(targets), _f64(sources), _f64(q_ws)
    es = np.empty(t.size)
    assert lib().orc_direct_sum_open_ext(_P(es), _P(t), t.size, _P(s), _P(w), s.size, eps) == 0
    return es


def external_field_add(es, targets, t: float, type_: int, Am: float, k: float,
                       omega: float) -> np.ndarray:
    """ExternalTanh/Sine/Logistic::operator() (src/FieldStructure.cpp:17-139) on a copy of es."""
    e, x = _f64(es).copy(), _f64(targets)
    assert lib().orc_external_field_add(_P(e), _P(x), x.size, t, type_, Am, k, omega) == 0
    return e


def external_amplitude(type_: int, t: float, Am: float, k: float) -> float:
    return float(lib().orc_external_amplitude(type_, t, Am, k))


def ref_direct_sum_open(targets, sources, q_ws, eps: float) -> np.ndarray:
    t, s, w = _f64(targets), _f64(sources), _f64(q_ws)
    es = np.empty(t.size)
    assert ref_lib().fsref_direct_sum_open(_P(es), _P(t), t.size, _P(s), _P(w), s.size, eps) == 0
    return es


def ref_external_field_add(es, targets, t: float, type_: int, Am: float, k: float,
                           omega: float) -> np.ndarray:
    e, x = _f64(es).copy(), _f64(targets)
    assert ref_lib().fsref_external_field_add(_P(e), _P(x), x.size, t, type_, Am, k,
                                              omega) == 0
    return e


def ref_treecode(sources, q_ws, L: float, eps: float, mac: float, degree: int,
                 max_source: int) -> np.ndarray:
    s, w = _f64(sources), _f64(q_ws)
    es = np.empty(s.size)
    rc = ref_lib().fsref_treecode(_P(es), _P(s), _P(w), s.size, L, eps, mac, degree,
                                  max_source)
    assert rc == 0
    return es


# ---------------------------------------------------------------- decks

@dataclass
class Species:
    name: str = "electrons"
    q: float = -1.0
    m: float = 1.0
    p_min: float = -1.0
    p_max: float = 1.0
    initial_height: int = 6
    p_height: int = 0
    max_height: Optional[int] = None
    do_adaptively_refine: bool = False
    amr_epsilons: Sequence[float] = ()
    ics_type: int = 1
    normalized_wavenumber: float = 1.0
    amp: float = 0.0
    pth: float = 1.0
    pstr: float = 0.0
    beam_frac: float = 0.0
    beam_v: float = 0.0
    beam_vth: float = 1.0

    def to_c(self) -> SpeciesParams:
        sp = SpeciesParams()
        sp.q, sp.m, sp.p_min, sp.p_max = self.q, self.m, self.p_min, self.p_max
        sp.initial_height, sp.p_height = self.initial_height, self.p_height
        sp.max_height = self.initial_height if self.max_height is None else self.max_height
        sp.do_adaptively_refine = int(bool(self.do_adaptively_refine))
        eps = list(self.amr_epsilons)[:5]
        sp.n_amr_epsilons = len(eps)
        for i, e in enumerate(eps):
            sp.amr_epsilons[i] = e
        sp.ics_type = self.ics_type
        sp.normalized_wavenumber, sp.amp = self.normalized_wavenumber, self.amp
        sp.pth, sp.pstr = self.pth, self.pstr
        sp.beam_frac, sp.beam_v, sp.beam_vth = self.beam_frac, self.beam_v, self.beam_vth
        return sp


@dataclass
class Deck:
    x_min: float = 0.0
    x_max: float = 4 * np.pi
    bcs: int = 0
    quadrature: int = 0
    greens_epsilon: float = 0.05
    use_treecode: int = 0
    mac: float = 0.7
    degree: int = 8
    max_source: int = 2000
    dt: float = 0.1
    remesh_period: int = 1
    use_external_field: int = 0
    external_type: int = 0
    Am: float = 0.0
    k: float = 0.0
    omega: float = 0.0
    species: List[Species] = field(default_factory=list)

    def to_c(self) -> SimParams:
        p = SimParams()
        p.x_min, p.x_max, p.bcs, p.quadrature = self.x_min, self.x_max, self.bcs, self.quadrature
        p.greens_epsilon, p.use_treecode = self.greens_epsilon, self.use_treecode
        p.mac, p.degree, p.max_source = self.mac, self.degree, self.max_source
        p.dt, p.remesh_period = self.dt, self.remesh_period
        p.use_external_field, p.external_type = self.use_external_field, self.external_type
        p.Am, p.k, p.omega = self.Am, self.k, self.omega
        return p


class OracleSpecies:
    """One species' mesh (AMRStructure) in the oracle."""

    PANEL_INTS = 20

    def __init__(self, deck: Deck, sp: Species, _handle=None, _owner=None):
        self._owner = _owner
        if _handle is not None:
            self.h = _handle
            self._own = False
        else:
            self._sim_c = deck.to_c()
            self._sp_c = sp.to_c()
            self.h = lib().orc_species_create(ctypes.byref(self._sim_c), ctypes.byref(self._sp_c))
            if not self.h:
                raise RuntimeError("oracle species creation failed")
            self._own = True

    def __del__(self):
        if getattr(self, "_own", False) and self.h:
            lib().orc_species_destroy(self.h)
            self.h = None

    @property
    def n(self) -> int:
        return lib().orc_species_num_points(self.h)

    def points(self):
        n = self.n
        xs, ps, fs, ws = (np.empty(n) for _ in range(4))
        lib().orc_species_get_points(self.h, _P(xs), _P(ps), _P(fs), _P(ws))
        return xs, ps, fs, ws

    def set_points(self, xs, ps):
        xs, ps = _f64(xs), _f64(ps)
        assert xs.size == self.n and ps.size == self.n
        lib().orc_species_set_points(self.h, _P(xs), _P(ps))

    def set_fs(self, fs):
        fs = _f64(fs)
        assert fs.size == self.n
        lib().orc_species_set_fs(self.h, _P(fs))

    def descend_points(self, x, p) -> np.ndarray:
        """The old-mesh leaf interp_from_mesh uses at each point (0: beyond
        a p boundary)."""
        x, p = _f64(x), _f64(p)
        out = np.empty(x.size, dtype=np.int32)
        if lib().orc_species_descend_points(self.h, _P(x), _P(p), x.size, _PI(out)) != 0:
            raise RuntimeError("too many shifts!")
        return out

    def panels(self) -> np.ndarray:
        n = lib().orc_species_num_panels(self.h)
        rec = np.empty((n, self.PANEL_INTS), dtype=np.int32)
        lib().orc_species_get_panels(self.h, _PI(rec))
        return rec

    def leaves(self) -> np.ndarray:
        n = lib().orc_species_num_leaves(self.h)
        out = np.empty(n, dtype=np.int32)
        lib().orc_species_get_leaves(self.h, _PI(out))
        return out

    @property
    def f_beyond_boundary(self) -> float:
        return lib().orc_species_f_beyond_boundary(self.h)

    def remesh(self):
        if lib().orc_species_remesh(self.h) != 0:
            raise RuntimeError("too many shifts!")

    def copy_to_old(self):
        lib().orc_species_copy_to_old(self.h)

    def interp_from_mesh(self, x, p) -> np.ndarray:
        x, p = _f64(x), _f64(p)
        out = np.empty(x.size)
        if lib().orc_species_interp_from_mesh(self.h, _P(x), _P(p), x.size, _P(out)) != 0:
            raise RuntimeError("too many shifts!")
        return out

    def interp_to_grid(self, xs, ps, nx: int, np_: int):
        xs, ps = _f64(xs), _f64(ps)
        assert xs.size == nx * np_
        f = np.empty(xs.size)
        leaf = np.empty(xs.size, dtype=np.int32)
        perm = np.empty(xs.size, dtype=np.int32)
        if lib().orc_species_interp_to_grid(self.h, _P(xs), _P(ps), nx, np_, _P(f), _PI(leaf),
                                            _PI(perm)) != 0:
            raise RuntimeError("too many shifts!")
        return f, leaf, perm


class OracleSim:
    """The whole FARRSIGHT step loop in the oracle."""

    def __init__(self, deck: Deck):
        self.deck = deck
        self._sim_c = deck.to_c()
        arr = (SpeciesParams * len(deck.species))(*[s.to_c() for s in deck.species])
        self._sp_c = arr
        self.h = lib().orc_sim_create(ctypes.byref(self._sim_c), arr, len(deck.species))
        if not self.h:
            raise RuntimeError("oracle simulation creation failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)
            self.h = None

    def step(self):
        if lib().orc_sim_step(self.h) != 0:
            raise RuntimeError("oracle step failed")

    @property
    def iter(self) -> int:
        return lib().orc_sim_iter(self.h)

    def species(self, i: int) -> OracleSpecies:
        return OracleSpecies(self.deck, self.deck.species[i],
                             _handle=lib().orc_sim_species(self.h, i), _owner=self)

    def es(self, i: int) -> np.ndarray:
        n = self.species(i).n
        out = np.empty(n)
        lib().orc_sim_get_es(self.h, i, _P(out))
        return out

    def diagnostics(self, i: int) -> np.ndarray:
        out = np.empty(4)
        lib().orc_sim_diagnostics(self.h, i, _P(out))
        return out

    def diagnostics_full(self, i: int) -> np.ndarray:
        out = np.empty(DIAG_COUNT)
        lib().orc_sim_diagnostics_full(self.h, i, _P(out))
        return out


DIAG_COUNT = 14
DIAG_NAMES = ("charge", "momentum", "kinetic", "electric", "entropy", "l1", "l2", "min_f",
              "max_f", "negative_area", "num_negative", "frac_negative", "negative_mass",
              "num_points")


def diagnostics(xs, ps, fs, q_ws, es, q: float, qm: float, L: float) -> np.ndarray:
    """The per-species diagnostic set of diagnostic_plot.py:116-143 on arrays."""
    a = [_f64(v) for v in (xs, ps, fs, q_ws, es)]
    n = a[0].size
    assert all(v.size == n for v in a)
    out = np.empty(DIAG_COUNT)
    lib().orc_diagnostics(n, *[_P(v) for v in a], q, qm, L, _P(out))
    return out


def landau_deck(initial_height: int = 5, p_height: int = 1, dt: float = 0.1,
                eps: float = 0.05, quadrature: int = 0) -> Deck:
    """Two-species weak Landau damping on [0, 4 pi) (electrons q=-1 m=1 with a
    1% cos(0.5x) perturbation; ions q=+1 m=1836 Maxwellian), p in [-6, 6]."""
    e = Species(name="electrons", q=-1.0, m=1.0, p_min=-6.0, p_max=6.0,
                initial_height=initial_height, p_height=p_height, ics_type=1,
                normalized_wavenumber=1.0, amp=0.01, pth=1.0)
    i = Species(name="ions", q=1.0, m=1836.0, p_min=-6.0, p_max=6.0,
                initial_height=initial_height, p_height=p_height, ics_type=1,
                normalized_wavenumber=1.0, amp=0.0, pth=1.0 / np.sqrt(1836.0))
    return Deck(x_min=0.0, x_max=4 * np.pi, bcs=0, quadrature=quadrature,
                greens_epsilon=eps, use_treecode=0, dt=dt, remesh_period=1,
                species=[e, i])
