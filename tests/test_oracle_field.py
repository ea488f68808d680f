"""Pins the oracle's field solvers (part A).

* Bit-exact against the reference's own E_MQ_DirectSum / E_MQ_Treecode
  (tests/golden/field_golden.npz, written by tests/golden/make_golden.py from
  src/FieldStructure.cpp compiled unmodified) and, when oracle/_ref is built,
  against fresh seeded inputs.
* Independently of the reference's expression: the algebraically rearranged
  single-source kernel, its zeros and oddness, the empty field of a uniform
  lattice, the Bessel-K1 closed form of a sinusoidal charge density, and the
  treecode's single-leaf identity (treecode = sqrt(1+4 delta^2) direct sum,
  src/FieldStructure.cpp:949 vs :150) and convergence in degree.
"""
import numpy as np
import pytest
from scipy.special import k1

import oracle

L0 = 4 * np.pi


def test_direct_sum_matches_reference_golden(golden):
    cases = [k for k in golden if k.startswith("ds_")]
    assert len(cases) >= 5
    for name in cases:
        c = golden[name]
        L, eps = c["params"]
        e = oracle.direct_sum(c["t"], c["s"], c["w"], L, eps)
        assert np.array_equal(e, c["es"]), name


def test_treecode_matches_reference_golden(golden):
    cases = [k for k in golden if k.startswith("tc_")]
    assert len(cases) >= 4
    for name in cases:
        c = golden[name]
        L, eps, mac, deg, ms = c["params"]
        e = oracle.treecode(c["s"], c["w"], L, eps, mac, int(deg), int(ms))
        assert np.array_equal(e, c["es"]), name


@pytest.mark.skipif(oracle.ref_lib(required=False) is None, reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_matches_live_reference(seed):
    rng = np.random.default_rng(seed)
    n = 1000 + 137 * seed
    x = rng.uniform(-2, L0 + 2, n)
    w = rng.normal(size=n)
    eps = [0.01, 0.05, 0.3][seed - 1]
    assert np.array_equal(oracle.direct_sum(x, x, w, L0, eps),
                          oracle.ref_direct_sum(x, x, w, L0, eps))
    for mac, deg, ms in [(0.7, 8, 100), (0.4, 3, 37)]:
        assert np.array_equal(oracle.treecode(x, w, L0, eps, mac, deg, ms),
                              oracle.ref_treecode(x, w, L0, eps, mac, deg, ms))


def kernel_closed_form(z, delta):
    """The reference kernel rewritten: z/(2 sqrt(z^2+d^2)) - z/sqrt(1+4 d^2)."""
    return z / (2 * np.sqrt(z * z + delta * delta)) - z / np.sqrt(1 + 4 * delta * delta)


@pytest.mark.parametrize("eps", [0.01, 0.05, 0.2])
def test_single_source_closed_form(eps):
    rng = np.random.default_rng(5)
    xs = np.array([1.3])
    t = rng.uniform(-L0, 2 * L0, 2000)
    e = oracle.direct_sum(t, xs, np.array([0.7]), L0, eps)
    z = (t - xs[0]) / L0
    z = z - np.round(z)
    ref = 0.7 * kernel_closed_form(z, eps / L0)
    assert np.max(np.abs(e - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_kernel_zeros_and_oddness():
    eps = 0.05
    s = np.array([2.0])
    w = np.array([1.0])
    # z = 0 and z = +-1/2 contribute exactly nothing (src/FieldStructure.cpp:135-137)
    e = oracle.direct_sum(np.array([2.0, 2.0 + L0 / 2, 2.0 - L0 / 2]), s, w, L0, eps)
    assert np.all(e == 0.0)
    # about a source at the origin the kernel is odd bit for bit
    d = np.linspace(0.01, 6.0, 301)
    s0 = np.array([0.0])
    ep = oracle.direct_sum(d, s0, w, L0, eps)
    em = oracle.direct_sum(-d, s0, w, L0, eps)
    assert np.array_equal(ep, -em)


def test_uniform_lattice_has_no_field():
    m = 512
    x = np.arange(m) * (L0 / m)
    e = oracle.direct_sum(x, x, np.full(m, 1.0 / m), L0, 0.05)
    assert np.max(np.abs(e)) < 1e-13


@pytest.mark.parametrize("eps,kmode", [(0.05, 1), (0.2, 2), (0.5, 3)])
def test_sinusoidal_density_bessel_closed_form(eps, kmode):
    """rho = A cos(kx) on a fine grid: E(x) = (A/k) sin(kx) * k eps K1(k eps)
    (the regularized kernel is the sawtooth smoothed by eps^2/(2(u^2+eps^2)^1.5),
    whose Fourier transform is |k| eps K1(|k| eps))."""
    m = 40000
    h = L0 / m
    x = (np.arange(m) + 0.5) * h
    k = 2 * np.pi / L0 * kmode
    A = 0.3
    w = A * np.cos(k * x) * h
    t = np.linspace(0.1, L0 - 0.1, 37)
    e = oracle.direct_sum(t, x, w, L0, eps)
    expect = A / k * np.sin(k * t) * (k * eps * k1(k * eps))
    # the kernel acts on the wrapped separation only: the tails of the smoothing
    # kernel beyond L/2 change the factor by O((eps/L)^2)
    tol = 5 * (eps / L0) ** 2 + 1e-7
    assert np.max(np.abs(e - expect)) <= tol * np.max(np.abs(expect))
    # and an eps twice as large would be far off: the test resolves the regularization
    wrong = A / k * np.sin(k * t) * (2 * k * eps * k1(2 * k * eps))
    assert np.max(np.abs(e - wrong)) > 20 * tol * np.max(np.abs(expect))


def test_linearity_and_periodic_translation():
    rng = np.random.default_rng(9)
    n = 800
    x = rng.uniform(0, L0, n)
    w1, w2 = rng.normal(size=n), rng.normal(size=n)
    e1 = oracle.direct_sum(x, x, w1, L0, 0.05)
    e2 = oracle.direct_sum(x, x, w2, L0, 0.05)
    e12 = oracle.direct_sum(x, x, w1 - 3 * w2, L0, 0.05)
    scale = np.max(np.abs(e1)) + 3 * np.max(np.abs(e2))
    assert np.max(np.abs(e12 - (e1 - 3 * e2))) < 1e-13 * scale
    eL = oracle.direct_sum(x + L0, x, w1, L0, 0.05)
    assert np.max(np.abs(eL - e1)) < 1e-13 * np.max(np.abs(e1))


def test_reference_sum_error_vs_extended():
    """The reference's own FP64 sum sits ~1e-14 (normwise) from the extended-
    precision value; that is the noise floor every FP64 comparison inherits."""
    rng = np.random.default_rng(11)
    n = 3000
    u = rng.uniform(0, L0, n)
    x = np.mod(u + 0.05 * np.sin(2 * np.pi / L0 * u), L0)
    v = rng.uniform(-6, 6, n)
    w = np.where(np.arange(n) < n // 2, -1.0, 1.0) * np.exp(-v * v / 2) / np.sqrt(2 * np.pi) * 1e-4
    e = oracle.direct_sum(x, x, w, L0, 0.05)
    ex = oracle.direct_sum_ext(x, x, w, L0, 0.05)
    rel = np.max(np.abs(e - ex)) / np.max(np.abs(ex))
    assert 0 < rel < 1e-12


def test_treecode_single_leaf_identity():
    rng = np.random.default_rng(12)
    n = 600
    x = rng.uniform(0, L0, n)
    w = rng.normal(size=n) * 1e-3
    eps = 0.05
    ds = oracle.direct_sum(x, x, w, L0, eps)
    tc = oracle.treecode(x, w, L0, eps, 0.7, 8, n + 1)
    st = oracle.treecode_stats()
    assert st["leaves"] == 1 and st["far_evals"] == 0
    factor = np.sqrt(1 + 4 * (eps / L0) ** 2)
    assert np.max(np.abs(tc - factor * ds)) < 1e-13 * np.max(np.abs(ds))


def test_treecode_converges_in_degree():
    rng = np.random.default_rng(13)
    n = 4000
    x = rng.uniform(0, L0, n)
    w = rng.normal(size=n) * 1e-3
    eps = 0.05
    exact = np.sqrt(1 + 4 * (eps / L0) ** 2) * oracle.direct_sum_ext(x, x, w, L0, eps)
    errs = []
    for deg in (2, 4, 8):
        tc = oracle.treecode(x, w, L0, eps, 0.5, deg, 64)
        assert oracle.treecode_stats()["far_evals"] > 0
        errs.append(np.max(np.abs(tc - exact)) / np.max(np.abs(exact)))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-6
