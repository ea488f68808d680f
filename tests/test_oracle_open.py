"""Pins the oracle's open-boundary field and external drive fields.

* Bit-exact against the reference's own E_MQ_DirectSum_openbcs and
  ExternalTanh / ExternalSine / ExternalLogistic (tests/golden/open_golden.npz,
  written by tests/golden/make_golden.py from src/FieldStructure.cpp compiled
  unmodified) and, when oracle/_ref is built, on fresh seeded inputs.
* Independently of the reference's expression: the single-source kernel in
  rearranged form, oddness, the eps -> 0 limit (Gauss's law in one dimension:
  E(x) = (Q left of x - Q right of x)/2 - x), the extended-precision sum, and
  the drive amplitudes' closed values (tanh drive zero at t = 0, sine ramp at
  its plateau and after it, logistic ramp's limits).
"""
import numpy as np
import pytest

import oracle


@pytest.fixture(scope="module")
def open_golden():
    from conftest import load_golden

    return load_golden("open_golden.npz")


def test_open_sum_matches_reference_golden(open_golden):
    cases = [k for k in open_golden if k.startswith("op_")]
    assert len(cases) >= 3
    for name in cases:
        c = open_golden[name]
        e = oracle.direct_sum_open(c["t"], c["s"], c["w"], float(c["params"][0]))
        assert np.array_equal(e, c["es"]), name


def test_external_fields_match_reference_golden(open_golden):
    cases = [k for k in open_golden if k.startswith("ex_")]
    assert len(cases) == 21
    for name in cases:
        c = open_golden[name]
        t_, typ, Am, k, om = c["params"]
        e = oracle.external_field_add(c["es0"], c["t"], t_, int(typ), Am, k, om)
        assert np.array_equal(e, c["es"]), name


@pytest.mark.parametrize("seed", [0, 1])
def test_matches_live_reference(seed):
    if oracle.ref_lib(required=False) is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(seed)
    s = rng.uniform(-3, 9, 700)
    t = np.concatenate([s[:300], rng.uniform(-5, 11, 100)])
    w = rng.normal(size=700) * 1e-3
    assert np.array_equal(oracle.direct_sum_open(t, s, w, 0.05),
                          oracle.ref_direct_sum_open(t, s, w, 0.05))
    es0 = rng.normal(size=t.size)
    for typ in (0, 1, 2, 7):
        for tm in rng.uniform(0, 420, 5):
            a = oracle.external_field_add(es0, t, tm, typ, 0.3, 0.4, 0.5)
            b = oracle.ref_external_field_add(es0, t, tm, typ, 0.3, 0.4, 0.5)
            assert np.array_equal(a, b), (typ, tm)


@pytest.mark.parametrize("eps", [0.01, 0.05, 0.3])
def test_single_source_closed_form(eps):
    t = np.linspace(-3, 3, 61)
    w = 0.37
    e = oracle.direct_sum_open(t, [0.25], [w], eps)
    z = t - 0.25
    expect = w * np.sign(z) / (2 * np.sqrt(1 + (eps / np.where(z == 0, 1, z)) ** 2)) - t
    expect[z == 0] = -t[z == 0]
    np.testing.assert_allclose(e, expect, rtol=1e-14, atol=1e-15)


def test_oddness():
    # sources symmetric about 0 with equal weights: no field from them at 0
    s = np.array([-1.5, -0.2, 0.2, 1.5])
    e = oracle.direct_sum_open([0.0], s, np.ones(4), 0.1)
    assert abs(e[0]) < 1e-16


def test_gauss_law_limit():
    rng = np.random.default_rng(3)
    s = np.sort(rng.uniform(-4, 4, 400))
    w = rng.normal(size=400) * 1e-2
    # targets at least 1e-3 from every source
    t = np.array([(s[i] + s[i + 1]) / 2 for i in range(0, 399, 7) if s[i + 1] - s[i] > 2e-3])
    eps = 1e-9
    e = oracle.direct_sum_open(t, s, w, eps)
    left = np.array([w[s < x].sum() for x in t])
    right = np.array([w[s > x].sum() for x in t])
    np.testing.assert_allclose(e, (left - right) / 2 - t, rtol=0, atol=1e-12)


def test_error_vs_extended_precision():
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 4 * np.pi, 3000)
    w = rng.normal(size=3000) * 1e-3
    e = oracle.direct_sum_open(x, x, w, 0.05)
    ex = oracle.direct_sum_open_ext(x, x, w, 0.05)
    assert np.max(np.abs(e - ex)) <= 1e-13 * np.max(np.abs(ex + x))


def test_drive_amplitudes():
    amp = oracle.external_amplitude
    assert amp(0, 0.0, 0.4, 0.26) == 0.0                # tanh drive starts at zero
    assert amp(0, 188.0, 0.4, 0.26) == pytest.approx(0.4 * 0.26, rel=1e-4)  # plateau (x k)
    assert amp(1, 100.0, 0.4, 0.26) == 0.4              # sine ramp plateau
    assert amp(1, 25.0, 0.4, 0.26) == pytest.approx(0.4 * np.sin(np.pi / 4), rel=1e-15)
    assert amp(1, 200.0, 0.4, 0.26) == 0.0              # after the ramp down
    assert amp(2, 60.0 - 1e-9, 0.4, 0.26) == pytest.approx(0.4, rel=1e-15)
    assert amp(2, 0.0, 0.4, 0.26) < 1e-170
    assert amp(2, 400.0, 0.4, 0.26) == 0.0
    assert amp(5, 30.0, 0.4, 0.26) == amp(0, 30.0, 0.4, 0.26)  # other types: tanh
    x = np.linspace(0, 10, 11)
    e = oracle.external_field_add(np.zeros(11), x, 100.0, 1, 0.4, 0.26, 0.37)
    np.testing.assert_allclose(e, 0.4 * np.sin(0.26 * x - 37.0), rtol=0, atol=1e-16)
