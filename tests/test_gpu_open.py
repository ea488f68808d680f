"""Open-boundary field (fs_emq_open_f64) and external drives
(fs_external_field_add) on the GPU against the oracle.

Bars (FP64): the open-boundary sum within 1e-12 of max|sum| of the oracle's
extended-precision sum, beyond two roundings at the magnitude of x (the bar is
on the pair sum itself, not on the -x background that dominates es; the oracle's FP64 sum, the reference's own order, is
printed beside it and is the less accurate of the two on decks whose charges
nearly cancel), for targets = sources, distinct targets, ragged sizes, a few
targets against many sources (source splits), and within 1e-12 of the
committed reference fixtures; the drive within 4.5e-16 (|a| + max|es|) of the
oracle's (the device sine against the host's: one ulp of the sine, which can
tip the rounding of the sum by one more).
"""
import numpy as np
import pytest

import multi_species_amd as ms
import oracle

pytestmark = pytest.mark.gpu


def _check_open(ctx, t, s, w, eps, tag):
    print(f"open-bc case {tag}: nt={len(t)} ns={len(s)} eps={eps}", flush=True)
    e = ctx.emq_open_f64(t, s, w, eps)
    ext = oracle.direct_sum_open_ext(t, s, w, eps)
    orc = oracle.direct_sum_open(t, s, w, eps)
    t = np.asarray(t, dtype=np.float64)
    scale = max(np.max(np.abs(ext + t)), 1e-300)
    # each result holds sum - x rounded once to FP64: beyond the bar on the
    # sum, two roundings at the magnitude of x
    slack = 4.5e-16 * np.abs(t)
    err = np.max(np.maximum(np.abs(e - ext) - slack, 0.0)) / scale
    err_orc = np.max(np.maximum(np.abs(orc - ext) - slack, 0.0)) / scale
    print(f"  rel err {err:.2e} (oracle FP64 {err_orc:.2e}) beyond the rounding of -x",
          flush=True)
    assert err <= 1e-12, (tag, err)


def test_open_matches_reference_fixtures(gpu_ctx):
    from conftest import load_golden

    g = load_golden("open_golden.npz")
    for name, c in g.items():
        if name.startswith("op_"):
            eps = float(c["params"][0])
            print(f"fixture {name}", flush=True)
            e = gpu_ctx.emq_open_f64(c["t"], c["s"], c["w"], eps)
            scale = np.max(np.abs(c["es"] + c["t"]))
            bound = 1e-12 * scale + 4.5e-16 * np.abs(c["t"])
            assert np.all(np.abs(e - c["es"]) <= bound), name
        elif name.startswith("ex_"):
            t_, typ, Am, k, om = c["params"]
            e = gpu_ctx.external_field_add(c["es0"], c["t"], t_, int(typ), Am, k, om)
            tol = 4.5e-16 * (abs(oracle.external_amplitude(int(typ), t_, Am, k)) +
                           np.max(np.abs(c["es"])))
            assert np.max(np.abs(e - c["es"])) <= tol, name


@pytest.mark.parametrize("n", [1, 63, 1000, 65537])
def test_open_targets_are_sources(gpu_ctx, n):
    rng = np.random.default_rng(n)
    x = rng.uniform(0, 4 * np.pi, n)
    w = np.where(np.arange(n) < n // 2, -1.0, 1.0) * np.exp(-rng.uniform(-6, 6, n) ** 2 / 2) * 1e-4
    _check_open(gpu_ctx, x, x, w, 0.05, f"same n={n}")


@pytest.mark.parametrize("n", [1 << 18, 1 << 20])
def test_open_pair_once_at_benchmark_size(gpu_ctx, n):
    """The benchmarked configuration (targets = sources, four targets per
    lane, N = 2^18 and 2^20): sampled targets against the extended oracle
    over all sources, and the whole field against the full sum over ordered
    pairs (fs_emq_set_symmetric(1)) within 1e-12 of max|sum|."""
    rng = np.random.default_rng(n + 3)
    x = rng.uniform(0, 4 * np.pi, n)
    w = np.where(np.arange(n) < n // 2, -1.0, 1.0) * np.exp(-rng.uniform(-6, 6, n) ** 2 / 2) * 1e-4
    print(f"open-bc pair-once n={n}", flush=True)
    e = gpu_ctx.emq_open_f64(x, x, w, 0.05)
    assert gpu_ctx.f64_pair_evals() < n * n  # the pair-once form ran
    pick = np.sort(rng.choice(n, 128, replace=False))
    ext = oracle.direct_sum_open_ext(x[pick], x, w, 0.05)
    scale = np.max(np.abs(e + x))
    slack = 4.5e-16 * np.abs(x[pick])
    assert np.max(np.maximum(np.abs(e[pick] - ext) - slack, 0.0)) <= 1e-12 * scale
    gpu_ctx.set_symmetric(1)
    try:
        full = gpu_ctx.emq_open_f64(x, x, w, 0.05)
        assert gpu_ctx.f64_pair_evals() == n * n
    finally:
        gpu_ctx.set_symmetric(0)
    assert np.max(np.maximum(np.abs(e - full) - 4.5e-16 * np.abs(x), 0.0)) <= 1e-12 * scale


def test_open_distinct_and_split(gpu_ctx):
    rng = np.random.default_rng(11)
    s = rng.uniform(-10, 20, 200000)
    w = rng.normal(size=s.size) * 1e-5
    _check_open(gpu_ctx, rng.uniform(-12, 22, 37), s, w, 0.01, "few targets, many sources")
    _check_open(gpu_ctx, rng.uniform(-12, 22, 3001), s[:5000], w[:5000], 0.2, "ragged")
    # no sources: only the background
    t = rng.uniform(0, 1, 10)
    e = gpu_ctx.emq_open_f64(t, np.zeros(0), np.zeros(0), 0.1)
    assert np.array_equal(e, -t)


def test_open_rejects_bad_arguments(gpu_ctx):
    with pytest.raises(ms.FarrsightError):
        gpu_ctx.emq_open_f64([0.0], [1.0], [1.0], 0.0)


@pytest.mark.parametrize("typ", [0, 1, 2])
def test_external_field_matches_oracle(gpu_ctx, typ):
    rng = np.random.default_rng(typ)
    x = rng.uniform(0, 10 * np.pi, 100003)
    es0 = rng.normal(size=x.size) * 1e-2
    for t_ in (0.0, 5.0, 61.0, 149.0, 170.0, 300.0):
        print(f"external type {typ} t={t_}", flush=True)
        e = gpu_ctx.external_field_add(es0, x, t_, typ, 0.4, 0.26, 0.37)
        ref = oracle.external_field_add(es0, x, t_, typ, 0.4, 0.26, 0.37)
        a = abs(oracle.external_amplitude(typ, t_, 0.4, 0.26))
        assert np.max(np.abs(e - ref)) <= 4.5e-16 * (a + np.max(np.abs(es0))) + 1e-300, t_


def test_device_entry_points_match_host(gpu_ctx):
    """fs_emq_open_f64_dev (targets = sources: the pair-once form) and
    fs_external_field_add_dev on device arrays, bitwise equal to the host
    entry points (the same kernels and plans). Device memory comes from the
    HIP runtime the library itself uses."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    h2d, d2h = 1, 2

    def upload(a):
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, h2d) == 0
        return p

    rng = np.random.default_rng(21)
    n = 20000
    x = rng.uniform(0, 10, n)
    w = rng.normal(size=n) * 1e-4
    host = gpu_ctx.emq_open_f64(x, x, w, 0.05)
    dx, dw, de = upload(x), upload(w), upload(np.zeros(n))
    try:
        print("open-bc dev entry point, n=20000", flush=True)
        gpu_ctx.emq_open_f64_dev(de.value, dx.value, n, dx.value, dw.value, n, 0.05)
        gpu_ctx.synchronize()
        out = np.empty(n)
        assert hip.hipMemcpy(out.ctypes.data, de, out.nbytes, d2h) == 0
        assert np.array_equal(out, host)
        print("external drive dev entry point", flush=True)
        want = gpu_ctx.external_field_add(host, x, 61.0, 2, 0.4, 0.26, 0.37)
        assert ms.lib().fs_external_field_add_dev(gpu_ctx.h, de, dx, n, 61.0, 2, 0.4, 0.26, 0.37,
                                                  None) == 0
        gpu_ctx.synchronize()
        assert hip.hipMemcpy(out.ctypes.data, de, out.nbytes, d2h) == 0
        assert np.array_equal(out, want)
    finally:
        for p in (dx, dw, de):
            hip.hipFree(p)
