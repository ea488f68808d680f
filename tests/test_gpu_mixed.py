"""GPU checks of the mixed-precision direct sum (fs_emq_periodic_mixed*).

The FP64 method is the judge: every case compares the mixed field with the
oracle's FP64 (or extended-precision) sum and requires
    max_i |E_i - E_i^FP64| / max_i |E_i^FP64| <= 1e-8.
Decks: the reference's golden inputs, ragged sizes, random-position decks
(half electron-, half ion-signed Maxwellian weights), lattice decks of the
two-species Landau problem (positions shifted by p*tau, where rounding errors
of any FP32 scheme add coherently), exact lattices with coincident and
antipodal points, unwrapped positions, large eps (every pair in the FP64 band),
and 2^20 points against sampled targets.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def mixed_kernels_at_every_size(gpu_ctx):
    """The cases below check the mixed kernels themselves, also at sizes
    where the entry points would otherwise return the FP64 sum."""
    import multi_species_amd as ms
    gpu_ctx.set_mixed_min_pairs(0)
    yield
    gpu_ctx.set_mixed_min_pairs(ms.MIXED_MIN_PAIRS_DEFAULT)

L0 = 4 * np.pi
TOL = 1e-8


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def random_deck(n, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(0, L0, n)
    x = np.mod(u + 0.05 * np.sin(2 * np.pi / L0 * u), L0)
    v = rng.uniform(-6, 6, n)
    q = np.where(np.arange(n) < n // 2, -1.0, 1.0)
    w = q * np.exp(-v * v / 2) / np.sqrt(2 * np.pi) * 1e-4
    return x, w


def lattice_deck(tau, mass_ratio, nx=64, np_=129, amp=0.01):
    """Electrons (amp cos(x/2) perturbation) and ions on a shared x-p lattice,
    each point moved by p/m * tau."""
    dx, dp = L0 / nx, 12.0 / (np_ - 1)
    xs, ws = [], []
    X, P = np.meshgrid(np.arange(nx) * dx, -6 + np.arange(np_) * dp, indexing="ij")
    X, P = X.ravel(), P.ravel()
    fe = (1 + amp * np.cos(0.5 * X)) / np.sqrt(2 * np.pi) * np.exp(-P * P / 2)
    vth = 1 / np.sqrt(mass_ratio)
    fi = np.exp(-P * P / (2 * vth * vth)) / (np.sqrt(2 * np.pi) * vth)
    for q, m, f in ((-1.0, 1.0, fe), (1.0, mass_ratio, fi)):
        xs.append(np.mod(X + P / m * tau, L0))
        ws.append(q * f * dx * dp)
    return np.concatenate(xs), np.concatenate(ws)


def test_golden_inputs(gpu_ctx, golden):
    for name, c in golden.items():
        if not name.startswith("ds_"):
            continue
        L, eps = c["params"]
        e = gpu_ctx.emq_periodic_mixed(c["t"], c["s"], c["w"], L, eps)
        print(name, "mixed-vs-ref", rel(e, c["es"]))
        assert rel(e, c["es"]) <= TOL, name


@pytest.mark.parametrize("n", [1, 2, 31, 32, 33, 63, 64, 65, 255, 257, 1000, 4097, 16770])
def test_ragged_sizes(gpu_ctx, n):
    x, w = random_deck(n, seed=n)
    e = gpu_ctx.emq_periodic_mixed(x, x, w, L0, 0.05, same=True)
    e2 = gpu_ctx.emq_periodic_mixed(x.copy(), x, w, L0, 0.05)  # separate target sort
    ex = oracle.direct_sum_ext(x, x, w, L0, 0.05)
    if n == 1:
        assert e[0] == 0.0 and e2[0] == 0.0
    else:
        assert rel(e, ex) <= TOL
        assert rel(e2, ex) <= TOL


@pytest.mark.parametrize("eps", [0.01, 0.05])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_deck_2e16(gpu_ctx, eps, seed):
    x, w = random_deck(1 << 16, seed)
    e = gpu_ctx.emq_periodic_mixed(x, x, w, L0, eps, same=True)
    ref = oracle.direct_sum(x, x, w, L0, eps)
    r = rel(e, ref)
    print(f"2^16 seed {seed} eps {eps}: rel {r:.2e}")
    assert r <= TOL


@pytest.mark.parametrize("tau,mass", [(0.0, 1836.0), (0.05, 1836.0), (0.3, 25.0), (0.7, 100.0)])
@pytest.mark.parametrize("eps", [0.01, 0.05, 0.2])
def test_lattice_decks(gpu_ctx, tau, mass, eps):
    x, w = lattice_deck(tau, mass)
    e = gpu_ctx.emq_periodic_mixed(x, x, w, L0, eps, same=True)
    ex = oracle.direct_sum_ext(x, x, w, L0, eps)
    r = rel(e, ex)
    print(f"lattice tau {tau} m {mass} eps {eps}: rel {r:.2e}")
    assert r <= TOL


def test_reduced_lattice_antipodal(gpu_ctx):
    """Unique x of a 64-column lattice: every point has an exact antipode."""
    x = np.arange(64) * (L0 / 64)
    w = np.cos(0.5 * x) * 0.1 + 0.01
    e = gpu_ctx.emq_periodic_mixed(x, x, w, L0, 0.05, same=True)
    ex = oracle.direct_sum_ext(x, x, w, L0, 0.05)
    assert rel(e, ex) <= TOL


def test_distinct_targets_unwrapped_and_empty(gpu_ctx):
    x, w = random_deck(3000, 7)
    t = np.linspace(-3 * L0, 4 * L0, 777)
    e = gpu_ctx.emq_periodic_mixed(t, x + L0 * np.round(np.sin(np.arange(3000))), w, L0, 0.05)
    ex = oracle.direct_sum_ext(t, x, w, L0, 0.05)
    assert rel(e, ex) <= TOL
    e0 = gpu_ctx.emq_periodic_mixed(t, np.empty(0), np.empty(0), L0, 0.05)
    assert np.all(e0 == 0.0)
    assert gpu_ctx.emq_periodic_mixed(np.empty(0), x, w, L0, 0.05).size == 0


def test_large_eps_all_fp64(gpu_ctx):
    x, w = random_deck(5000, 11)
    for eps in (0.5, 1.0, 3.0):
        e = gpu_ctx.emq_periodic_mixed(x, x, w, L0, eps, same=True)
        ex = oracle.direct_sum_ext(x, x, w, L0, eps)
        assert rel(e, ex) <= 1e-12, eps  # every pair takes the FP64 band


def test_coincident_points(gpu_ctx):
    rng = np.random.default_rng(5)
    x = np.repeat(rng.uniform(0, L0, 300), 7)
    w = rng.standard_normal(x.size) * 1e-3
    e = gpu_ctx.emq_periodic_mixed(x, x, w, L0, 0.05, same=True)
    ex = oracle.direct_sum_ext(x, x, w, L0, 0.05)
    assert rel(e, ex) <= TOL


def test_deterministic_and_tuning(gpu_ctx):
    x, w = random_deck(20000, 3)
    e1 = gpu_ctx.emq_periodic_mixed(x, x, w, L0, 0.05, same=True)
    e2 = gpu_ctx.emq_periodic_mixed(x, x, w, L0, 0.05, same=True)
    assert np.array_equal(e1, e2)
    ex = oracle.direct_sum_ext(x, x, w, L0, 0.05)
    try:
        for tpt, splits in ((1, 1), (2, 3), (4, 7), (1, 40)):
            gpu_ctx.set_tuning(tpt, splits)
            e = gpu_ctx.emq_periodic_mixed(x, x, w, L0, 0.05, same=True)
            assert rel(e, ex) <= TOL, (tpt, splits)
        # targets per thread is fixed in mixed precision: asking for another
        # changes no bit
        gpu_ctx.set_tuning(1, 3)
        e13 = gpu_ctx.emq_periodic_mixed(x, x, w, L0, 0.05, same=True)
        tpt = gpu_ctx.mixed_shape()[0]
        gpu_ctx.set_tuning(2, 3)
        assert np.array_equal(gpu_ctx.emq_periodic_mixed(x, x, w, L0, 0.05, same=True), e13)
        assert gpu_ctx.mixed_shape()[0] == tpt
    finally:
        gpu_ctx.set_tuning(0, 0)


def test_full_size_sampled(gpu_ctx):
    n = 1 << 20
    x, w = random_deck(n, 0)
    for eps in (0.01, 0.05):
        e = gpu_ctx.emq_periodic_mixed(x, x, w, L0, eps, same=True)
        idx = np.arange(0, n, n // 256)
        ex = oracle.direct_sum_ext(x[idx], x, w, L0, eps)
        r = np.max(np.abs(e[idx] - ex)) / np.max(np.abs(ex))
        print(f"2^20 eps {eps}: sampled rel {r:.2e}")
        assert r <= TOL


@pytest.mark.parametrize("amp", [1e-3, 1e-2])
def test_error_target(gpu_ctx, amp):
    """Error target of the mixed calls: weak perturbations (max|E| small
    against the weights) are where the FP32 far field's error is largest
    against the field. With the target (the band chosen from the call's own
    probed error) the result stays within 1e-8; target 0 restores the
    minimum-band result bit for bit."""
    import multi_species_amd as msa
    x, w = lattice_deck(0.3, 25.0, amp=amp)
    ex = oracle.direct_sum_ext(x, x, w, L0, 0.05)
    e0 = gpu_ctx.emq_periodic_mixed(x, x, w, L0, 0.05, same=True)
    try:
        gpu_ctx.set_mixed_target(4e-9)
        e1 = gpu_ctx.emq_periodic_mixed(x, x, w, L0, 0.05, same=True)
        e2 = gpu_ctx.emq_periodic_mixed(x.copy(), x, w, L0, 0.05)
    finally:
        gpu_ctx.set_mixed_target(0.0)
    e3 = gpu_ctx.emq_periodic_mixed(x, x, w, L0, 0.05, same=True)
    r0, r1, r2 = rel(e0, ex), rel(e1, ex), rel(e2, ex)
    print(f"amp {amp}: minimum band {r0:.2e}, target 4e-9: {r1:.2e} / {r2:.2e}")
    assert r1 <= TOL and r2 <= TOL
    assert np.array_equal(e3, e0)
    for bad in (-1.0, 1.0):
        with pytest.raises(msa.FarrsightError):
            gpu_ctx.set_mixed_target(bad)


@pytest.mark.parametrize("eps", [0.2, 0.5])
@pytest.mark.parametrize("kind", ["random", "lattice"])
def test_all_band_expansions_fp64(gpu_ctx, eps, kind):
    """When the minimum band reaches a quarter period every pair is FP64: the
    step part from prefix sums, compact near tiles and groups from their FP64
    Legendre expansions (truncation below 16^-12 of their weight), the rest
    pair by pair. The result then agrees with the extended-precision sum to
    FP64 accuracy, not just 1e-8."""
    if kind == "random":
        x, w = random_deck(30000, 11)
    else:
        x, w = lattice_deck(0.3, 25.0)
    e = gpu_ctx.emq_periodic_mixed(x, x, w, L0, eps, same=True)
    ex = oracle.direct_sum_ext(x, x, w, L0, eps)
    r = rel(e, ex)
    print(f"all-band {kind} eps {eps}: rel {r:.2e}")
    assert r <= 1e-11


def test_rejects_bad_eps(gpu_ctx):
    import multi_species_amd as msa
    x, w = random_deck(10, 1)
    with pytest.raises(msa.FarrsightError):
        gpu_ctx.emq_periodic_mixed(x, x, w, L0, 0.0, same=True)


def test_c2_twenty_seeds_full_field(gpu_ctx):
    """The c2 error study: N = 2^20 decks of 20 seeds at eps 0.01 and 0.05,
    the full mixed field (every target, minimum band and error target 4e-9)
    against the full FP64 field of the same deck (the pair-once FP64 path,
    pinned to the reference and the extended oracle by test_gpu_field)."""
    n = 1 << 20
    worst = {}
    try:
        for eps in (0.01, 0.05):
            for seed in range(20):
                x, w = random_deck(n, seed)
                f64 = gpu_ctx.emq_periodic_f64(x, x, w, L0, eps)
                gpu_ctx.set_mixed_target(0.0)
                e0 = gpu_ctx.emq_periodic_mixed(x, x, w, L0, eps, same=True)
                gpu_ctx.set_mixed_target(4e-9)
                e1 = gpu_ctx.emq_periodic_mixed(x, x, w, L0, eps, same=True)
                r0, r1 = rel(e0, f64), rel(e1, f64)
                worst[eps] = max(worst.get(eps, (0, 0))[0], r0), max(worst.get(eps, (0, 0))[1], r1)
                print(f"c2 seed {seed:2d} eps {eps}: minimum band {r0:.2e}, target 4e-9 {r1:.2e}",
                      flush=True)
                assert r0 <= TOL and r1 <= TOL, (seed, eps, r0, r1)
    finally:
        gpu_ctx.set_mixed_target(0.0)
    print("c2 worst over 20 seeds:", {k: f"{a:.2e}/{b:.2e}" for k, (a, b) in worst.items()})


@pytest.mark.parametrize("n", [1, 3000, 16770])
def test_small_calls_return_the_fp64_sum(gpu_ctx, n):
    """Below the size threshold the mixed entry point returns the FP64 sum bit
    for bit (targets = sources and separate targets); at threshold 0 the same
    call goes through the mixed kernels and stays within the mixed bar. A
    threshold test with the comparison reversed (< against <=) or a fallback
    to a different FP64 kernel would fail the bitwise check."""
    x, w = random_deck(n, 7)
    t = np.mod(x[::-1] + 0.3, L0)
    e64_same = gpu_ctx.emq_periodic_f64(x, x, w, L0, 0.05)
    e64_sep = gpu_ctx.emq_periodic_f64(t, x, w, L0, 0.05)
    gpu_ctx.set_mixed_min_pairs(n * n + 1)
    assert np.array_equal(gpu_ctx.emq_periodic_mixed(x, x, w, L0, 0.05, same=True), e64_same)
    assert np.array_equal(gpu_ctx.emq_periodic_mixed(t, x, w, L0, 0.05), e64_sep)
    gpu_ctx.set_mixed_min_pairs(n * n)  # nt * ns == min_pairs: mixed kernels
    em = gpu_ctx.emq_periodic_mixed(x, x, w, L0, 0.05, same=True)
    assert rel(em, e64_same) <= TOL
    gpu_ctx.set_mixed_min_pairs(0)
