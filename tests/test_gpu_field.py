"""GPU parity of the FP64 periodic direct sum (fs_emq_periodic_f64*).

Bar: normwise max|dE| <= 1e-12 max|E| against the reference's own outputs
(golden) and against the extended-precision oracle, plus pointwise
|dE_i| <= 1e-12 max|E| everywhere. At the full 2^20-point size the check is
sampled targets against the extended oracle and size-independent properties
(exact oddness in w, linearity, periodic translation, determinism).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

L0 = 4 * np.pi
TOL = 1e-12


def c2_deck(n, seed, eps_unused=None):
    rng = np.random.default_rng(seed)
    u = rng.uniform(0, L0, n)
    x = np.mod(u + 0.05 * np.sin(2 * np.pi / L0 * u), L0)
    v = rng.uniform(-6, 6, n)
    q = np.where(np.arange(n) < n // 2, -1.0, 1.0)
    w = q * np.exp(-v * v / 2) / np.sqrt(2 * np.pi) * 1e-4
    return x, w


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_matches_reference_golden(gpu_ctx, golden):
    for name, c in golden.items():
        if not name.startswith("ds_"):
            continue
        L, eps = c["params"]
        e = gpu_ctx.emq_periodic_f64(c["t"], c["s"], c["w"], L, eps)
        ex = oracle.direct_sum_ext(c["t"], c["s"], c["w"], L, eps)
        print(name, "gpu-vs-ref", rel(e, c["es"]), "gpu-vs-ext", rel(e, ex),
              "ref-vs-ext", rel(c["es"], ex))
        assert rel(e, c["es"]) <= TOL, name
        # at least as close to the exact sum as the reference's own FP64 sum
        assert rel(e, ex) <= max(1.5 * rel(c["es"], ex), 4e-15), name


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 255, 257, 1000, 4097, 16770])
def test_ragged_sizes(gpu_ctx, n):
    x, w = c2_deck(n, seed=n)
    e = gpu_ctx.emq_periodic_f64(x, x, w, L0, 0.05)
    ex = oracle.direct_sum_ext(x, x, w, L0, 0.05)
    if n == 1:
        assert e[0] == 0.0
    else:
        assert rel(e, ex) <= 1e-14


def test_empty_and_distinct_targets(gpu_ctx):
    x, w = c2_deck(500, 3)
    assert gpu_ctx.emq_periodic_f64(np.empty(0), x, w, L0, 0.05).size == 0
    e0 = gpu_ctx.emq_periodic_f64(x[:10], np.empty(0), np.empty(0), L0, 0.05)
    assert np.all(e0 == 0.0)
    t = np.linspace(-L0, 2 * L0, 333)  # unwrapped targets, not the sources
    e = gpu_ctx.emq_periodic_f64(t, x, w, L0, 0.05)
    assert rel(e, oracle.direct_sum_ext(t, x, w, L0, 0.05)) <= 1e-14


def test_coincident_points_give_exact_zero(gpu_ctx):
    s = np.array([2.0])
    e = gpu_ctx.emq_periodic_f64(np.array([2.0, 2.0 + L0 / 2, 2.0 - L0 / 2]), s,
                                 np.array([1.0]), L0, 0.05)
    assert np.all(e == 0.0)


@pytest.mark.parametrize("tpt,splits", [(1, 1), (1, 7), (2, 1), (2, 3), (4, 1), (4, 16)])
def test_launch_shapes_agree(gpu_ctx, tpt, splits):
    x, w = c2_deck(20000, 4)
    ex = oracle.direct_sum_ext(x, x, w, L0, 0.05)
    gpu_ctx.set_tuning(tpt, splits)
    try:
        e1 = gpu_ctx.emq_periodic_f64(x, x, w, L0, 0.05)
        e2 = gpu_ctx.emq_periodic_f64(x, x, w, L0, 0.05)
    finally:
        gpu_ctx.set_tuning(0, 0)
    assert np.array_equal(e1, e2)  # deterministic
    assert rel(e1, ex) <= 1e-14


@pytest.mark.parametrize("n,tpt", [(1, 1), (2, 1), (63, 1), (64, 2), (65, 4), (65, 8),
                                   (511, 1), (512, 1), (513, 1), (1025, 1), (1536, 2),
                                   (2049, 1), (4097, 4), (4097, 8), (8193, 8), (16770, 0),
                                   (40000, 0)])
def test_pair_once_ragged(gpu_ctx, n, tpt):
    """Targets = sources, each pair evaluated once and given to both points:
    ragged sizes around the 64-point tile and the row block (512, 1024, 2048
    points), odd and even block counts, against the extended oracle."""
    x, w = c2_deck(n, seed=n + 7)
    gpu_ctx.set_symmetric(2)
    gpu_ctx.set_tuning(tpt, 0)
    try:
        e = gpu_ctx.emq_periodic_f64(x, x, w, L0, 0.05)
        e2 = gpu_ctx.emq_periodic_f64(x, x, w, L0, 0.05)
        em = gpu_ctx.emq_periodic_f64(x, x, -w, L0, 0.05)
    finally:
        gpu_ctx.set_symmetric(0)
        gpu_ctx.set_tuning(0, 0)
    assert np.array_equal(e, e2)
    assert np.array_equal(em, -e)
    ex = oracle.direct_sum_ext(x, x, w, L0, 0.05)
    if n == 1:
        assert e[0] == 0.0
    else:
        assert rel(e, ex) <= 1e-14, rel(e, ex)


@pytest.mark.parametrize("tpt", [2, 8])
def test_pair_once_positions_beyond_the_period(gpu_ctx, tpt):
    """The reference does not wrap positions (the RK4 step moves them out of
    [0, L)): half the points spread over [-1.3L, 2.2L), a cluster straddling
    the periodic seam at L and one at L/2 (tiles whose displacements reach
    +-L/2 and +-3L/2). The pair-once sum sorts the points and gives each tile
    of sources one wrap where its displacement range allows, per pair
    otherwise; against the extended oracle. A mistake this catches: a tile
    given one wrap while its pairs need two (its far pairs then enter with
    the wrong image, an O(1) error)."""
    rng = np.random.default_rng(3)
    n = 40000
    x = np.concatenate([rng.uniform(-1.3 * L0, 2.2 * L0, n // 2),
                        L0 + rng.normal(0.0, 0.01, n // 4),
                        0.5 * L0 + rng.normal(0.0, 0.01, n - n // 2 - n // 4)])
    rng.shuffle(x)
    w = rng.normal(size=n) * 1e-3
    gpu_ctx.set_symmetric(2)
    gpu_ctx.set_tuning(tpt, 0)
    try:
        e = gpu_ctx.emq_periodic_f64(x, x, w, L0, 0.05)
    finally:
        gpu_ctx.set_symmetric(0)
        gpu_ctx.set_tuning(0, 0)
    ex = oracle.direct_sum_ext(x, x, w, L0, 0.05)
    assert rel(e, ex) <= 1e-13, rel(e, ex)


def test_pair_once_with_copied_positions(gpu_ctx):
    """The reference passes two copies of the same positions as targets and
    sources: identical contents take the pair-once form like one array."""
    n = 40000
    x, w = c2_deck(n, seed=5)
    e1 = gpu_ctx.emq_periodic_f64(x, x, w, L0, 0.05)
    ev1 = gpu_ctx.f64_pair_evals()
    e2 = gpu_ctx.emq_periodic_f64(x.copy(), x.copy(), w, L0, 0.05)
    ev2 = gpu_ctx.f64_pair_evals()
    assert np.array_equal(e1, e2)
    assert ev1 == ev2 and ev2 < n * n * 0.6
    y = x.copy()
    y[123] += 1e-9  # not the same positions: the full sum
    gpu_ctx.emq_periodic_f64(x, y, w, L0, 0.05)
    assert gpu_ctx.f64_pair_evals() >= n * n


@pytest.mark.parametrize("splits", [0, 1, 3, 1000])
def test_pair_once_matches_full_sum(gpu_ctx, splits):
    """The pair-once form against the full sum of every ordered pair (same
    device arrays), for several chunkings of the source blocks."""
    n = 70001
    x, w = c2_deck(n, seed=11)
    gpu_ctx.set_symmetric(1)
    try:
        full = gpu_ctx.emq_periodic_f64(x, x, w, L0, 0.05)
    finally:
        gpu_ctx.set_symmetric(0)
    gpu_ctx.set_tuning(0, splits)
    try:
        e = gpu_ctx.emq_periodic_f64(x, x, w, L0, 0.05)
    finally:
        gpu_ctx.set_tuning(0, 0)
    assert rel(e, full) <= 1e-14, rel(e, full)
    pick = np.arange(0, n, 997)
    ex = oracle.direct_sum_ext(x[pick], x, w, L0, 0.05)
    assert np.max(np.abs(e[pick] - ex)) <= 1e-14 * np.max(np.abs(ex))


@pytest.mark.parametrize("eps", [0.01, 0.05])
def test_full_size_c2(gpu_ctx, eps):
    """N = 2^20 targets = sources (the pair-once form): sampled targets
    against the extended oracle over all sources, and size-independent
    properties of the whole field."""
    n = 1 << 20
    x, w = c2_deck(n, seed=0)
    e = gpu_ctx.emq_periodic_f64(x, x, w, L0, eps)
    rng = np.random.default_rng(1)
    pick = np.sort(rng.choice(n, 96, replace=False))
    ex = oracle.direct_sum_ext(x[pick], x, w, L0, eps)
    scale = np.max(np.abs(e))
    assert np.max(np.abs(e[pick] - ex)) <= TOL * scale
    em = gpu_ctx.emq_periodic_f64(x, x, -w, L0, eps)
    assert np.array_equal(em, -e)
    e2 = gpu_ctx.emq_periodic_f64(x, x, 2.0 * w, L0, eps)
    assert np.array_equal(e2, 2.0 * e)  # scaling by 2 is exact
    eL = gpu_ctx.emq_periodic_f64(x + L0, x, w, L0, eps)
    assert np.max(np.abs(eL - e)) <= TOL * scale


def test_rsq_refinement_is_full_precision(gpu_ctx):
    """Single-source fields sweep the kernel over z in [-1/2, 1/2]: every value
    within a few ulp of the extended oracle (pins the refined v_rsq_f64)."""
    z = np.linspace(-0.5, 0.5, 20001)
    t = z * L0
    for eps in (1e-3, 0.05, 1.0):
        e = gpu_ctx.emq_periodic_f64(t, np.array([0.0]), np.array([1.0]), L0, eps)
        ex = oracle.direct_sum_ext(t, np.array([0.0]), np.array([1.0]), L0, eps)
        scale = np.max(np.abs(ex))
        assert np.max(np.abs(e - ex)) <= 8e-16 * scale, eps


def test_pair_once_budget_changes_no_bit(gpu_ctx):
    """The pair-once partials in passes under a small budget (3 and 16 passes
    at N = 2^20) against one pass: bitwise the same field, periodic and
    open-boundary alike."""
    import multi_species_amd as ms

    n = 1 << 20
    x, w = c2_deck(n, seed=5)
    assert ms.pair_once_plan(n)[1] == 1
    assert ms.pair_once_plan(n, 1 << 30)[1] == 3
    assert ms.pair_once_plan(n, 1)[1] == 16
    ref = gpu_ctx.emq_periodic_f64(x, x, w, L0, 0.05)
    ref_open = gpu_ctx.emq_open_f64(x, x, w, 0.05)
    try:
        for budget in (1 << 30, 1):
            gpu_ctx.set_pair_once_budget(budget)
            e = gpu_ctx.emq_periodic_f64(x, x, w, L0, 0.05)
            assert np.array_equal(e, ref), budget
            eo = gpu_ctx.emq_open_f64(x, x, w, 0.05)
            assert np.array_equal(eo, ref_open), budget
    finally:
        gpu_ctx.set_pair_once_budget(0)


def test_pair_once_at_2p22_within_budget(gpu_ctx):
    """N = 2^22 targets = sources: one pass would need 32 GiB of partials; the
    default budget runs it in 10 passes within 4 GiB. Sampled targets against
    the full-sum kernel over all sources (distinct target array), the uniform
    lattice's zero field, and the exact oddness in w."""
    import multi_species_amd as ms

    n = 1 << 22
    b, passes, evals = ms.pair_once_plan(n)
    assert passes == 10 and b <= 4 << 30
    assert evals == n * n // 2 + n * 2048 // 2
    x, w = c2_deck(n, seed=11)
    e = gpu_ctx.emq_periodic_f64(x, x, w, L0, 0.05)
    assert gpu_ctx.f64_pair_evals() == evals
    pick = np.sort(np.random.default_rng(2).choice(n, 2048, replace=False))
    ref = gpu_ctx.emq_periodic_f64(x[pick].copy(), x, w, L0, 0.05)
    assert np.max(np.abs(e[pick] - ref)) <= TOL * np.max(np.abs(e))
    assert np.array_equal(gpu_ctx.emq_periodic_f64(x, x, -w, L0, 0.05), -e)
    lat = np.arange(n) * (L0 / n)
    e0 = gpu_ctx.emq_periodic_f64(lat, lat, np.full(n, 1.0 / n), L0, 0.05)
    assert np.max(np.abs(e0)) <= 1e-13
