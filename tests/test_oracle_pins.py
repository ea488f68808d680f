"""Pins, from the mathematics, the three oracle functions that no reference
binary can run for (Eigen and Boost are absent here): adaptive refinement
(Species::test_panel and the hanging-node branches of refine_panels,
AMRStructure_amr.cpp:506-839, 971-1081), the root descent of
interpolate_from_mesh (AMRStructure_interp.cpp:232-426, 977-1143), and the
diagnostics (standard_interface/diagnostic_plot.py:116-143).

Each check is computed here, independently of the oracle's code, from what
the reference's definitions fix:

* Refinement, on small adaptive decks with every combination of the five
  amr_epsilons criteria: (1) no panel above max_height; (2) every leaf's nine
  points lie on its 3x3 sub-grid (children at x0 + dx/2 i, p0 + dp/2 j); (3)
  f at every point is f0 there (the initial mesh evaluates f0 at new points);
  (4) every leaf below max_height fails all criteria -- refinement stopped
  only when nothing was flagged; (5) every refined panel beyond the base
  either meets a criterion or borders a refined panel one level finer (the
  2:1 balance flag); (6) edge-adjacent leaves differ by at most one level;
  (7) no two point ids share coordinates (hanging-node points are shared).
  Mistakes these catch: a criterion on the wrong epsilon or with the wrong
  comparison, a refinement pass that stops early or refines unflagged
  panels, a missing balance flag, children at wrong coordinates, interpolated
  instead of exact f at new points, an off-by-one max_height.
* Root descent, on refined meshes deformed by shear and a sinusoid: the leaf
  it picks contains the point (brute-force point-in-quadrilateral over every
  leaf and periodic image, with the reference's half-plane tie rules); with
  random f the interpolated value equals the 9-monomial interpolant of that
  leaf solved here with numpy; f(x, p) = a + b p + c p^2 is reproduced at
  every point inside. Mistakes caught: a wrong child choice or neighbour
  redirect, a missed periodic shift, a transposed monomial or wrong centre.
* Diagnostics: closed forms -- the charge and kinetic energy of a Maxwellian
  deck (q L and m vth^2 L / 2 up to the domain's erfc(6/sqrt 2) tail), zero
  momentum of a p-symmetric f, the electric energy of E = A sin(2 pi m x/L)
  on a uniform grid (A^2 L / 4 exactly, duplicates and points outside [0, L)
  ignored), entropy, L1/L2, extrema and negative parts of synthetic f.
  Mistakes caught: a missing 1/2 or 1/q, p instead of p^2, duplicates
  counted twice, the periodic spacing at the ends wrong.
"""
import itertools

import numpy as np
import pytest

import oracle

L0 = 4 * np.pi


# ---------------------------------------------------------------- refinement

def f0(sp, kx, x, p):
    def M(v, th, c):
        return np.exp(-(v - c) ** 2 / (2 * th * th)) / (np.sqrt(2 * np.pi) * th)
    pert = 1 + sp.amp * np.cos(kx * x)
    if sp.ics_type == 4:
        return 0.5 * (M(p, sp.pth, sp.pstr) + M(p, sp.pth, -sp.pstr)) * pert
    if sp.ics_type == 3:
        return M(p, sp.pth, 0.0) * pert * p * p
    return M(p, sp.pth, sp.pstr) * pert


def criteria(f9, x9, p9, eps):
    """test_panel's refinement criteria on a panel's nine values."""
    met = False
    if len(eps) > 0:
        mx, mn = f9.max(), f9.min()
        met |= (mx - mn) > eps[0]
        if len(eps) > 1:
            met |= (mx - mn) / f9[4] > eps[1]
        if len(eps) > 2:
            met |= mx / mn > eps[2]
    if len(eps) > 3:
        dx = x9[3] - x9[0]
        met |= max(abs((f9[3 + i] - f9[i]) / dx) for i in range(6)) > eps[3]
    if len(eps) > 4:
        # the reference fills entries 0 and 1 of its six p-differences in each
        # of three passes (its last pass, the column at x1, survives) and
        # leaves entries 2..5 unset, taken as 0
        dp = p9[1] - p9[0]
        met |= max(abs((f9[7] - f9[6]) / dp), abs((f9[8] - f9[7]) / dp), 0.0) > eps[4]
    return met


def criteria_violations(rec, xs, ps, fs, eps, sp, adjacent_to):
    """(leaves below max_height that meet a criterion, refined panels beyond
    the base that neither meet one nor border a refined panel one level
    finer, refined panels justified by balance alone)."""
    lvl = rec[:, 10]
    refined = (rec[:, 17] & 8) != 0
    leaf = (rec[:, 17] & 24) == 0
    pts = rec[:, 1:10]
    base = sp.initial_height + sp.p_height
    flagged = np.array([criteria(fs[pts[k]], xs[pts[k]], ps[pts[k]], eps)
                        for k in range(len(rec))])
    bad = int(np.sum(leaf & (lvl < sp.max_height) & flagged))
    unjustified = n_balance = 0
    for k in np.nonzero(refined & (lvl >= base))[0]:
        if flagged[k]:
            continue
        finer = np.nonzero(refined & (lvl == lvl[k] + 1))[0]
        if adjacent_to(k, finer).any():
            n_balance += 1
        else:
            unjustified += 1
    return bad, unjustified, n_balance


def adaptive_species(eps, ics=4, h=3, ph=1, extra=3, amp=5e-2, pth=0.5, pstr=2.0, L=10 * np.pi):
    sp = oracle.Species(name="e", q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, initial_height=h,
                        p_height=ph, max_height=h + ph + extra, do_adaptively_refine=True,
                        amr_epsilons=list(eps), ics_type=ics, normalized_wavenumber=1.0,
                        amp=amp, pth=pth, pstr=pstr)
    deck = oracle.Deck(x_min=0.0, x_max=L, quadrature=0, greens_epsilon=0.05, dt=0.05,
                       remesh_period=1, species=[sp])
    return deck, sp, oracle.OracleSpecies(deck, sp)


EPS_SETS = [
    [2e-3],
    [1e-2, 0.05],
    [1e9, 1e9, 1.2],
    [1e9, 1e9, 1e9, 0.003],
    [1e9, 1e9, 1e9, 1e9, 0.05],
    [5e-3, 0.1, 1e9, 0.02, 0.1],
]


@pytest.mark.parametrize("eps", EPS_SETS, ids=[str(i) for i in range(len(EPS_SETS))])
@pytest.mark.parametrize("ics", [1, 4])
def test_adaptive_refinement_invariants(eps, ics):
    deck, sp, s = adaptive_species(eps, ics=ics, pstr=2.0 if ics == 4 else 0.0)
    rec = s.panels()
    xs, ps, fs, _ = s.points()
    kx = 2 * np.pi / deck.x_max * sp.normalized_wavenumber
    L = deck.x_max - deck.x_min
    base = sp.initial_height + sp.p_height
    lvl = rec[:, 10]
    refined = (rec[:, 17] & 8) != 0
    leaf = (rec[:, 17] & 24) == 0
    pts = rec[:, 1:10]
    # (1)
    assert lvl.max() <= sp.max_height
    # (3) f at every point is f0
    ex = f0(sp, kx, xs, ps)
    assert np.max(np.abs(fs - ex)) <= 1e-14 * np.max(np.abs(ex))
    # (2) sub-grid of every leaf (a point shared with a neighbour carries the
    # coordinate the neighbour computed for it: equal to rounding)
    X, P = xs[pts], ps[pts]
    tx, tp = 1e-13 * L, 1e-13 * 12.0
    for col in range(3):
        assert np.all(np.abs(X[:, 3 * col] - X[:, 3 * col + 1]) <= tx)
        assert np.all(np.abs(X[:, 3 * col] - X[:, 3 * col + 2]) <= tx)
        assert np.all(np.abs(P[:, col] - P[:, col + 3]) <= tp)
        assert np.all(np.abs(P[:, col] - P[:, col + 6]) <= tp)
    dx0, dp0 = L / 2 ** sp.initial_height, 12.0 / 2 ** (sp.initial_height + sp.p_height)
    for k in np.nonzero(lvl >= base)[0]:
        d = 2.0 ** -(lvl[k] - base)
        assert np.allclose(np.diff(X[k, ::3]), 0.5 * dx0 * d, rtol=1e-12, atol=0)
        assert np.allclose(np.diff(P[k, :3]), 0.5 * dp0 * d, rtol=1e-12, atol=0)
    # boxes of every panel (undeformed mesh)
    x0, x1, p0, p1 = X[:, 0], X[:, 6], P[:, 0], P[:, 2]

    def adjacent_to(a, cand):
        """Which of the panels cand share an edge segment with panel a
        (periodic in x)."""
        px = np.minimum(p1[a], p1[cand]) - np.maximum(p0[a], p0[cand])
        xo = np.minimum(x1[a], x1[cand]) - np.maximum(x0[a], x0[cand])

        def near(u, v):
            d = np.abs(u - v)
            return np.minimum(d, np.abs(d - L)) <= tx

        vert = (px > tp) & (near(x1[a], x0[cand]) | near(x1[cand], x0[a]))
        hor = (xo > tx) & ((np.abs(p1[a] - p0[cand]) <= tp) | (np.abs(p1[cand] - p0[a]) <= tp))
        return vert | hor

    bad, unjustified, n_balance = criteria_violations(rec, xs, ps, fs, eps, sp, adjacent_to)
    # (4) leaves below max_height fail every criterion
    assert bad == 0
    # (5) refined panels beyond the base: a criterion, or a finer refined neighbour
    assert unjustified == 0
    # (6) 2:1 balance between edge-adjacent leaves
    lv = np.nonzero(leaf)[0]
    for a in lv:
        far = lv[lvl[lv] >= lvl[a] + 2]
        if far.size:
            assert not adjacent_to(a, far).any(), a
    # (7) distinct coordinates for distinct ids (beyond rounding)
    key = set(zip(np.round(xs / (1e-9 * L)).astype(np.int64).tolist(),
                  np.round(ps / 1e-9).astype(np.int64).tolist()))
    assert len(key) == xs.size
    assert refined[lvl >= base].sum() > 0, "the deck must refine"
    print(f"eps {eps} ics {ics}: {len(rec)} panels, {leaf.sum()} leaves, "
          f"{refined[lvl >= base].sum()} refined beyond the base ({n_balance} for balance)")


def test_invariants_reject_a_misread_criterion():
    """The mesh refined on the p-gradient criterion, checked as if the
    threshold were the x-gradient's (a plausible misreading of the
    amr_epsilons order): the checks report violations."""
    eps = [1e9, 1e9, 1e9, 1e9, 0.05]
    deck, sp, s = adaptive_species(eps, ics=1, pstr=0.0)
    rec = s.panels()
    xs, ps, fs, _ = s.points()
    everywhere = lambda a, cand: np.ones(len(cand), bool)  # noqa: E731
    assert criteria_violations(rec, xs, ps, fs, eps, sp, everywhere)[:2] == (0, 0)
    bad, unjustified, _ = criteria_violations(rec, xs, ps, fs, [1e9, 1e9, 1e9, 0.05], sp,
                                              lambda a, cand: np.zeros(len(cand), bool))
    assert bad + unjustified > 0


def test_refinement_depth_stops_at_max_height():
    # a criterion every panel meets: refined to max_height everywhere, no further
    deck, sp, s = adaptive_species([0.0], ics=1, h=2, ph=1, extra=2, amp=0.3, pth=1.0, pstr=0.0)
    rec = s.panels()
    leaf = (rec[:, 17] & 24) == 0
    assert np.all(rec[leaf, 10] == sp.max_height)
    n_base = 2 ** sp.initial_height * 2 ** (sp.initial_height + sp.p_height)
    assert leaf.sum() == n_base * 4 ** 2
    xs = s.points()[0]
    nx = 2 ** (sp.initial_height + 2 + 1) + 1
    assert xs.size == nx * (2 ** (sp.initial_height + sp.p_height + 2 + 1) + 1)


# ------------------------------------------------------------------ descent

def inside(X, P, x, p):
    """The reference's containment tests against corners 0, 2, 6, 8."""
    r = (X[8] - X[6]) * (p - P[6]) > (P[8] - P[6]) * (x - X[6])
    l_ = (X[2] - X[0]) * (p - P[0]) <= (P[2] - P[0]) * (x - X[0])
    t = (X[8] - X[2]) * (p - P[2]) < (P[8] - P[2]) * (x - X[2])
    b = (X[6] - X[0]) * (p - P[0]) >= (P[6] - P[0]) * (x - X[0])
    return r and l_ and t and b


def monomials(dx, dp):
    return np.array([np.ones_like(dx), dx, dx * dp, dp, dx * dx, dx * dx * dp,
                     dx * dx * dp * dp, dx * dp * dp, dp * dp])


def deformed(h=3, ph=1, eps=(2e-3,), seed=0):
    deck, sp, s = adaptive_species(eps, ics=4, h=h, ph=ph, extra=2)
    xs, ps, fs, _ = s.points()
    L = deck.x_max
    ax = xs + 0.15 * ps + 0.4 * np.sin(2 * np.pi * xs / L)
    ap = ps + 0.1 * np.cos(2 * np.pi * xs / L)
    s.set_points(ax, ap)
    rng = np.random.default_rng(seed)
    fr = rng.normal(size=xs.size)
    s.set_fs(fr)
    s.copy_to_old()
    return deck, sp, s, ax, ap, fr, L


def test_descent_finds_the_containing_leaf_and_its_interpolant():
    deck, sp, s, ax, ap, fr, L = deformed()
    rec = s.panels()
    leaves = np.nonzero((rec[:, 17] & 24) == 0)[0]
    assert (rec[leaves, 10] > sp.initial_height + sp.p_height).any()  # refined leaves present
    rng = np.random.default_rng(3)
    m = 400
    x = rng.uniform(-0.5 * L, 1.5 * L, m)
    p = rng.uniform(-5.0, 5.0, m)
    got_leaf = s.descend_points(x, p)
    got_f = s.interp_from_mesh(x, p)
    checked = 0
    for i in range(m):
        hits = []
        for k in leaves:
            pt = rec[k, 1:10]
            for sh in (-2, -1, 0, 1, 2):
                if inside(ax[pt], ap[pt], x[i] + sh * L, p[i]):
                    hits.append((k, x[i] + sh * L))
        if len(hits) != 1:
            continue  # on a seam or outside: the tie rules alone decide
        k, xsh = hits[0]
        assert got_leaf[i] == k, (i, got_leaf[i], k)
        pt = rec[k, 1:10]
        A = monomials(ax[pt] - ax[pt[4]], ap[pt] - ap[pt[4]]).T
        c = np.linalg.solve(A, fr[pt])
        ref = monomials(np.array(xsh - ax[pt[4]]), np.array(p[i] - ap[pt[4]])) @ c
        assert abs(got_f[i] - ref) <= 1e-9 * max(1.0, np.max(np.abs(c))), (i, got_f[i], ref)
        checked += 1
    assert checked > 0.8 * m


def test_descent_reproduces_quadratics_in_p():
    deck, sp, s, ax, ap, _, L = deformed(seed=1)
    a, b, c = 0.7, -0.3, 0.05
    s.set_fs(a + b * ap + c * ap * ap)
    s.copy_to_old()
    rng = np.random.default_rng(5)
    x = rng.uniform(0, L, 300)
    p = rng.uniform(-5.0, 5.0, 300)
    leaf = s.descend_points(x, p)
    f = s.interp_from_mesh(x, p)
    ok = leaf > 0
    assert ok.mean() > 0.95
    ex = a + b * p + c * p * p
    assert np.max(np.abs(f[ok] - ex[ok])) <= 1e-10


# --------------------------------------------------------------- diagnostics

def maxwellian_species(simpson=False, q=-1.0, m=1.0, pth=1.0):
    sp = oracle.Species(name="e", q=q, m=m, p_min=-6.0, p_max=6.0, initial_height=5,
                        p_height=1, ics_type=1, normalized_wavenumber=1.0, amp=0.0, pth=pth)
    deck = oracle.Deck(x_min=0.0, x_max=L0, quadrature=1 if simpson else 0,
                       greens_epsilon=0.05, dt=0.1, remesh_period=1, species=[sp])
    return deck, sp, oracle.OracleSpecies(deck, sp)


@pytest.mark.parametrize("simpson", [False, True])
def test_diagnostics_of_a_maxwellian(simpson):
    from math import erf, exp, pi, sqrt
    deck, sp, s = maxwellian_species(simpson=simpson)
    xs, ps, fs, qw = s.points()
    es = np.zeros_like(xs)
    d = oracle.diagnostics(xs, ps, fs, qw, es, sp.q, sp.q / sp.m, L0)
    mass = erf(6 / sqrt(2))                       # int f dp over [-6, 6]
    second = mass - 12 * exp(-18) / sqrt(2 * pi)  # int p^2 f dp over [-6, 6]
    assert abs(d[0] - sp.q * L0 * mass) <= 1e-9 * L0
    assert abs(d[1]) <= 1e-13 * np.sum(np.abs(ps * qw))
    assert abs(d[2] - sp.m / 2 * L0 * second) <= 1e-8 * L0
    assert d[3] == 0.0
    assert d[5] == pytest.approx(L0 * mass, rel=1e-9)
    assert d[7] == fs.min() and d[8] == fs.max() and d[7] > 0
    assert d[9] == 0.0 and d[10] == 0 and d[11] == 0 and d[12] == 0
    assert d[13] == xs.size


@pytest.mark.parametrize("m", [1, 3])
def test_electric_energy_of_a_sine(m):
    n, A = 64, 0.37
    xg = np.arange(n) * (L0 / n)
    rows = 5  # each x repeated on five p rows, in shuffled order, plus points at L and < 0
    xs = np.concatenate([np.tile(xg, rows), [L0, -0.5]])
    es = A * np.sin(2 * np.pi * m * xs / L0)
    es[-2:] = 100.0  # outside [0, L): must not count
    perm = np.random.default_rng(m).permutation(xs.size)
    xs, es = xs[perm], es[perm]
    one = np.ones_like(xs)
    d = oracle.diagnostics(xs, one, one, one, es, -1.0, -1.0, L0)
    assert abs(d[3] - A * A * L0 / 4) <= 1e-14 * A * A * L0


def test_electric_energy_first_occurrence_and_spacing():
    # unequal spacing: 1/2 sum dx_k E_k^2 with the periodic midpoint spacing;
    # a duplicate x takes the E of its first occurrence
    xs = np.array([1.0, 0.0, 3.0, 1.0, 2.5])
    es = np.array([2.0, 1.0, 3.0, 50.0, 4.0])
    L = 4.0
    one = np.ones_like(xs)
    d = oracle.diagnostics(xs, one, one, one, es, 1.0, 1.0, L)
    # unique x 0, 1, 2.5, 3 with E 1, 2, 4, 3
    dx = [0.5 * (1.0 + L - 3.0), 0.5 * (2.5 - 0.0), 0.5 * (3.0 - 1.0), 0.5 * (0.0 + L - 2.5)]
    assert d[3] == pytest.approx(0.5 * (dx[0] * 1 + dx[1] * 4 + dx[2] * 16 + dx[3] * 9),
                                 rel=1e-15)


def test_entropy_norms_and_negative_parts():
    rng = np.random.default_rng(9)
    n = 1000
    fs = rng.uniform(-0.2, 2.0, n)
    qw = rng.uniform(0.5, 1.5, n) * -2.0  # q = -2
    ps = rng.uniform(-3, 3, n)
    xs = rng.uniform(0, 1, n)
    q, qm = -2.0, -2.0 / 3.0
    d = oracle.diagnostics(xs, ps, fs, qw, np.zeros(n), q, qm, 1.0)
    pos, neg = fs > 0, fs < 0
    assert d[1] == pytest.approx(np.sum(ps * qw) / q, rel=1e-12)
    assert d[2] == pytest.approx(np.sum(ps * ps * qw) / (2 * qm), rel=1e-12)
    assert d[4] == pytest.approx(np.sum(fs[pos] * np.log10(fs[pos]) * qw[pos]) / q, rel=1e-12)
    assert d[5] == pytest.approx(np.sum(np.abs(qw)) / 2.0, rel=1e-13)
    assert d[6] == pytest.approx(np.sum(qw * fs) / q, rel=1e-12)
    assert d[7] == fs.min() and d[8] == fs.max()
    assert d[9] == pytest.approx(np.sum(qw[neg]) / q, rel=1e-12)
    assert d[10] == neg.sum() and d[11] == pytest.approx(neg.mean(), rel=1e-15)
    assert d[12] == pytest.approx(abs(np.sum(qw[neg])) / np.sum(np.abs(qw)), rel=1e-12)
    # entropy of a constant f: c log10(c) times the charge over q
    c = 0.3
    d = oracle.diagnostics(xs, ps, np.full(n, c), qw, np.zeros(n), q, qm, 1.0)
    assert d[4] == pytest.approx(c * np.log10(c) * np.sum(qw) / q, rel=1e-12)


def test_simulation_diagnostics_use_the_same_definitions():
    d = oracle.landau_deck(initial_height=4, p_height=1)
    sim = oracle.OracleSim(d)
    sim.step()
    for k in range(2):
        sp = d.species[k]
        xs, ps, fs, qw = sim.species(k).points()
        full = oracle.diagnostics(xs, ps, fs, qw, sim.es(k), sp.q, sp.q / sp.m, d.x_max)
        assert np.array_equal(sim.diagnostics_full(k), full)
        assert np.array_equal(sim.diagnostics(k), full[:4])
