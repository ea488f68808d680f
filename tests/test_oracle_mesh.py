"""Pins the oracle's remesh interpolation (part B) where no reference binary
can run (Eigen and Boost are absent): properties the mathematics fixes.

* the 9-point interpolant reproduces every function in its monomial span
  (1, dx, dp, dx dp, dx^2, dx^2 dp, dx^2 dp^2, dx dp^2, dp^2 about any centre)
  on any deformed panel -- so f(x, p) = a + b p + c p^2 (x-independent, hence
  periodic) is reproduced exactly after any shear, and a biquadratic in x is
  reproduced away from the periodic seam;
* a zero-displacement remesh returns f at the nodes;
* the leaf the neighbour walk picks contains the point (brute-force
  point-in-quadrilateral search over all leaves);
* points pushed beyond the p boundary get f_beyond_boundary; more than ten
  periods away raises "too many shifts!" (AMRStructure_interp.cpp:469-471).
"""
import numpy as np
import pytest

import oracle

L0 = 4 * np.pi


def fresh(h=4, ph=1, amp=0.05, pth=1.0):
    d = oracle.landau_deck(initial_height=h, p_height=ph)
    sp = d.species[0]
    sp.amp, sp.pth = amp, pth
    return d, sp, oracle.OracleSpecies(d, sp)


def grid(h, ph, d, sp):
    g = oracle.OracleSpecies(d, sp)
    xs, ps, _, _ = g.points()
    return xs, ps, 2 ** (h + 1) + 1, 2 ** (h + ph + 1) + 1


def test_zero_displacement_reproduces_nodes():
    d, sp, s = fresh()
    xs, ps, fs, _ = s.points()
    s.copy_to_old()
    gx, gp, nx, np_ = grid(4, 1, d, sp)
    f, leaf, _ = s.interp_to_grid(gx, gp, nx, np_)
    top = gp == sp.p_max
    # every other node is reproduced up to the rounding of the 9x9 solve
    assert np.max(np.abs(f[~top] - fs[: nx * np_][~top])) < 1e-12 * np.max(fs)
    # the reference's top-edge test is strict (ineq_top uses <,
    # AMRStructure_interp.cpp:62), so the p = p_max row is "beyond the boundary"
    # and takes f_beyond_boundary (:212-224, :799-803) even without motion
    assert np.all(f[top] == s.f_beyond_boundary)
    assert np.sum(leaf == 0) == nx


@pytest.mark.parametrize("t", [0.05, 0.3, 1.5])
def test_shear_preserves_x_independent_polynomial(t):
    # f0 with amp = 0 and a huge pth is ~ constant; instead build f = a + b p + c p^2 by
    # choosing a deck whose f0 is x-independent and quadratic in p on each panel: the
    # strong two-stream IC times a huge thermal width is F0_LD * p^2 ~ const * p^2.
    d = oracle.landau_deck(initial_height=4, p_height=1)
    sp = d.species[0]
    sp.ics_type, sp.amp, sp.pth = 3, 0.0, 1e8
    s = oracle.OracleSpecies(d, sp)
    xs, ps, fs, _ = s.points()
    c = fs[0] / ps[0] ** 2
    assert np.allclose(fs, c * ps ** 2, rtol=1e-11, atol=0)
    s.set_points(xs + t * ps, ps)  # pure shear: f(x, p) unchanged as a function
    s.copy_to_old()
    gx, gp, nx, np_ = grid(4, 1, d, sp)
    f, leaf, _ = s.interp_to_grid(gx, gp, nx, np_)
    inside = leaf > 0
    assert inside.mean() > 0.9
    _, _, perm = s.interp_to_grid(gx, gp, nx, np_)
    f_exact = c * gp ** 2
    # points that landed beyond the sheared p boundary take f_beyond_boundary
    ok = np.zeros(gx.size, bool)
    ok[perm[inside]] = True
    assert np.max(np.abs(f[ok] - f_exact[ok])) < 1e-9 * np.max(np.abs(f_exact))


def test_walk_leaf_contains_point():
    d, sp, s = fresh(h=3, ph=1)
    xs, ps, fs, _ = s.points()
    ax = xs + 0.2 * ps + 0.3 * np.sin(0.5 * xs)
    ap = ps + 0.2 * np.cos(0.5 * xs)
    s.set_points(ax, ap)
    s.copy_to_old()
    gx, gp, nx, np_ = grid(3, 1, d, sp)
    f, leaf, perm = s.interp_to_grid(gx, gp, nx, np_)
    rec = s.panels()
    leaves = np.nonzero((rec[:, 17] & 24) == 0)[0]

    def inside(k, x, p):
        pt = rec[k, 1:10]
        X, P = ax[pt], ap[pt]
        r = (X[8] - X[6]) * (p - P[6]) > (P[8] - P[6]) * (x - X[6])
        l = (X[2] - X[0]) * (p - P[0]) <= (P[2] - P[0]) * (x - X[0])
        t = (X[8] - X[2]) * (p - P[2]) < (P[8] - P[2]) * (x - X[2])
        b = (X[6] - X[0]) * (p - P[0]) >= (P[6] - P[0]) * (x - X[0])
        return r and l and t and b

    checked = 0
    for k_sorted in range(0, gx.size, 7):
        lf = leaf[k_sorted]
        if lf == 0:
            continue
        i = perm[k_sorted]
        x, p = gx[i], gp[i]
        hits = [k for k in leaves if any(inside(k, x + m * L0, p) for m in (-2, -1, 0, 1, 2))]
        if len(hits) == 1:  # strictly inside exactly one leaf (not on a seam)
            assert lf == hits[0]
            checked += 1
    assert checked > 20


def test_beyond_boundary_and_too_many_shifts():
    d, sp, s = fresh(h=3, ph=1)
    s.copy_to_old()
    fb = s.f_beyond_boundary
    f = s.interp_from_mesh(np.array([1.0, 3.0]), np.array([6.5, -7.0]))
    assert np.all(f == fb)
    with pytest.raises(RuntimeError, match="too many shifts"):
        s.interp_from_mesh(np.array([1.0 + 30 * L0]), np.array([0.0]))


def test_remesh_keeps_structure_and_charge_close():
    d, sp, s = fresh(h=4, ph=1)
    xs, ps, fs, qw = s.points()
    rec0 = s.panels()
    s.set_points(xs + 0.1 * ps, ps)
    s.remesh()
    xs2, ps2, fs2, qw2 = s.points()
    assert np.array_equal(s.panels(), rec0)
    assert np.array_equal(xs2, xs) and np.array_equal(ps2, ps)
    assert abs(qw2.sum() - qw.sum()) < 1e-4 * abs(qw.sum())
