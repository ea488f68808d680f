"""GPU parity of the remesh interpolation (part B) against the oracle.

The oracle restates interpolate_to_initial_xps / interpolate_from_mesh
(AMRStructure_interp.cpp) with the reference's tie rules; the GPU keeps
the same expression order with contraction off, so f agrees bit for bit
except at points whose sorted order among identical keys differs
(checked here to stay within 1e-15 of the oracle). Meshes: the weak-Landau
base mesh and adaptively refined meshes (hanging nodes, -1 neighbours),
advected by a shear plus a smooth warp.
"""
import numpy as np
import pytest

import multi_species_amd as ms
import oracle

pytestmark = pytest.mark.gpu

L0 = 4 * np.pi


def advect(xs, ps, t, warp):
    x = xs + t * ps + warp * np.sin(0.5 * xs) * t
    p = ps + warp * np.cos(0.5 * xs + 0.3) * t
    return x, p


def species(h=4, ph=1, adaptive=False, eps=(), max_h=None, q=-1.0, pth=1.0, amp=0.05):
    d = oracle.landau_deck(initial_height=h, p_height=ph)
    sp = d.species[0]
    sp.q, sp.amp, sp.pth = q, amp, pth
    sp.do_adaptively_refine = adaptive
    sp.amr_epsilons = list(eps)
    sp.max_height = max_h
    return d, sp


def grid_of(d, sp):
    """Base-grid points of a fresh non-adaptive mesh of the same heights."""
    import copy
    sp2 = copy.deepcopy(sp)
    sp2.do_adaptively_refine = False
    sp2.amr_epsilons = []
    g = oracle.OracleSpecies(d, sp2)
    xs, ps, _, _ = g.points()
    nx = 2 ** (sp.initial_height + 1) + 1
    np_ = 2 ** (sp.initial_height + sp.p_height + 1) + 1
    assert xs.size == nx * np_
    return xs, ps, nx, np_


def old_mesh(d, sp, t, warp):
    s = oracle.OracleSpecies(d, sp)
    xs, ps, fs, _ = s.points()
    ax, ap = advect(xs, ps, t, warp)
    s.set_points(ax, ap)
    s.copy_to_old()
    return s, s.panels(), ax, ap, fs


@pytest.mark.parametrize("t,warp", [(0.0, 0.0), (0.1, 0.0), (0.1, 0.3), (0.5, 0.5), (2.0, 0.2)])
@pytest.mark.parametrize("h,ph", [(3, 1), (4, 1), (5, 0)])
def test_grid_interp_matches_oracle(gpu_ctx, t, warp, h, ph):
    d, sp = species(h, ph)
    s, panels, ax, ap, fs = old_mesh(d, sp, t, warp)
    gx, gp, nx, np_ = grid_of(d, sp)
    f_orc, _, _ = s.interp_to_grid(gx, gp, nx, np_)
    f_gpu = ms.remesh_interp_grid(gpu_ctx, panels, ax, ap, fs, gx, gp, nx, np_, 0.0, L0, True,
                                  s.f_beyond_boundary)
    same = np.mean(f_gpu == f_orc)
    err = np.max(np.abs(f_gpu - f_orc)) / np.max(np.abs(f_orc))
    print(f"h={h} ph={ph} t={t} warp={warp}: bitwise-equal fraction {same:.6f}, max rel {err:.2e}")
    assert err <= 1e-15
    assert same >= 0.995


@pytest.mark.parametrize("adaptive_eps,max_h", [((1e-3,), 6), ((2e-3, 0.5), 7)])
def test_points_interp_on_adaptive_mesh(gpu_ctx, adaptive_eps, max_h):
    d, sp = species(4, 1, adaptive=True, eps=adaptive_eps, max_h=max_h, amp=0.3, pth=0.6)
    s, panels, ax, ap, fs = old_mesh(d, sp, 0.2, 0.3)
    nref = int(np.sum((panels[:, 17] & 8) != 0))
    assert nref > 21  # refined beyond the uniform base
    rng = np.random.default_rng(3)
    x = rng.uniform(-1.0, L0 + 1.0, 20000)
    p = rng.uniform(-6.5, 6.5, 20000)
    f_orc = s.interp_from_mesh(x, p)
    f_gpu = ms.remesh_interp_points(gpu_ctx, panels, ax, ap, fs, x, p, 0.0, L0, True,
                                    s.f_beyond_boundary)
    assert np.array_equal(f_gpu, f_orc)
    # and the grid path on the same refined mesh (neighbour walk across levels)
    gx, gp, nx, np_ = grid_of(d, sp)
    f_orc_g, _, _ = s.interp_to_grid(gx, gp, nx, np_)
    f_gpu_g = ms.remesh_interp_grid(gpu_ctx, panels, ax, ap, fs, gx, gp, nx, np_, 0.0, L0, True,
                                    s.f_beyond_boundary)
    assert np.max(np.abs(f_gpu_g - f_orc_g)) <= 1e-15 * np.max(np.abs(f_orc_g))


def test_beyond_boundary_and_range_error(gpu_ctx):
    d, sp = species(3, 1)
    s, panels, ax, ap, fs = old_mesh(d, sp, 0.0, 0.0)
    fb = s.f_beyond_boundary
    f = ms.remesh_interp_points(gpu_ctx, panels, ax, ap, fs, np.array([1.0, 2.0]),
                                np.array([7.0, -9.0]), 0.0, L0, True, fb)
    assert np.all(f == fb)
    with pytest.raises(ms.FarrsightError, match="too many shifts"):
        ms.remesh_interp_points(gpu_ctx, panels, ax, ap, fs, np.array([1.0 + 20 * L0]),
                                np.array([0.0]), 0.0, L0, True, fb)


def test_rejects_malformed_panels(gpu_ctx):
    d, sp = species(3, 1)
    s, panels, ax, ap, fs = old_mesh(d, sp, 0.0, 0.0)
    bad = panels.copy()
    bad[5, 3] = ax.size + 7
    with pytest.raises(ms.FarrsightError, match="out of range"):
        ms.remesh_interp_points(gpu_ctx, bad, ax, ap, fs, np.array([1.0]), np.array([0.0]), 0.0,
                                L0, True, 0.0)
