"""Adaptive refinement on the device against the host builder.

The device passes (fs_amr.hip) must give, after every remesh, the tree the
sequential host builder gives from the same base mesh and grid f: panel
records, points, f, weights (through q_ws), the reduced-x map (through E) --
all bit for bit. The host builder is itself pinned against the oracle's
restatement of refine_panels (tests/test_mesh_build.py, test_gpu_sim.py) and
by the mathematics (tests/test_oracle_pins.py). Decks cover every
refinement criterion, Simpson quadrature, open boundaries, p-refined bases
of height 0..2, small bases whose periodic neighbours wrap, several
refining species in one launch, and the two-stream workload's own deck.
"""
import numpy as np
import pytest

import multi_species_amd as ms
from multi_species_amd import decks

pytestmark = pytest.mark.gpu


def adaptive_deck(h=4, ph=1, extra=2, eps=(2e-3,), bcs=0, quad=0, amp=0.3, kind=1, L=4 * np.pi,
                  dt=0.1, ions_adaptive=False):
    sim = ms.sim_params(x_min=0.0, x_max=L, bcs=bcs, quadrature=quad, greens_epsilon=0.05,
                        use_treecode=0, dt=dt, remesh_period=1)
    e = ms.species_params(q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, initial_height=h, p_height=ph,
                          max_height=h + extra, do_adaptively_refine=True,
                          amr_epsilons=list(eps), ics_type=kind, normalized_wavenumber=1.0,
                          amp=amp, pth=1.0 if kind != 4 else 0.5, pstr=2.4 if kind == 4 else 0.0)
    i = ms.species_params(q=1.0, m=100.0, p_min=-6.0, p_max=6.0, initial_height=h, p_height=ph,
                          max_height=h + 1, do_adaptively_refine=ions_adaptive,
                          amr_epsilons=[0.5], ics_type=1, normalized_wavenumber=1.0, amp=0.0,
                          pth=0.1)
    return sim, [e, i]


DECKS = {
    "variation": dict(extra=3),
    "two_stream": dict(kind=4, amp=1e-3, eps=(1e-3,), L=10 * np.pi, dt=0.05),
    "all_criteria": dict(eps=(0.02, 0.5, 1e3, 0.05, 0.02), amp=0.5),
    "ratio_only": dict(eps=(1e9, 1e9, 1.3), amp=0.5),
    "simpson_open": dict(quad=1, bcs=1, amp=0.2),
    "p_height_0": dict(ph=0, h=5, extra=3),
    "p_height_2": dict(ph=2, h=3, extra=4),
    "small_base": dict(h=2, ph=0, extra=3, amp=0.5),
    "smallest_base": dict(h=1, ph=1, extra=3, amp=0.5),
    "both_species": dict(ions_adaptive=True, amp=0.5),
}


def compare_sims(a, b, nsp, tag):
    for k in range(nsp):
        assert a.counts(k) == b.counts(k), f"{tag} species {k}: counts"
        assert np.array_equal(a.panels(k), b.panels(k)), f"{tag} species {k}: panels"
        for name, u, v in zip(("x", "p", "f", "q_w", "E"), a.points(k), b.points(k)):
            assert np.array_equal(u, v), f"{tag} species {k}: {name}"


@pytest.mark.parametrize("name", sorted(DECKS))
def test_device_refinement_equals_host(gpu_ctx, name):
    sim, sps = adaptive_deck(**DECKS[name])
    dev = ms.Simulation(gpu_ctx, sim, sps)
    host = ms.Simulation(gpu_ctx, sim, sps)
    host.set_refinement(on_device=False)
    compare_sims(dev, host, 2, f"{name} start")
    most = 0
    for step in range(1, 5):
        dev.step()
        host.step()
        compare_sims(dev, host, 2, f"{name} step {step}")
        most = max(most, dev.counts(0)[0])
    nd, nh = dev.refinement_counts()
    assert nh == 0 and nd == 4, (nd, nh)
    assert host.refinement_counts() == (0, 4)
    # the remeshes refined beyond the base grid
    e = sps[0]
    base = ((1 << (e.initial_height + 1)) + 1) * ((1 << (e.initial_height + e.p_height + 1)) + 1)
    assert most > base
    dev.close()
    host.close()


def test_batch_of_refining_decks(gpu_ctx):
    """Several refining species in one launch, with static decks beside them."""
    ds = [adaptive_deck(**DECKS[n]) for n in ("variation", "all_criteria", "p_height_0")]
    ds.append(decks.landau_batch(1, seed=2)[0])
    ds.append(adaptive_deck(**DECKS["both_species"]))
    dev = ms.Simulation.batch(gpu_ctx, ds)
    host = ms.Simulation.batch(gpu_ctx, ds)
    host.set_refinement(on_device=False)
    nsp = dev.n_species
    for step in range(1, 4):
        dev.step()
        host.step()
        compare_sims(dev, host, nsp, f"batch step {step}")
    assert dev.refinement_counts() == (3, 0)
    dev.close()
    host.close()


def test_two_stream_workload_deck(gpu_ctx):
    """The adaptive two-stream deck at its own size (base 129x257 points per
    species, electrons refined three levels, ~3e5 points), treecode field."""
    sim, sps = decks.adaptive_two_stream_deck(treecode=True)
    dev = ms.Simulation(gpu_ctx, sim, sps)
    host = ms.Simulation(gpu_ctx, sim, sps)
    host.set_refinement(on_device=False)
    for step in range(1, 3):
        dev.step()
        host.step()
        compare_sims(dev, host, 2, f"two-stream step {step}")
    assert dev.refinement_counts() == (2, 0)
    assert dev.counts(0)[0] > 250_000
    dev.close()
    host.close()


def test_served_requests_refine_on_device(gpu_ctx):
    """Bump-on-tail and strong Landau requests of the serving stream, stepped
    in one batch with device refinement, equal to the host builder's."""
    reqs = [r for r in decks.request_stream(12, seed=5) if r[1] in ("bump", "strong")][:3]
    ds = [r[2] for r in reqs]
    dev = ms.Simulation.batch(gpu_ctx, ds)
    host = ms.Simulation.batch(gpu_ctx, ds)
    host.set_refinement(on_device=False)
    for step in range(1, 3):
        dev.step()
        host.step()
        compare_sims(dev, host, dev.n_species, f"served step {step}")
    assert dev.refinement_counts() == (2, 0)
    dev.close()
    host.close()
