"""Driven and open-boundary decks in the device step loop (fs_sim_*), against
the oracle's step loop (itself pinned by single-particle dynamics in
test_oracle_drive_open.py).

* External drives (ExternalTanh / Sine / Logistic, AMRSimulation_field.cpp:
  70-72, 169-171): every field evaluation of a step adds a(t) sin(k x - omega
  t) at the reference's stage times; after every step the trees and grids
  bit-exact, f within 1e-12, E within 1e-12 max|E| (+ 1e-14 of the terms'
  magnitude), diagnostics within 1e-10 of their scale. The drive's sine is
  the device's and the host's: they may differ in the last bit, far inside
  the bar.
* Open boundaries (bcs 1): the open-boundary direct sum
  (E_MQ_DirectSum_openbcs) for every evaluation, no periodic shift in the
  remesh; with use_treecode the reference's loader takes the periodic direct
  sum (AMRSimulation_load.cpp:61-73), and so does the device.
* A batch of driven decks (graph-captured step: the drive table is updated
  in device memory before every replay) equals each deck run alone eagerly,
  bit for bit, and driven decks joining a running batch at different times
  keep their own clocks.
* Open-boundary decks are FP64 only: a mixed batch refuses them.
"""
import numpy as np
import pytest

import multi_species_amd as ms
import oracle
from test_gpu_sim import compare, contribution_scale, to_ms

pytestmark = pytest.mark.gpu


def driven_deck(kind, Am, dt=0.1, remesh_period=1, tracer=False):
    """The two-species Landau deck with a drive; tracer: one species of
    negligible charge (q = m = 1e-10, a 1% density perturbation so that its
    own field is resolved before the drive acts) instead, so that long steps
    stay well conditioned (its points move in the drive alone)."""
    deck = oracle.landau_deck(initial_height=4, p_height=1, dt=dt)
    if tracer:
        deck.species = [oracle.Species(name="tracer", q=1e-10, m=1e-10, p_min=-2.0, p_max=2.0,
                                       initial_height=4, p_height=1, ics_type=1,
                                       normalized_wavenumber=1.0, amp=0.01, pth=1.0)]
    deck.remesh_period = remesh_period
    deck.use_external_field, deck.external_type = 1, kind
    deck.Am, deck.k, deck.omega = Am, 0.5, 0.37
    return deck


def _run_against_oracle(gpu_ctx, deck, steps, periodic=True):
    osim = oracle.OracleSim(deck)
    gsim = ms.Simulation(gpu_ctx, *to_ms(deck))
    S = 0.0
    if periodic:
        pts = [osim.species(k).points() for k in range(len(deck.species))]
        xs = np.concatenate([p[0] for p in pts])
        ws = np.concatenate([p[3] for p in pts])
        S = contribution_scale(xs, xs, ws, deck.x_max - deck.x_min, deck.greens_epsilon)
    compare(gsim, osim, len(deck.species), 0, e_scale=S)
    worst = 0.0
    for step in range(1, steps + 1):
        osim.step()
        gsim.step()
        assert gsim.iter == osim.iter == step
        worst = max(worst, compare(gsim, osim, len(deck.species), step, e_scale=S,
                                   expect_grid=(step % deck.remesh_period == 0)))
    gsim.close()
    return worst


# amplitudes that act within the run: the sine drive from t = 0, the tanh
# one slowly from t = 0, the logistic one from t = 10 (13 steps of 1.0 on the
# tracer deck)
@pytest.mark.parametrize("kind,Am,dt,remesh_period,steps,tracer", [
    (1, 0.4, 0.1, 1, 7, False), (1, 0.4, 0.1, 2, 7, False), (0, 40.0, 0.1, 1, 7, False),
    (2, 0.4, 1.0, 1, 13, True)])
def test_driven_deck_matches_oracle(gpu_ctx, kind, Am, dt, remesh_period, steps, tracer):
    deck = driven_deck(kind, Am, dt, remesh_period, tracer)
    print(f"drive {kind} Am {Am} dt {dt} remesh_period {remesh_period}", flush=True)
    worst = _run_against_oracle(gpu_ctx, deck, steps)
    print(f"  worst rel E/f diff {worst:.2e}")


def open_deck(use_treecode=0):
    deck = oracle.landau_deck(initial_height=4, p_height=1, dt=0.1)
    deck.bcs = 1
    deck.use_treecode, deck.mac, deck.degree, deck.max_source = use_treecode, 0.7, 8, 200
    return deck


def test_open_boundary_deck_matches_oracle(gpu_ctx):
    _run_against_oracle(gpu_ctx, open_deck(), 6, periodic=False)


def test_open_deck_with_treecode_takes_the_periodic_direct_sum(gpu_ctx):
    deck = open_deck(use_treecode=1)
    _run_against_oracle(gpu_ctx, deck, 4, periodic=False)
    # the same as a periodic-field deck would give at the first evaluation:
    # E at the start equals the periodic direct sum of the initial points
    g = ms.Simulation(gpu_ctx, *to_ms(deck))
    pts = [g.points(k) for k in range(2)]
    x = np.concatenate([p[0] for p in pts])
    w = np.concatenate([p[3] for p in pts])
    e = np.concatenate([p[4] for p in pts])
    ref = oracle.direct_sum(x, x, w, deck.x_max - deck.x_min, deck.greens_epsilon)
    assert np.max(np.abs(e - ref)) <= 1e-12 * np.max(np.abs(ref)) + 1e-14 * contribution_scale(
        x, x, w, deck.x_max - deck.x_min, deck.greens_epsilon)
    g.close()


def _state(s, d):
    first, nsp, _, _ = s.deck_info(d)
    return [s.points(first + k) for k in range(nsp)]


def test_driven_batch_equals_each_alone(gpu_ctx):
    decks = [driven_deck(1, 0.4), driven_deck(0, 40.0), oracle.landau_deck(initial_height=4),
             driven_deck(2, 0.4)]
    ms_decks = [to_ms(d) for d in decks]
    b = ms.Simulation.batch(gpu_ctx, ms_decks, ms.PRECISION_F64)   # graph-captured step
    b.step(4)
    got = [_state(b, d) for d in range(len(decks))]
    b.close()
    for d, deck in enumerate(ms_decks):
        s = ms.Simulation.batch(gpu_ctx, [deck], ms.PRECISION_F64)
        s.set_options(ms.PRECISION_F64, use_graph=False)           # eager
        s.step(4)
        ref = _state(s, 0)
        s.close()
        for k in range(len(ref)):
            for a in range(5):
                assert np.array_equal(got[d][k][a], ref[k][a]), (d, k, a)


def test_driven_decks_joining_later_keep_their_own_time(gpu_ctx):
    """A driven deck that joins after 3 steps of another is at its own t = 0:
    after 2 more steps it equals the same deck run alone for 2 steps."""
    a, b = to_ms(driven_deck(1, 0.4)), to_ms(driven_deck(1, 0.4))
    s = ms.Simulation.batch(gpu_ctx, [a], ms.PRECISION_F64)
    s.step(3)
    s.add_deck(*b)
    s.step(2)
    got = _state(s, 1)
    s.close()
    r = ms.Simulation.batch(gpu_ctx, [b], ms.PRECISION_F64)
    r.step(2)
    ref = _state(r, 0)
    r.close()
    for k in range(len(ref)):
        for i in range(5):
            assert np.array_equal(got[k][i], ref[k][i]), (k, i)


def test_open_deck_refused_in_mixed_precision(gpu_ctx):
    with pytest.raises(ms.FarrsightError, match="FP64 only"):
        ms.Simulation.batch(gpu_ctx, [to_ms(open_deck())], ms.PRECISION_MIXED)
    s = ms.Simulation.batch(gpu_ctx, [to_ms(oracle.landau_deck(initial_height=4))],
                            ms.PRECISION_MIXED)
    with pytest.raises(ms.FarrsightError, match="FP64 only"):
        s.add_deck(*to_ms(open_deck()))
    assert s.n_decks == 1
    s.step()
    s.close()
    f = ms.Simulation.batch(gpu_ctx, [to_ms(open_deck())], ms.PRECISION_F64)
    with pytest.raises(ms.FarrsightError, match="FP64 only"):
        f.set_options(ms.PRECISION_MIXED)
    f.step()   # still FP64 and stepping
    f.close()
