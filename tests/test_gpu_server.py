"""The serving loop (fs_server): continuous batching of decks.

Bars:
* every served deck's final state is bitwise the state of the same deck run
  alone for the same number of steps (same precision), whatever else shared
  its steps: decks arrive while others are stepping, join and leave between
  steps, and the mix includes an adaptively refined deck (eager steps);
* admission respects the slot limit (never more resident decks than
  max_active) and the HBM budget (a deck larger than the budget fails, the
  others complete);
* the statistics count every step and request.
"""
import time

import numpy as np
import pytest

import multi_species_amd as ms

pytestmark = pytest.mark.gpu


def landau(h, amp=0.01, k=1.0, mass=1836.0, eps=0.05, dt=0.1, L=4 * np.pi):
    sim = ms.sim_params(x_min=0.0, x_max=L, bcs=0, quadrature=0, greens_epsilon=eps,
                        use_treecode=0, dt=dt, remesh_period=1)
    e = ms.species_params(q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, initial_height=h, p_height=1,
                          ics_type=1, normalized_wavenumber=k, amp=amp, pth=1.0)
    i = ms.species_params(q=1.0, m=mass, p_min=-6.0, p_max=6.0, initial_height=h, p_height=1,
                          ics_type=1, normalized_wavenumber=k, amp=0.0, pth=1.0 / np.sqrt(mass))
    return sim, [e, i]


def adaptive_two_stream():
    sim = ms.sim_params(x_min=0.0, x_max=10 * np.pi, bcs=0, quadrature=0, greens_epsilon=0.05,
                        use_treecode=0, dt=0.05, remesh_period=1)
    e = ms.species_params(q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, initial_height=3, p_height=1,
                          max_height=5, do_adaptively_refine=True, amr_epsilons=[2e-3],
                          ics_type=4, normalized_wavenumber=1.0, amp=1e-3, pth=0.5, pstr=2.4)
    i = ms.species_params(q=1.0, m=100.0, p_min=-1.0, p_max=1.0, initial_height=3, p_height=1,
                          ics_type=1, amp=0.0, pth=0.1)
    return sim, [e, i]


def alone(ctx, deck, steps, precision):
    s = ms.Simulation.batch(ctx, [deck], precision)
    s.step(steps)
    out = [s.points(k) for k in range(len(deck[1]))]
    s.close()
    return out


@pytest.mark.parametrize("precision", [ms.PRECISION_F64, ms.PRECISION_MIXED])
def test_served_equals_alone(precision):
    decks = [(landau(4), 5), (landau(3, amp=0.3, k=1.5, mass=25.0), 3),
             (adaptive_two_stream(), 4), (landau(5, amp=0.1, mass=100.0), 6),
             (landau(3, eps=0.08, dt=0.05), 2)]
    ctx = ms.Context(0)
    srv = ms.Server(ctx, hbm_budget_bytes=8e9, max_active=0, precision=precision)
    ids = []
    for k, (deck, n) in enumerate(decks):
        ids.append(srv.submit(*deck, n))
        if k == 1:
            # let the first two start stepping before the others arrive
            t0 = time.time()
            while srv.status(ids[0])[1] < 1 and time.time() - t0 < 60:
                time.sleep(0.005)
    for rid in ids:
        srv.wait(rid, 300)
    results = {rid: [srv.result(rid, k) for k in range(2)] for rid in ids}
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(decks) and st["failed"] == 0
    assert st["steps"] >= max(n for _, n in decks)
    assert st["step_p50_ms"] <= st["step_p99_ms"]
    for rid, (deck, n) in zip(ids, decks):
        ref = alone(ctx, deck, n, precision)
        for k in range(2):
            got = results[rid][k]
            for a in range(5):  # xs, ps, fs, q_ws, es
                assert np.array_equal(got[a], ref[k][a]), (precision, rid, k, a)
    ctx.close()


def test_slots_and_budget():
    ctx = ms.Context(0)
    srv = ms.Server(ctx, hbm_budget_bytes=25e6, max_active=2, precision=ms.PRECISION_MIXED)
    ids = [srv.submit(*landau(3 + (k % 2)), 3) for k in range(5)]
    big = srv.submit(*landau(7), 1)  # ~150 MB estimated: beyond the budget
    max_active = 0
    t0 = time.time()
    while time.time() - t0 < 120:
        max_active = max(max_active, srv.stats()["active"])
        if all(srv.status(r)[0] in (ms.REQ_DONE, ms.REQ_FAILED) for r in ids + [big]):
            break
        time.sleep(0.002)
    for rid in ids:
        srv.wait(rid, 120)
    assert srv.status(big)[0] == ms.REQ_FAILED
    with pytest.raises(ms.FarrsightError):
        srv.wait(big, 1)
    st = srv.stats()
    srv.close()
    ctx.close()
    assert st["completed"] == 5 and st["failed"] == 1
    assert 1 <= max_active <= 2


def test_failed_requests_leave_the_others_intact():
    """Bad decks beside good ones: a request of zero steps is refused at
    submit, a deck whose mesh cannot be built (p_max < p_min) fails in the
    builder; the good decks (one of them remeshing every second step),
    submitted before and after them, complete bitwise equal to the same decks
    run alone."""
    ctx = ms.Context(0)
    srv = ms.Server(ctx, hbm_budget_bytes=8e9, max_active=0, precision=ms.PRECISION_F64)
    good = [(landau(4), 3), (landau(3, amp=0.2, mass=25.0), 5)]
    good[1][0][0].remesh_period = 2
    with pytest.raises(ms.FarrsightError):
        srv.submit(*landau(3), 0)
    bad_mesh = landau(3)
    bad_mesh[1][0].p_max = -7.0
    ids = [srv.submit(*good[0][0], good[0][1]), srv.submit(*bad_mesh, 2),
           srv.submit(*good[1][0], good[1][1])]
    srv.wait(ids[0], 120)
    srv.wait(ids[2], 120)
    with pytest.raises(ms.FarrsightError):
        srv.wait(ids[1], 60)
    assert srv.status(ids[1])[0] == ms.REQ_FAILED
    results = {rid: [srv.result(rid, k) for k in range(2)] for rid in (ids[0], ids[2])}
    st = srv.stats()
    srv.close()
    assert st["completed"] == 2 and st["failed"] == 1
    for rid, (deck, n) in zip((ids[0], ids[2]), good):
        ref = alone(ctx, deck, n, ms.PRECISION_F64)
        for k in range(2):
            for a in range(5):
                assert np.array_equal(results[rid][k][a], ref[k][a]), (rid, k, a)
    ctx.close()


def test_refused_join_leaves_the_batch_unchanged():
    """fs_sim_add_deck refuses an open-boundary deck in a mixed-precision batch
    before anything joins: the batch keeps stepping bitwise as the deck
    alone."""
    ctx = ms.Context(0)
    deck = landau(4)
    s = ms.Simulation.batch(ctx, [deck], ms.PRECISION_MIXED)
    s.step(1)
    bad = landau(3)
    bad[0].bcs = 1
    with pytest.raises(ms.FarrsightError, match="FP64 only"):
        s.add_deck(*bad)
    assert s.n_decks == 1
    s.step(2)
    got = [s.points(k) for k in range(2)]
    s.close()
    ref = alone(ctx, deck, 3, ms.PRECISION_MIXED)
    for k in range(2):
        for a in range(5):
            assert np.array_equal(got[k][a], ref[k][a]), (k, a)
    ctx.close()


def test_fp64_admission_counts_pair_once_partials():
    """The server's estimate of an FP64 deck includes its pair-once partials:
    a budget between the estimate without them and with them refuses the
    deck (it is larger than the whole budget)."""
    deck = landau(6)
    f64, n = ms.deck_footprint(*deck, precision=ms.PRECISION_F64)
    sym = ms.pair_once_plan(n, 1 << 62)[0]
    assert sym > 0
    ctx = ms.Context(0)
    srv = ms.Server(ctx, hbm_budget_bytes=f64 - sym // 2, max_active=0,
                    precision=ms.PRECISION_F64)
    rid = srv.submit(*deck, 1)
    with pytest.raises(ms.FarrsightError):
        srv.wait(rid, 60)
    assert srv.status(rid)[0] == ms.REQ_FAILED
    ok = srv.submit(*landau(4), 1)
    srv.wait(ok, 60)
    assert srv.stats()["completed"] == 1
    srv.close()
    ctx.close()


def test_context_close_closes_its_simulations_and_servers():
    """A context closed while a simulation and a server built on it are still
    referenced closes them first (their C objects use the context until they
    are destroyed); closing them again afterwards does nothing."""
    ctx = ms.Context(0)
    s = ms.Simulation.batch(ctx, [landau(3)], ms.PRECISION_F64)
    s.step(1)
    srv = ms.Server(ctx, hbm_budget_bytes=1e9, max_active=0, precision=ms.PRECISION_F64)
    ctx.close()
    assert s.h is None and srv.h is None
    s.close()
    srv.close()
