"""Batches of independent decks in one fs_sim (fs_batch_create).

Bars:
* a deck's state in a batch is bitwise that of the same deck run alone, in
  FP64 and in mixed precision, for decks of different sizes, domains, time
  steps and eps;
* the captured step graph: every replay bitwise equal to the eager step, with
  host reads and eager work between replays;
* the mixed batch against the FP64 runs of the same decks: E within 1e-8
  (relative, normwise) after the first step;
* the governed step loop: E of weak decks within 1e-8 of the FP64 field of
  the same points, step after step, at moved points (full-field evaluation)
  and after remesh;
* continuous batching: a deck that joins a running batch evolves bitwise as
  it does alone; removing a deck leaves the others' states unchanged.
"""
import numpy as np
import pytest

import multi_species_amd as ms

pytestmark = pytest.mark.gpu


def landau(h, ph, L=4 * np.pi, dt=0.1, eps=0.05, amp=0.01, k=1.0, mass=1836.0):
    sim = ms.sim_params(x_min=0.0, x_max=L, bcs=0, quadrature=0, greens_epsilon=eps,
                        use_treecode=0, dt=dt, remesh_period=1)
    e = ms.species_params(q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, initial_height=h, p_height=ph,
                          ics_type=1, normalized_wavenumber=k, amp=amp, pth=1.0)
    i = ms.species_params(q=1.0, m=mass, p_min=-6.0, p_max=6.0, initial_height=h, p_height=ph,
                          ics_type=1, normalized_wavenumber=k, amp=0.0, pth=1.0 / np.sqrt(mass))
    return sim, [e, i]


def two_stream(h, ph):
    sim = ms.sim_params(x_min=0.0, x_max=10 * np.pi, bcs=0, quadrature=1, greens_epsilon=0.1,
                        use_treecode=0, dt=0.05, remesh_period=1)
    e = ms.species_params(q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, initial_height=h, p_height=ph,
                          ics_type=4, normalized_wavenumber=1.0, amp=1e-3, pth=0.5, pstr=2.4)
    i = ms.species_params(q=1.0, m=100.0, p_min=-2.0, p_max=2.0, initial_height=h, p_height=ph,
                          ics_type=1, normalized_wavenumber=1.0, amp=0.0, pth=0.1)
    return sim, [e, i]


DECKS = [landau(3, 1), landau(4, 1, dt=0.05, eps=0.08, amp=0.3, k=1.5, mass=25.0),
         two_stream(4, 0), landau(5, 1)]


def state(sim, first, nsp):
    return [np.concatenate(sim.points(first + k)) for k in range(nsp)]


def run_alone(ctx, deck, steps, precision):
    s = ms.Simulation.batch(ctx, [deck], precision)
    s.step(steps)
    out = state(s, 0, len(deck[1]))
    s.close()
    return out


@pytest.mark.parametrize("precision", [ms.PRECISION_F64, ms.PRECISION_MIXED])
def test_batch_equals_alone(gpu_ctx, precision):
    steps = 3
    b = ms.Simulation.batch(gpu_ctx, DECKS, precision)
    assert b.n_decks == len(DECKS)
    b.step(steps)
    for d, deck in enumerate(DECKS):
        first, nsp, it, _ = b.deck_info(d)
        assert it == steps
        alone = run_alone(gpu_ctx, deck, steps, precision)
        mixed_in = state(b, first, nsp)
        for k in range(nsp):
            assert np.array_equal(mixed_in[k], alone[k]), (precision, d, k)
    b.close()


def test_graph_replays_match_eager(gpu_ctx):
    g = ms.Simulation.batch(gpu_ctx, DECKS[:3], ms.PRECISION_MIXED)
    e = ms.Simulation.batch(gpu_ctx, DECKS[:3], ms.PRECISION_MIXED)
    e.set_options(ms.PRECISION_MIXED, use_graph=False)
    for step in range(4):
        g.step()
        e.step()
        # eager device work and host reads between replays
        _ = gpu_ctx.emq_periodic_f64(np.linspace(0, 1, 100), np.linspace(0, 1, 50),
                                     np.ones(50), 1.0, 0.1)
        for k in range(6):
            assert np.array_equal(np.concatenate(g.points(k)), np.concatenate(e.points(k))), \
                (step, k)
    g.close()
    e.close()


def test_mixed_field_within_1e8_of_f64(gpu_ctx):
    f = ms.Simulation.batch(gpu_ctx, DECKS, ms.PRECISION_F64)
    m = ms.Simulation.batch(gpu_ctx, DECKS, ms.PRECISION_MIXED)
    # initial reduced-grid fields, then after one step
    for stage in range(2):
        for d in range(len(DECKS)):
            first, nsp, _, _ = f.deck_info(d)
            ef = np.concatenate([f.points(first + k)[4] for k in range(nsp)])
            em = np.concatenate([m.points(first + k)[4] for k in range(nsp)])
            r = np.max(np.abs(em - ef)) / np.max(np.abs(ef))
            print(f"stage {stage} deck {d}: mixed vs FP64 E rel {r:.2e}")
            assert r <= 1e-8
        f.step()
        m.step()
    f.close()
    m.close()


def test_governed_step_loop_error(gpu_ctx):
    """The mixed step loop sets each deck's FP64 band from the error its
    previous full-field evaluation measured. With remesh every second step,
    E after an odd step is the loop's own full-field evaluation at moved
    points (the positions an RK stage sees), after an even step the remesh
    field: both within 1e-8 of the FP64 field of the same points, for weak
    decks (max|E| small against the charges) and step after step."""
    decks = [landau(5, 1, amp=1e-3, k=0.5), landau(4, 1, amp=1e-2, k=0.5), two_stream(4, 0)]
    for sim, _ in decks:
        sim.remesh_period = 2
    m = ms.Simulation.batch(gpu_ctx, decks, ms.PRECISION_MIXED)
    for step in range(6):
        m.step()
        for d, (sim, sps) in enumerate(decks):
            first, nsp, _, _ = m.deck_info(d)
            pts = [m.points(first + k) for k in range(nsp)]
            x = np.concatenate([p[0] for p in pts])
            w = np.concatenate([p[3] for p in pts])
            es = np.concatenate([p[4] for p in pts])
            ef = gpu_ctx.emq_periodic_f64(x - sim.x_min, x - sim.x_min, w,
                                          sim.x_max - sim.x_min, sim.greens_epsilon)
            r = np.max(np.abs(es - ef)) / np.max(np.abs(ef))
            em, errs, bands = m.field_governor(d)
            print(f"step {step + 1} deck {d}: E vs FP64 of the same points {r:.2e}; governor "
                  f"max|E| {em:.3e} (FP64 {np.max(np.abs(ef)):.3e}), rel errors "
                  + " ".join(f"{b:.4f}:{e / max(em, 1e-300):.1e}" for b, e in zip(bands, errs)))
            assert r <= 1e-8, (step, d)
    m.close()


@pytest.mark.parametrize("precision", [ms.PRECISION_F64, ms.PRECISION_MIXED])
def test_continuous_batching(gpu_ctx, precision):
    b = ms.Simulation.batch(gpu_ctx, DECKS[:2], precision)
    b.step(2)
    j = b.add_deck(*DECKS[2])
    assert j == 2 and b.n_decks == 3
    b.step(3)
    first, nsp, it, _ = b.deck_info(2)
    assert it == 3
    joined = state(b, first, nsp)
    alone = run_alone(gpu_ctx, DECKS[2], 3, precision)
    for k in range(nsp):
        assert np.array_equal(joined[k], alone[k])
    # deck 0 has taken 5 steps; remove deck 1 and check deck 0 and 2 continue unchanged
    b.remove_deck(1)
    assert b.n_decks == 2
    b.step(1)
    d0 = state(b, 0, 2)
    a0 = run_alone(gpu_ctx, DECKS[0], 6, precision)
    for k in range(2):
        assert np.array_equal(d0[k], a0[k])
    first, nsp, it, _ = b.deck_info(1)
    assert it == 4
    d2 = state(b, first, nsp)
    a2 = run_alone(gpu_ctx, DECKS[2], 4, precision)
    for k in range(nsp):
        assert np.array_equal(d2[k], a2[k])
    b.close()


def _with_period(deck, period):
    sim, sps = deck
    sim = ms.sim_params(**{f: getattr(sim, f) for f, _ in ms.SimParams._fields_})
    sim.remesh_period = period
    return sim, sps


@pytest.mark.parametrize("precision", [ms.PRECISION_F64, ms.PRECISION_MIXED])
def test_mixed_remesh_periods_equal_each_alone(gpu_ctx, precision):
    """Decks that remesh every step, every second and every third step in one
    batch (graph replays for the steps where all, none and some of them
    remesh, and eagerly): after every step each deck's state equals the deck
    run alone, bit for bit (AMRSimulation_step.cpp:74-90 per deck)."""
    decks = [_with_period(DECKS[0], 1), _with_period(DECKS[1], 2), _with_period(DECKS[2], 3),
             _with_period(DECKS[3], 2)]
    g = ms.Simulation.batch(gpu_ctx, decks, precision)
    e = ms.Simulation.batch(gpu_ctx, decks, precision)
    e.set_options(precision, use_graph=False)
    alone = [ms.Simulation.batch(gpu_ctx, [d], precision) for d in decks]
    for step in range(1, 7):
        g.step()
        e.step()
        for d, deck in enumerate(decks):
            alone[d].step()
            first, nsp, it, _ = g.deck_info(d)
            assert it == step
            ref = state(alone[d], 0, nsp)
            for k in range(nsp):
                assert np.array_equal(state(g, first, nsp)[k], ref[k]), (precision, step, d, k)
                assert np.array_equal(state(e, first, nsp)[k], ref[k]), (precision, step, d, k)
    for s in [g, e] + alone:
        s.close()


def test_deck_with_other_period_joins_mid_run(gpu_ctx):
    """A deck that remeshes every second step joins a batch after 3 steps of
    a deck that remeshes every step (and one that remeshes every third): it
    keeps its own iteration count, and after 4 more steps every deck equals
    its run alone."""
    a, c = _with_period(DECKS[0], 1), _with_period(DECKS[2], 3)
    b = _with_period(DECKS[1], 2)
    s = ms.Simulation.batch(gpu_ctx, [a, c], ms.PRECISION_F64)
    s.step(3)
    s.add_deck(*b)
    s.step(4)
    for d, (deck, steps) in enumerate([(a, 7), (c, 7), (b, 4)]):
        first, nsp, it, _ = s.deck_info(d)
        assert it == steps
        ref = run_alone(gpu_ctx, deck, steps, ms.PRECISION_F64)
        got = state(s, first, nsp)
        for k in range(nsp):
            assert np.array_equal(got[k], ref[k]), (d, k)
    s.close()


def adaptive_landau(h, period, amp=0.3):
    sim, sps = landau(h, 1, amp=amp)
    sim.remesh_period = period
    sps[0].max_height = h + 3
    sps[0].do_adaptively_refine = True
    sps[0].amr_epsilons[0] = 2e-3
    sps[0].n_amr_epsilons = 1
    return sim, sps


@pytest.mark.parametrize("precision", [ms.PRECISION_F64, ms.PRECISION_MIXED])
def test_adaptive_batch_replays_graphs_between_its_remeshes(gpu_ctx, precision):
    """Adaptive decks remeshing every third step beside static decks remeshing
    every step: the steps in which no adaptive deck remeshes (static remeshes
    only, or none) replay captured graphs; every state bitwise equal to the
    same batch stepped eagerly, with host reads and eager work between
    replays."""
    decks = [adaptive_landau(4, 3), DECKS[0], adaptive_landau(3, 3, amp=0.5), DECKS[1]]
    g = ms.Simulation.batch(gpu_ctx, decks, precision)
    e = ms.Simulation.batch(gpu_ctx, decks, precision)
    e.set_options(precision, use_graph=False)
    nsp = g.n_species
    for step in range(7):
        g.step()
        e.step()
        _ = gpu_ctx.emq_periodic_f64(np.linspace(0, 1, 100), np.linspace(0, 1, 50),
                                     np.ones(50), 1.0, 0.1)
        for k in range(nsp):
            assert np.array_equal(np.concatenate(g.points(k)), np.concatenate(e.points(k))), \
                (step, k)
            assert np.array_equal(g.panels(k), e.panels(k)), (step, k)
    # steps 3 and 6 remesh the adaptive decks (eager); the other five replay
    assert g.step_counts() == (5, 2), g.step_counts()
    assert e.step_counts() == (0, 7)
    assert g.refinement_counts()[0] == 2
    g.close()
    e.close()
