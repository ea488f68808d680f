"""Checkpoints in the reference's files on the GPU step loop.

Bars: a deck written after step 3 and restarted into a new batch
(fs_sim_restart_deck) continues bitwise as the run that wrote it, for
2 more steps, in FP64 and mixed precision, for a Landau deck and an
adaptively refined two-stream deck; the deck's reduced set files hold each
species' distinct x and the field there; a served request's final state
written as a checkpoint (fs_server_write_result) and resumed by a second
request (fs_server_submit_restart) ends bitwise where the deck run alone for
the sum of the steps ends; a restart at a step that is not a remesh step is
refused.
"""
import os

import numpy as np
import pytest

import multi_species_amd as ms

pytestmark = pytest.mark.gpu


def landau(h=4):
    sim = ms.sim_params(x_min=0.0, x_max=4 * np.pi, quadrature=0, greens_epsilon=0.05, dt=0.1)
    e = ms.species_params(q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, initial_height=h, p_height=1,
                          ics_type=1, normalized_wavenumber=1.0, amp=0.05, pth=1.0)
    i = ms.species_params(q=1.0, m=100.0, p_min=-6.0, p_max=6.0, initial_height=h, p_height=1,
                          ics_type=1, normalized_wavenumber=1.0, amp=0.0, pth=0.1)
    return sim, [e, i]


def two_stream():
    sim = ms.sim_params(x_min=0.0, x_max=10 * np.pi, quadrature=0, greens_epsilon=0.05, dt=0.05)
    e = ms.species_params(q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, initial_height=3, p_height=1,
                          max_height=5, do_adaptively_refine=True, amr_epsilons=[2e-3],
                          ics_type=4, normalized_wavenumber=1.0, amp=1e-3, pth=0.5, pstr=2.4)
    i = ms.species_params(q=1.0, m=100.0, p_min=-1.0, p_max=1.0, initial_height=3, p_height=1,
                          ics_type=1, amp=0.0, pth=0.1)
    return sim, [e, i]


NAMES = ["electrons", "ions"]


def states(s, d=0):
    first, nsp, _, _ = s.deck_info(d)
    return [s.points(first + k) for k in range(nsp)]


@pytest.mark.parametrize("precision", [ms.PRECISION_F64, ms.PRECISION_MIXED])
@pytest.mark.parametrize("make", [landau, two_stream], ids=["landau", "two_stream"])
def test_restart_continues_bitwise(gpu_ctx, tmp_path, precision, make):
    d = str(tmp_path)
    deck = make()
    a = ms.Simulation.batch(gpu_ctx, [deck], precision)
    a.step(3)
    a.write_output(0, d, NAMES)
    b = ms.Simulation.empty(gpu_ctx, precision)
    assert b.restart_deck(*make(), d, NAMES, 3) == 0
    assert b.deck_info(0)[2] == 3
    for k in range(2):
        assert np.array_equal(b.panels(k), a.panels(k))
    # the restart evaluates E on the reduced set of the loaded state: the same
    # evaluation that ended step 3
    for ga, gb in zip(states(a), states(b)):
        for q in range(5):
            assert np.array_equal(ga[q], gb[q]), ("after restart", q)
    # the reduced-set files: each species' distinct x, then E there
    rx = np.fromfile(os.path.join(d, "simulation_output", "reduce_xs", "xs_3"))
    re = np.fromfile(os.path.join(d, "simulation_output", "reduced_es", "es_3"))
    ux = np.concatenate([np.unique(p[0]) for p in states(a)])
    assert np.array_equal(rx, ux) and re.size == rx.size
    xs0, _, _, _, es0 = states(a)[0]
    first = {x: e for x, e in zip(xs0[::-1], es0[::-1])}
    nu0 = np.unique(xs0).size
    assert all(re[j] == first[rx[j]] for j in range(nu0))
    a.step(2)
    b.step(2)
    for ga, gb in zip(states(a), states(b)):
        for q in range(5):
            assert np.array_equal(ga[q], gb[q]), ("after 2 more steps", q)
    a.close()
    b.close()


def test_restart_needs_a_remesh_step(gpu_ctx, tmp_path):
    d = str(tmp_path)
    sim, sps = landau(3)
    sim.remesh_period = 2
    a = ms.Simulation.batch(gpu_ctx, [(sim, sps)], ms.PRECISION_F64)
    a.step(1)
    a.write_output(0, d, NAMES)
    a.close()
    b = ms.Simulation.empty(gpu_ctx, ms.PRECISION_F64)
    with pytest.raises(ms.FarrsightError, match="not on its grid"):
        b.restart_deck(*landau(3), d, NAMES, 1)  # written between two remeshes
    with pytest.raises(ms.FarrsightError):
        b.restart_deck(*landau(3), d, NAMES, 4)  # no files of step 4
    b.close()


def test_served_checkpoint_and_resume(tmp_path):
    d = str(tmp_path)
    ctx = ms.Context(0)
    srv = ms.Server(ctx, hbm_budget_bytes=8e9, precision=ms.PRECISION_MIXED)
    r1 = srv.submit(*two_stream(), 3)
    srv.wait(r1, 120)
    srv.write_result(r1, d, NAMES)
    r2 = srv.submit_restart(*two_stream(), d, NAMES, 3, 2)
    srv.wait(r2, 120)
    got = [srv.result(r2, k) for k in range(2)]
    srv.close()
    ref = ms.Simulation.batch(ctx, [two_stream()], ms.PRECISION_MIXED)
    ref.step(5)
    for k, st in enumerate(states(ref)):
        for q in range(5):
            assert np.array_equal(got[k][q], st[q]), (k, q)
    ref.close()
    ctx.close()
