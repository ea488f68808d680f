"""Pins the oracle's step loop (AMRSimulation_step.cpp) by physics and
kinematics it must reproduce: free streaming of a field-free plasma, and
the weak Landau damping rate gamma = 0.1533 at k = 0.5 that the reference's
own diagnostics compare against (standard_interface/diagnostic_plot.py:171-174)."""
import numpy as np

import oracle


def test_free_streaming_without_field():
    # a nearly charge-free species (q = 1e-8): the acceleration q/m E is O(q^2),
    # so RK4 must reproduce x(t) = x0 + t p, p(t) = p0 to rounding
    d = oracle.landau_deck(initial_height=3, p_height=1)
    d.species = d.species[:1]
    d.species[0].q = -1e-8
    d.remesh_period = 1000
    sim = oracle.OracleSim(d)
    x0, p0, _, _ = sim.species(0).points()
    assert np.max(np.abs(sim.es(0))) < 1e-9
    for _ in range(3):
        sim.step()
    x, p, _, _ = sim.species(0).points()
    assert np.max(np.abs(p - p0)) < 1e-13
    assert np.max(np.abs(x - (x0 + 3 * d.dt * p0))) < 1e-12


def test_weak_landau_damping_rate():
    """Electrons only (the two-species benchmark deck's ion Maxwellian,
    vth = 1/sqrt(1836), is narrower than its p-grid spacing, and each remesh's
    quadratic interpolation across that spike feeds a small ion density
    perturbation back into E): the k = 0.5 mode of E must oscillate at
    Re(omega) = 1.4156 (zero crossings pi/omega apart) and its envelope decay
    at gamma = 0.1533."""
    d = oracle.landau_deck(initial_height=5, p_height=1, dt=0.1)
    d.species = d.species[:1]
    sim = oracle.OracleSim(d)
    t, amp = [], []
    for k in range(0, 91):
        if k:
            sim.step()
        xs, ps, fs, qw = sim.species(0).points()
        a = np.linalg.lstsq(np.stack([np.sin(0.5 * xs), np.cos(0.5 * xs)], 1), sim.es(0),
                            rcond=None)[0][0]
        t.append(k * d.dt)
        amp.append(a)
    t, amp = np.array(t), np.array(amp)
    zc = [t[i - 1] + d.dt * amp[i - 1] / (amp[i - 1] - amp[i])
          for i in range(1, len(amp)) if amp[i - 1] * amp[i] < 0]
    assert len(zc) >= 4
    omega = np.pi / np.mean(np.diff(zc))
    assert abs(omega - 1.4156) < 0.02 * 1.4156, omega
    ext = [i for i in range(1, len(amp) - 1)
           if abs(amp[i]) > abs(amp[i - 1]) and abs(amp[i]) >= abs(amp[i + 1])]
    gamma = -np.polyfit(t[ext], np.log(np.abs(amp[ext])), 1)[0]
    assert abs(gamma - 0.1533) < 0.08 * 0.1533, gamma
