"""The oracle's step loop with an external drive and with open boundaries,
pinned by single-particle dynamics (AMRSimulation_step.cpp:122-263,
AMRSimulation_field.cpp:70-72, 164-173, AMRSimulation_load.cpp:49-136).

A species with q = m = 1e-10 carries charges of order 1e-10: its self-field is
negligible, so every mesh point moves as a single particle in the field the
deck adds to it, and with no remesh (remesh_period beyond the run) each point
must follow the classical RK4 integration of that one particle, computed here
with numpy from the drive's closed forms:

* a driven periodic deck: dp/dt = (q/m) a(t) sin(k x - omega t), the four
  RK4 stages at t, t + dt/2, t + dt/2, t + dt (the first from the field left
  by the previous step). A wrong stage time, a drive added twice or missing
  from the field after a step, or the wrong amplitude law shifts p by
  O(dt a) -- far above the 1e-8 bar;
* an open-boundary deck: E = 0.5 sum_j w_j z / sqrt(z^2 + eps^2) - x, so with
  negligible charges dp/dt = -(q/m) x, the harmonic oscillator. Selecting
  the periodic sum instead (no -x background) leaves the points at rest.
"""
import numpy as np
import pytest

import oracle


def _tiny_species():
    return oracle.Species(name="tracer", q=1e-10, m=1e-10, p_min=-2.0, p_max=2.0,
                          initial_height=3, p_height=1, ics_type=1, normalized_wavenumber=1.0,
                          amp=0.0, pth=1.0)


def _amplitude(kind, t, Am, k):
    if kind == 1:                      # ExternalSine, FieldStructure.cpp:57-75
        if t < 50:
            return Am * np.sin(t * np.pi / 100)
        if t < 150:
            return Am
        if t < 200:
            return Am * np.cos((t - 150) * np.pi / 100)
        return 0.0
    if kind == 2:                      # ExternalLogistic, :86-101
        if t < 60:
            return Am / (1.0 + np.exp(-40.0 * (t - 10.0)))
        return Am * (1.0 - 1.0 / (1.0 + np.exp(-40.0 * (t - 110.0))))

    def drive(s):                      # ExternalTanh, :17-39
        return 0.5 * (np.tanh((s - 69) / 20) - np.tanh((s - 307) / 20))
    eps = drive(0.0)
    return (drive(t) - eps) / (1 - eps) * Am * k


def _rk4(x, p, accel, t, dt, steps):
    for _ in range(steps):
        k1x, k1p = p, accel(x, t)
        k2x, k2p = p + 0.5 * dt * k1p, accel(x + 0.5 * dt * k1x, t + 0.5 * dt)
        k3x, k3p = p + 0.5 * dt * k2p, accel(x + 0.5 * dt * k2x, t + 0.5 * dt)
        k4x, k4p = p + dt * k3p, accel(x + dt * k3x, t + dt)
        x = x + dt / 6.0 * (k1x + 2 * k2x + 2 * k3x + k4x)
        p = p + dt / 6.0 * (k1p + 2 * k2p + 2 * k3p + k4p)
        t += dt
    return x, p


# the logistic drive switches on at t = 10: steps of 2 reach it
@pytest.mark.parametrize("kind,Am,dt", [(1, 0.4, 0.5), (2, 0.4, 2.0), (0, 40.0, 0.5)])
def test_driven_points_follow_single_particle_rk4(kind, Am, dt):
    k, omega, steps = 0.5, 0.37, 6
    deck = oracle.Deck(x_min=0.0, x_max=4 * np.pi, bcs=0, quadrature=0, greens_epsilon=0.2,
                       dt=dt, remesh_period=1000, use_external_field=1, external_type=kind,
                       Am=Am, k=k, omega=omega, species=[_tiny_species()])
    sim = oracle.OracleSim(deck)
    x0, p0, _, _ = sim.species(0).points()
    for _ in range(steps):
        sim.step()
    x1, p1, _, _ = sim.species(0).points()
    acc = lambda x, t: _amplitude(kind, t, Am, k) * np.sin(k * x - omega * t)
    xe, pe = _rk4(x0.copy(), p0.copy(), acc, 0.0, dt, steps)
    scale = max(1e-3, np.max(np.abs(pe - p0)))
    assert np.max(np.abs(p1 - pe)) <= 1e-8 * scale + 1e-13, np.max(np.abs(p1 - pe))
    assert np.max(np.abs(x1 - xe)) <= 1e-8 * scale * dt * steps + 1e-13
    # the drive did act: without it the points would not have moved in p
    assert np.max(np.abs(pe - p0)) > 1e-4


def test_open_boundary_points_oscillate():
    dt, steps = 0.1, 10
    deck = oracle.Deck(x_min=0.0, x_max=4 * np.pi, bcs=1, quadrature=0, greens_epsilon=0.2,
                       dt=dt, remesh_period=1000, species=[_tiny_species()])
    sim = oracle.OracleSim(deck)
    x0, p0, _, _ = sim.species(0).points()
    for _ in range(steps):
        sim.step()
    x1, p1, _, _ = sim.species(0).points()
    xe, pe = _rk4(x0.copy(), p0.copy(), lambda x, t: -x, 0.0, dt, steps)
    assert np.max(np.abs(p1 - pe)) <= 1e-8 * np.max(np.abs(x0))
    assert np.max(np.abs(x1 - xe)) <= 1e-8 * np.max(np.abs(x0))
    assert np.max(np.abs(p1 - p0)) > 0.1   # the -x background acted
