"""Seeded decks of the benchmark workloads, shared by bench.py and the tests.

Every builder returns the library's own deck structures (``SimParams`` and a
list of ``SpeciesParams``, see ``sim_params``/``species_params``) or plain
arrays; nothing here runs on the device.

* ``field_deck``: the field evaluation alone at scale (N targets = sources).
* ``landau_deck``: two-species weak Landau damping, 65x129 points per species
  (the CPU-runnable plain FP64 deck).
* ``adaptive_two_stream_deck``: adaptive two-stream with ions and the treecode.
* ``landau_batch``: independent Landau decks of uneven sizes for one batch.
* ``request_stream``: Poisson arrivals of seeded decks for the serving loop.
"""
from __future__ import annotations

import numpy as np

from . import sim_params, species_params

L0 = 4 * np.pi


def field_deck(n: int, seed: int = 0):
    """(x, w) of n points: x uniform in [0, 4pi) plus 0.05 sin(2pi u / L),
    wrapped; w = q f_M(v) 1e-4 with v ~ U[-6, 6], q = -1 for the first half
    and +1 for the second."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(0, L0, n)
    x = np.mod(u + 0.05 * np.sin(2 * np.pi / L0 * u), L0)
    v = rng.uniform(-6, 6, n)
    q = np.where(np.arange(n) < n // 2, -1.0, 1.0)
    w = q * np.exp(-v * v / 2) / np.sqrt(2 * np.pi) * 1e-4
    return x, w


def landau_deck(initial_height: int = 5, p_height: int = 1):
    """Weak Landau damping with ions on [0, 4pi): electrons q = -1, m = 1 with
    f = (1 + 0.01 cos 0.5x) exp(-v^2/2)/sqrt(2pi); ions q = +1, m = 1836, a
    Maxwellian of vth = 1/sqrt(1836); v in [-6, 6]; 32x64 quadratic panels
    (65x129 points) per species at the default heights; eps 0.05, direct sum,
    trapezoid weights, dt 0.1, remesh every step, no refinement."""
    sim = sim_params(x_min=0.0, x_max=L0, bcs=0, quadrature=0, greens_epsilon=0.05,
                     use_treecode=0, dt=0.1, remesh_period=1)
    e = species_params(q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, initial_height=initial_height,
                       p_height=p_height, ics_type=1, normalized_wavenumber=1.0, amp=0.01,
                       pth=1.0)
    i = species_params(q=1.0, m=1836.0, p_min=-6.0, p_max=6.0, initial_height=initial_height,
                       p_height=p_height, ics_type=1, normalized_wavenumber=1.0, amp=0.0,
                       pth=1.0 / np.sqrt(1836.0))
    return sim, [e, i]


def adaptive_two_stream_deck(treecode: bool = True):
    """Adaptive two-stream with ions: electron beams at v = +-2.4 (vth 0.5) with
    amplitude 1e-3 cos(0.2x) on [0, 10pi), ions of mass ratio 100 (vth 0.1); base
    64x128 panels (129x257 points) per species, electrons refined up to three levels
    beyond initial_height 6 (max_height 9) where a panel's f varies by more than 1e-3
    (290,837 electron points at the start); the treecode with theta 0.7, degree 8,
    leaves under 2000 points; dt 0.05."""
    sim = sim_params(x_min=0.0, x_max=10 * np.pi, bcs=0, quadrature=0, greens_epsilon=0.05,
                     use_treecode=1 if treecode else 0, mac=0.7, degree=8, max_source=2000,
                     dt=0.05, remesh_period=1)
    e = species_params(q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, initial_height=6, p_height=1,
                       max_height=9, do_adaptively_refine=True, amr_epsilons=[1e-3],
                       ics_type=4, normalized_wavenumber=1.0, amp=1e-3, pth=0.5, pstr=2.4)
    i = species_params(q=1.0, m=100.0, p_min=-6.0, p_max=6.0, initial_height=6, p_height=1,
                       ics_type=1, normalized_wavenumber=1.0, amp=0.0, pth=0.1)
    return sim, [e, i]


def landau_batch(n: int = 64, seed: int = 4):
    """Independent two-species Landau decks: amplitude log-uniform in [0.01, 0.5],
    k uniform in [0.3, 0.6] (one wavelength per box), ion mass ratio in {25, 100, 1836},
    base panels 16x32 to 128x256 (initial_height 4..7, p_height 1), eps 0.05, dt 0.1."""
    rng = np.random.default_rng(seed)
    decks = []
    for _ in range(n):
        amp = float(np.exp(rng.uniform(np.log(0.01), np.log(0.5))))
        k = float(rng.uniform(0.3, 0.6))
        mass = float(rng.choice([25.0, 100.0, 1836.0]))
        h = int(rng.integers(4, 8))
        sim = sim_params(x_min=0.0, x_max=2 * np.pi / k, bcs=0, quadrature=0,
                         greens_epsilon=0.05, use_treecode=0, dt=0.1, remesh_period=1)
        e = species_params(q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, initial_height=h, p_height=1,
                           ics_type=1, normalized_wavenumber=1.0, amp=amp, pth=1.0)
        i = species_params(q=1.0, m=mass, p_min=-6.0, p_max=6.0, initial_height=h, p_height=1,
                           ics_type=1, normalized_wavenumber=1.0, amp=0.0,
                           pth=1.0 / np.sqrt(mass))
        decks.append((sim, [e, i]))
    return decks


def request_stream(n: int, seed: int = 5):
    """Seeded two-species decks: weak Landau (amp 0.01), strong Landau (amp 0.5) or
    bump-on-tail (beam density 0.1 at v=4.5, vth 0.5), sizes log-uniform in about
    5k..500k base points, adaptive refinement of the electrons up to 2 levels;
    arrival times of a Poisson process at 20 decks/s. Returns
    [(arrival time, kind, (sim, species)), ...]."""
    rng = np.random.default_rng(seed)
    shapes = [(4, 0), (4, 1), (5, 0), (5, 1), (6, 0), (6, 1), (7, 0), (7, 1), (8, 0)]
    sizes = np.array([2 * ((1 << (h + 1)) + 1) * ((1 << (h + p + 1)) + 1) for h, p in shapes])
    out = []
    t = 0.0
    for _ in range(n):
        t += rng.exponential(1.0 / 20.0)
        target = np.exp(rng.uniform(np.log(5e3), np.log(5e5)))
        h, ph = shapes[int(np.argmin(np.abs(np.log(sizes) - np.log(target))))]
        kind = str(rng.choice(["weak", "strong", "bump"]))
        amp = {"weak": 0.01, "strong": 0.5, "bump": 0.01}[kind]
        k = 0.5 if kind != "bump" else 0.3
        sim = sim_params(x_min=0.0, x_max=2 * np.pi / k, bcs=0, quadrature=0,
                         greens_epsilon=0.05, use_treecode=0, dt=0.1, remesh_period=1)
        e = species_params(q=-1.0, m=1.0, p_min=-8.0, p_max=8.0, initial_height=h,
                           p_height=ph, max_height=h + 2, do_adaptively_refine=True,
                           amr_epsilons=[0.1], ics_type=11 if kind == "bump" else 1,
                           normalized_wavenumber=1.0, amp=amp, pth=1.0,
                           beam_frac=0.1 if kind == "bump" else 0.0, beam_v=4.5,
                           beam_vth=0.5)
        i = species_params(q=1.0, m=1836.0, p_min=-8.0, p_max=8.0, initial_height=h,
                           p_height=ph, ics_type=1, normalized_wavenumber=1.0, amp=0.0,
                           pth=1.0 / np.sqrt(1836.0))
        out.append((t, kind, (sim, [e, i])))
    return out
