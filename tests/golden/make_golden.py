"""Writes tests/golden/field_golden.npz and open_golden.npz from the
reference's own field solvers.

The reference's src/FieldStructure.cpp compiles standalone; oracle/Makefile
builds it (unmodified, from the reference tree) into oracle/_ref/libfsref.so
behind oracle/ref_harness.cpp. This script evaluates E_MQ_DirectSum and
E_MQ_Treecode (field_golden.npz), E_MQ_DirectSum_openbcs and the external
drive fields ExternalTanh / ExternalSine / ExternalLogistic (open_golden.npz)
on small seeded inputs and stores inputs and outputs, so the tests can pin the
oracle and the GPU path where the reference tree is absent.

Run: make -C oracle ref && python tests/golden/make_golden.py [field|open]
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "field_golden.npz")
OUT_OPEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "open_golden.npz")


def maxwell_weights(rng, n, dA=1e-4):
    v = rng.uniform(-6, 6, n)
    f = np.exp(-v * v / 2) / np.sqrt(2 * np.pi)
    q = np.where(np.arange(n) < n // 2, -1.0, 1.0)
    return q * f * dA


def cases():
    L = 4 * np.pi
    rng = np.random.default_rng(20240601)
    out = {}

    # 1. c2-like deck, small: perturbed uniform x, half electron/half ion weights
    n = 777
    u = rng.uniform(0, L, n)
    x = np.mod(u + 0.05 * np.sin(2 * np.pi / L * u), L)
    out["ds_c2like"] = dict(t=x, s=x, w=maxwell_weights(rng, n), L=L, eps=0.05)
    out["ds_c2like_eps001"] = dict(t=x, s=x, w=maxwell_weights(rng, n), L=L, eps=0.01)

    # 2. unwrapped positions (RK4 never wraps x) and distinct targets
    ns, nt = 513, 301
    s = rng.uniform(-L, 2 * L, ns)
    t = rng.uniform(-0.5 * L, 1.5 * L, nt)
    out["ds_unwrapped"] = dict(t=t, s=s, w=rng.normal(size=ns) * 1e-3, L=L, eps=0.05)

    # 3. exact coincidences: z = 0 and z = +-1/2 (the reference's three skips)
    base = np.linspace(0, L, 16, endpoint=False)
    s = np.concatenate([base, base + L / 2, base[:4]])
    out["ds_coincident"] = dict(t=s.copy(), s=s, w=rng.normal(size=s.size), L=L, eps=0.05)

    # 4. the reduced set after a uniform remesh: 65 unique x per species, twice
    gx = np.linspace(0, L, 65)
    xs = np.concatenate([gx, gx])
    w = np.concatenate([-(1 + 0.01 * np.cos(0.5 * gx)) * 0.1, np.full(65, 0.1)])
    out["ds_reduced_grid"] = dict(t=xs, s=xs, w=w, L=L, eps=0.05)

    # 5. a different period and regularization
    L2 = 10 * np.pi
    s = rng.uniform(0, L2, 400)
    out["ds_L10pi"] = dict(t=s, s=s, w=rng.normal(size=400) * 1e-3, L=L2, eps=0.2)

    # treecode cases (targets are the sources)
    n = 1500
    u = rng.uniform(0, L, n)
    x = np.mod(u + 0.05 * np.sin(2 * np.pi / L * u), L)
    w = maxwell_weights(rng, n)
    out["tc_c2like"] = dict(s=x, w=w, L=L, eps=0.05, mac=0.7, degree=8, max_source=100)
    out["tc_deg4_mac05"] = dict(s=x, w=w, L=L, eps=0.05, mac=0.5, degree=4, max_source=64)
    s = rng.uniform(-2.0, L + 1.0, 900)  # negative x stay negative under fmod
    out["tc_negative_x"] = dict(s=s, w=rng.normal(size=900) * 1e-3, L=L, eps=0.1, mac=0.7,
                                degree=6, max_source=50)
    out["tc_single_leaf"] = dict(s=x[:300], w=w[:300], L=L, eps=0.05, mac=0.7, degree=8,
                                 max_source=2000)
    return out


def open_cases():
    """Open-boundary sums (op_*) and external drives (ex_*)."""
    rng = np.random.default_rng(20241203)
    L = 4 * np.pi
    out = {}
    n = 640
    u = rng.uniform(0, L, n)
    x = u + 0.05 * np.sin(2 * np.pi / L * u)
    out["op_c2like"] = dict(t=x, s=x, w=maxwell_weights(rng, n), eps=0.05)
    ns, nt = 500, 333
    s = rng.uniform(-L, 2 * L, ns)
    t = rng.uniform(-0.5 * L, 1.5 * L, nt)
    out["op_distinct"] = dict(t=t, s=s, w=rng.normal(size=ns) * 1e-3, eps=0.2)
    base = np.linspace(-2.0, 2.0, 33)
    out["op_coincident"] = dict(t=base, s=np.concatenate([base, base[::3]]),
                                w=rng.normal(size=base.size + base[::3].size), eps=0.01)
    xe = rng.uniform(0, 10 * np.pi, 257)
    e0 = rng.normal(size=257) * 1e-2
    for typ, name in ((0, "tanh"), (1, "sine"), (2, "logistic")):
        for t_ in (0.0, 12.5, 55.0, 100.0, 180.0, 250.0, 400.0):
            out[f"ex_{name}_t{t_:g}"] = dict(t=xe, es0=e0, params=np.array(
                [t_, typ, 0.4, 0.26, 0.37]))
    return out


def main_open():
    flat = {}
    for name, c in open_cases().items():
        if name.startswith("op_"):
            e = oracle.ref_direct_sum_open(c["t"], c["s"], c["w"], c["eps"])
            for k in ("t", "s", "w"):
                flat[f"{name}__{k}"] = c[k]
            flat[f"{name}__params"] = np.array([c["eps"]])
        else:
            t_, typ, Am, k, om = c["params"]
            e = oracle.ref_external_field_add(c["es0"], c["t"], t_, int(typ), Am, k, om)
            flat[f"{name}__t"] = c["t"]
            flat[f"{name}__es0"] = c["es0"]
            flat[f"{name}__params"] = c["params"]
        flat[f"{name}__es"] = e
    np.savez_compressed(OUT_OPEN, **flat)
    print(f"wrote {OUT_OPEN} ({os.path.getsize(OUT_OPEN)} bytes)")


def main():
    flat = {}
    for name, c in cases().items():
        if name.startswith("ds_"):
            e = oracle.ref_direct_sum(c["t"], c["s"], c["w"], c["L"], c["eps"])
            for k in ("t", "s", "w"):
                flat[f"{name}__{k}"] = c[k]
            flat[f"{name}__params"] = np.array([c["L"], c["eps"]])
        else:
            e = oracle.ref_treecode(c["s"], c["w"], c["L"], c["eps"], c["mac"], c["degree"],
                                    c["max_source"])
            flat[f"{name}__s"] = c["s"]
            flat[f"{name}__w"] = c["w"]
            flat[f"{name}__params"] = np.array([c["L"], c["eps"], c["mac"], c["degree"],
                                                c["max_source"]])
        flat[f"{name}__es"] = e
    np.savez_compressed(OUT, **flat)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what in ("all", "field"):
        main()
    if what in ("all", "open"):
        main_open()
