#!/usr/bin/env python3
"""Benchmark: periodic regularized E-field evaluation at N = 2^20 (FP64).

Workload (the field-evaluation-alone-at-scale deck): N = 2^20 targets =
sources, x uniform in [0, 4 pi) plus a 0.05 sin perturbation, weights
q f(v) dA with f Maxwellian in v ~ U[-6, 6], dA = 1e-4, half electron- and
half ion-signed, eps = 0.05. One step = one full direct-sum evaluation
(N^2 pair interactions) through fs_emq_periodic_f64_dev on device-resident
inputs; with targets = sources the library evaluates each pair once and
gives it to both points. Prints one JSON line.

The roofline block uses HIP events (torch.cuda.Event on the stream the
library launches on) around the timed launches; the kernel is VALU-bound
(no MFMA: not a contraction), so it is priced at 12 FLOP per kernel
evaluation with one accumulation (14 with the second accumulation of the
pair-once form) times the evaluations the library reports, against
MI355X's FP64 vector peak. The CPU baseline times the reference's own
compiled E_MQ_DirectSum (oracle/_ref) on a bounded target sample on this
host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from multi_species_amd.decks import (adaptive_two_stream_deck as c3_deck,  # noqa: E402
                                     field_deck as c2_inputs,
                                     landau_batch as c4_decks,
                                     landau_deck as c1_deck,
                                     request_stream as c5_requests)

L0 = 4 * np.pi
FLOP_PER_PAIR = 12
FP64_PEAK_TFLOPS = 78.6
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "direct_f64_n1048576_traffic.txt")


def read_traffic():
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(x, w, eps, budget_s: float = 12.0):
    """Reference E_MQ_DirectSum on a target sample x all N sources."""
    import oracle

    kind = "reference"
    lib = oracle.ref_lib(required=False)
    fn = oracle.ref_direct_sum if lib is not None else oracle.direct_sum
    if lib is None:
        kind = "port"
    n = x.size
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    probe = 256
    t0 = time.perf_counter()
    fn(x[:probe], x, w, L0, eps)
    dt = time.perf_counter() - t0
    rate = probe * n / dt
    m = int(min(n, max(probe, rate * budget_s / n)))
    m = max(probe, (m // 256) * 256)
    t0 = time.perf_counter()
    fn(x[:m], x, w, L0, eps)
    dt = time.perf_counter() - t0
    return {
        "value": m * n / dt,
        "unit": "pair-interactions/s",
        "cores": cores,
        "kind": kind,
        "sample": f"{m} targets x {n} sources of the same deck (eps={eps}), "
                  f"{dt:.1f} s on {cores} host threads",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--eps", type=float, default=0.05)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["field", "open", "treecode", "c1", "c3", "c4", "c5"],
                    default="field",
                    help="field: N=2^20 direct-sum evaluation (default); c1: the two-species "
                         "weak-Landau deck, 65x129 points per species, RK4 dt=0.1 with a remesh "
                         "every step (a step = one RK4 step + remesh); treecode: the treecode "
                         "field of the N=2^20 deck; c3: adaptive two-stream "
                         "with ions and the treecode; c4: 64 independent "
                         "two-species Landau decks (16x32 to 128x256 panels) stepped as one "
                         "batch in mixed precision, against the same decks one by one in FP64; "
                         "c5: Poisson arrivals of decks served with continuous batching; "
                         "open: the open-boundary direct sum (E_MQ_DirectSum_openbcs) of the "
                         "N=2^20 deck")
    ap.add_argument("--precision", choices=["f64", "mixed"], default="f64",
                    help="c1: field precision of the step loop")
    ap.add_argument("--decks", type=int, default=64, help="c4: decks in the batch")
    ap.add_argument("--no-compare", action="store_true",
                    help="c4: only the batched steps (no FP64 comparison runs, no error checks; "
                         "for profiling the batched step alone)")
    ap.add_argument("--requests", type=int, default=12, help="c5: requests served")
    ap.add_argument("--req-steps", type=int, default=50, help="c5: steps per request")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration "
                         "for the HBM traffic figures) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design note, not a measured path.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import multi_species_amd as ms

    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy << 17, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 8}))
        return

    if args.workload == "open":
        return bench_open(args)
    if args.workload == "c1":
        return bench_c1(args)
    if args.workload == "treecode":
        return bench_treecode(args)
    if args.workload == "c3":
        return bench_c3(args)
    if args.workload == "c4":
        return bench_c4(args)
    if args.workload == "c5":
        return bench_c5(args)

    n = args.n
    x, w = c2_inputs(n)
    dev = torch.device("cuda:0")
    tx = torch.from_numpy(x).to(dev)
    tw = torch.from_numpy(w).to(dev)
    te = torch.empty(n, dtype=torch.float64, device=dev)
    ctx = ms.Context(0)
    # a dedicated default-priority stream: the library launches on it and the
    # HIP events below are recorded on it
    stream = torch.cuda.Stream(device=dev)
    sptr = stream.cuda_stream
    assert sptr, "need a non-null stream handle"

    def step():
        ctx.emq_periodic_f64_dev(te.data_ptr(), tx.data_ptr(), n, tx.data_ptr(), tw.data_ptr(),
                                 n, L0, args.eps, stream=sptr)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps
    pairs = float(n) * float(n)
    value = pairs * args.steps / wall
    # targets = sources: the library evaluates each pair once (the kernel is
    # odd) and gives it to both points, two accumulations per evaluation
    evals = float(ctx.f64_pair_evals())
    flop_per_eval = FLOP_PER_PAIR + (2 if evals < pairs else 0)
    achieved_tflops = flop_per_eval * evals / (kernel_ms * 1e-3) / 1e12
    traffic = read_traffic()

    result = {
        "metric": "E-field pair-interactions/s",
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "periodic regularized E-field direct sum, N=2^20 targets=sources "
                        "(x uniform on [0,4pi) + 0.05 sin, Maxwellian weights dA=1e-4, "
                        "half electron/half ion), one evaluation per step",
            "n": n, "L": L0, "eps": args.eps, "precision": "plain FP64",
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved_tflops,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel_ms": kernel_ms,
            "kernel_evals_per_launch": evals,
            "flop_per_eval": flop_per_eval,
            # 12 FLOP per interaction: the rate a kernel evaluating all N^2
            # ordered pairs would need for the same interactions/s
            "interaction_equiv_tflops": FLOP_PER_PAIR * pairs / (kernel_ms * 1e-3) / 1e12,
        },
    }
    # the mixed-precision form on the same resident deck, against the FP64 result
    tm = torch.empty_like(te)

    def step_mixed():
        ctx.emq_periodic_mixed_dev(tm.data_ptr(), tx.data_ptr(), n, tx.data_ptr(),
                                   tw.data_ptr(), n, L0, args.eps, stream=sptr)

    for _ in range(args.warmup):
        step_mixed()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step_mixed()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall_m = time.perf_counter() - t0
    mixed_ms = ev0.elapsed_time(ev1) / args.steps
    err = float((tm - te).abs().max() / te.abs().max())
    result["mixed"] = {
        "value": pairs * args.steps / wall_m,
        "unit": "pair-interactions/s",
        "ms_per_step": wall_m / args.steps * 1e3,
        "gpu_ms": mixed_ms,
        "speedup_vs_f64": kernel_ms / mixed_ms,
        "rel_err_vs_f64": err,
        "tuned_shape": dict(zip(("targets_per_thread", "wgs_per_cu"), ctx.mixed_shape())),
        "precision": "FP64 wrap + 64-bit fixed-point displacement; FP64 step/neutralizing "
                     "part from sorted prefix sums; FP64 band |z| < max(2^-6, 16 d, 4000 d^2), "
                     "d = eps/L (Legendre expansions of 6 to 12 terms, by distance, of compact "
                     "one-sided ranges of sorted sources, exact pairs elsewhere); beyond it "
                     "FP32 8-moment expansions of the largest compact range (2048, 512, 128 or "
                     "32 sorted sources), else the FP64 expansion where a range is compact "
                     "enough for it, else FP32 pairs; compensated FP64 accumulation",
    }
    # the same with the error target 4e-9 (the band from each call's own probe)
    ctx.set_mixed_target(4e-9)
    step_mixed()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step_mixed()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall_g = time.perf_counter() - t0
    ctx.set_mixed_target(0.0)
    result["mixed"]["governed_4e-9"] = {
        "value": pairs * args.steps / wall_g,
        "gpu_ms": ev0.elapsed_time(ev1) / args.steps,
        "rel_err_vs_f64": float((tm - te).abs().max() / te.abs().max()),
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(x, w, args.eps)
    ctx.close()
    print(json.dumps(result))


FLOP_PER_OPEN_PAIR = 7  # sub, FMA (2), rsqrt (1), product, FMA (2)
# (the CPU baseline of this workload is the reference's E_MQ_DirectSum_openbcs)


def bench_open(args):
    """The open-boundary direct sum (E_MQ_DirectSum_openbcs) of the N=2^20 deck,
    targets = sources: each unordered pair evaluated once and given to both
    points (the kernel is odd); the rate counts n^2 interaction-equivalents."""
    import torch

    import multi_species_amd as ms

    n = args.n
    x, w = c2_inputs(n)
    dev = torch.device("cuda:0")
    tx = torch.from_numpy(x).to(dev)
    tw = torch.from_numpy(w).to(dev)
    te = torch.empty(n, dtype=torch.float64, device=dev)
    ctx = ms.Context(0)
    stream = torch.cuda.Stream(device=dev)
    sptr = stream.cuda_stream

    def step():
        ctx.emq_open_f64_dev(te.data_ptr(), tx.data_ptr(), n, tx.data_ptr(), tw.data_ptr(), n,
                             args.eps, stream=sptr)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps
    pairs = float(n) * float(n)
    # targets = sources: each pair evaluated once and given to both points
    # (a second accumulation per evaluation)
    evals = float(ctx.f64_pair_evals())
    flop_per_eval = FLOP_PER_OPEN_PAIR + (2 if evals < pairs else 0)
    achieved = flop_per_eval * evals / (kernel_ms * 1e-3) / 1e12
    result = {
        "metric": "E-field pair-interactions/s",
        "value": pairs * args.steps / wall,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "open-boundary regularized E-field direct sum "
                               "(E_MQ_DirectSum_openbcs), N=2^20 targets=sources of the "
                               "periodic deck's positions and weights, one evaluation per step",
                   "n": n, "eps": args.eps},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel_ms": kernel_ms, "kernel_evals_per_launch": evals,
                     "flop_per_eval": flop_per_eval,
                     "interaction_equiv_tflops":
                         FLOP_PER_OPEN_PAIR * pairs / (kernel_ms * 1e-3) / 1e12},
    }
    if not args.no_cpu_baseline:
        import oracle

        lib = oracle.ref_lib(required=False)
        fn = oracle.ref_direct_sum_open if lib is not None else oracle.direct_sum_open
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        t0 = time.perf_counter()
        fn(x[:256], x, w, args.eps)
        rate = 256 * n / (time.perf_counter() - t0)
        m = max(256, int(min(n, rate * 10.0 / n)) // 256 * 256)
        t0 = time.perf_counter()
        fn(x[:m], x, w, args.eps)
        dt = time.perf_counter() - t0
        result["cpu_baseline"] = {
            "value": m * n / dt, "unit": "pair-interactions/s", "cores": cores,
            "kind": "reference" if lib is not None else "port",
            "sample": f"{m} targets x {n} sources of the same deck, {dt:.1f} s on {cores} "
                      "host threads"}
    print(json.dumps(result))


def bench_c1(args):
    """The whole step loop of the two-species weak-Landau deck on the device."""
    import multi_species_amd as ms

    ctx = ms.Context(0)
    sim_p, sps = c1_deck()
    prec = ms.PRECISION_MIXED if args.precision == "mixed" else ms.PRECISION_F64
    sim = ms.Simulation.batch(ctx, [(sim_p, sps)], prec)
    npts = sum(sim.counts(k)[0] for k in range(2))
    for _ in range(args.warmup):
        sim.step()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        sim.step()
        lat.append(time.perf_counter() - t1)
    wall = time.perf_counter() - t0
    lat = np.array(lat) * 1e3
    result = {
        "metric": "RK4 point-steps/s",
        "value": npts * args.steps / wall,
        "unit": "point-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == ms.PRECISION_F64 else "f64+f32", "data": "synthetic",
        "config": {"workload": "two-species weak Landau damping deck (electrons q=-1 m=1 "
                               "amp 0.01 k=0.5; ions q=1 m=1836 vth=1/sqrt(1836)), 32x64 "
                               "quadratic panels per species (65x129 points), eps=0.05, "
                               "direct sum, RK4 dt=0.1, remesh every step",
                   "points": npts, "precision": args.precision},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
    }
    if not args.no_cpu_baseline:
        import oracle

        deck = oracle.landau_deck(initial_height=5, p_height=1)
        osim = oracle.OracleSim(deck)
        # steps until about 12 s of host work (at least 3)
        k = 0
        t0 = time.perf_counter()
        while k < 3 or time.perf_counter() - t0 < 12.0:
            osim.step()
            k += 1
        dt = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        result["cpu_baseline"] = {
            "value": npts * k / dt, "unit": "point-steps/s", "cores": cores, "kind": "port",
            "sample": f"{k} steps of the same deck with the oracle's OpenMP FP64 restatement "
                      f"of the reference step loop, {dt:.1f} s on {cores} host threads",
        }
    sim.close()
    ctx.close()
    print(json.dumps(result))


def bench_treecode(args):
    """E_MQ_Treecode on the N=2^20 deck (theta 0.7, degree 8, leaves under 2000 points):
    tree and interaction lists on the host, moments and far/near sums on the device,
    host arrays in and out. Reported as kernel evaluations/s (far + near) and as the
    equivalent direct-sum pairs/s; the reference's own CPU treecode (serial) on a 2^16
    sample of the deck beside it."""
    import multi_species_amd as ms

    n = args.n
    x, w = c2_inputs(n)
    mac, deg, msrc = 0.7, 8, 2000
    ctx = ms.Context(0)
    es, ev = ctx.emq_treecode_f64(x, w, L0, args.eps, mac, deg, msrc, return_evals=True)
    for _ in range(max(args.warmup - 1, 0)):
        ctx.emq_treecode_f64(x, w, L0, args.eps, mac, deg, msrc)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.emq_treecode_f64(x, w, L0, args.eps, mac, deg, msrc)
    dt = (time.perf_counter() - t0) / args.steps
    ef = ctx.emq_periodic_f64(x, x, w, L0, args.eps)
    err = float(np.max(np.abs(es - ef)) / np.max(np.abs(ef)))
    evals = ev[0] + ev[1]
    result = {
        "metric": "treecode kernel evaluations/s",
        "value": evals / dt,
        "unit": "kernel-evaluations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "E_MQ_Treecode of the N=2^20 deck (c2 inputs), theta 0.7, "
                               "degree 8, max_source 2000, host arrays (tree on the host)",
                   "n": n, "eps": args.eps},
        "evals": {"far": int(ev[0]), "near": int(ev[1])},
        "equivalent_pairs_per_s": n * n / dt,
        "rel_e_err_vs_direct_f64": err,
    }
    if not args.no_cpu_baseline:
        import oracle

        m = 1 << 16
        xs, ws = c2_inputs(m)
        if oracle.ref_lib(required=False) is not None:
            oracle.treecode(xs, ws, L0, args.eps, mac, deg, msrc)
            st = oracle.treecode_stats()
            t0 = time.perf_counter()
            oracle.ref_treecode(xs, ws, L0, args.eps, mac, deg, msrc)
            dtc = time.perf_counter() - t0
            result["cpu_baseline"] = {
                "value": (st["far_evals"] + st["near_evals"]) / dtc,
                "unit": "kernel-evaluations/s", "cores": 1, "kind": "reference",
                "sample": f"the reference's E_MQ_Treecode (far and near sums serial) on a 2^16 "
                          f"point deck of the same kind, {dtc:.1f} s; evaluation counts from "
                          f"the oracle's restatement of the same tree"}
    ctx.close()
    print(json.dumps(result))


def bench_c3(args):
    """The adaptive two-stream deck stepped with the treecode field (host tree,
    device moments and sums) and adaptive remesh; at the end the treecode E of the
    final state against the FP64 direct sum of the same points."""
    import multi_species_amd as ms

    ctx = ms.Context(0)
    s = ms.Simulation(ctx, *c3_deck())
    counts0 = [s.counts(k)[0] for k in range(2)]
    for _ in range(args.warmup):
        s.step()
    lat, pts = [], 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        n = sum(s.counts(k)[0] for k in range(2))
        t1 = time.perf_counter()
        s.step()
        lat.append(time.perf_counter() - t1)
        pts += n
    wall = time.perf_counter() - t0
    counts = [s.counts(k)[0] for k in range(2)]
    sim_p = c3_deck()[0]
    L = sim_p.x_max - sim_p.x_min
    P = [s.points(k) for k in range(2)]
    x = np.concatenate([p[0] for p in P])
    w = np.concatenate([p[3] for p in P])
    es = np.concatenate([p[4] for p in P])
    ef = ctx.emq_periodic_f64(x, x, w, L, sim_p.greens_epsilon)
    err = float(np.max(np.abs(es - ef)) / np.max(np.abs(ef)))
    s.close()
    ctx.close()
    lat = np.array(lat) * 1e3
    print(json.dumps({
        "metric": "RK4 point-steps/s",
        "value": pts / sum(lat) * 1e3,
        "unit": "point-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": float(np.mean(lat)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "adaptive two-stream with ions (beams +-2.4, vth 0.5, amp 1e-3, "
                               "L=10pi; ions m=100), 129x257 base points per species, electrons "
                               "refined up to 3 levels on panel variation > 1e-3, treecode "
                               "theta 0.7 degree 8 max_source 2000, dt 0.05, remesh every step",
                   "points_start": counts0, "points_end": counts},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "treecode_rel_e_err_vs_direct_f64": err,
    }))


def bench_c4(args):
    """64 decks as one batch (mixed precision, graph-captured step) against the same decks
    stepped one by one in plain FP64."""
    import torch  # noqa: F401  (device init through the library)

    import multi_species_amd as ms

    ctx = ms.Context(0)
    decks = c4_decks(args.decks)
    b = ms.Simulation.batch(ctx, decks, ms.PRECISION_MIXED)
    npts = sum(b.deck_info(d)[3] for d in range(b.n_decks))
    for _ in range(args.warmup):
        b.step()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        b.step()
        lat.append(time.perf_counter() - t1)
    wall = time.perf_counter() - t0
    value = npts * args.steps / wall
    # algorithmic bytes of one remesh of every species (target x, p in and f out;
    # old x, p, f; nine point ids and child/flag words per old panel)
    remesh_bytes = 0
    for k in range(sum(b.deck_info(d)[1] for d in range(b.n_decks))):
        n_k, p_k = b.counts(k)
        remesh_bytes += 24 * n_k + 24 * n_k + 48 * p_k
    if args.no_compare:
        b.close()
        ctx.close()
        lat = np.array(lat) * 1e3
        print(json.dumps({
            "metric": "batched RK4 point-steps/s", "value": value, "unit": "point-steps/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
            "config": {"workload": f"{args.decks} Landau decks, batched step only",
                       "decks": args.decks, "points": int(npts)},
            "latency_ms": {"p50": float(np.percentile(lat, 50)),
                           "p99": float(np.percentile(lat, 99))},
            "remesh_algorithmic_bytes_per_step": remesh_bytes}))
        return
    # E error of the mixed batch against the FP64 field of the same points,
    # every deck: the served E at the end (remesh field), and the full-field
    # evaluation at moved points (the positions RK stages see) after three
    # steps of the same decks with remesh every second step; informational,
    # the E difference to FP64 runs of the same decks after the same steps
    # (trajectory drift included)
    def field_err(sim_p, pts):
        x = np.concatenate([p[0] for p in pts])
        w = np.concatenate([p[3] for p in pts])
        es = np.concatenate([p[4] for p in pts])
        ef = ctx.emq_periodic_f64(x, x, w, sim_p.x_max - sim_p.x_min, sim_p.greens_epsilon)
        return float(np.max(np.abs(es - ef)) / np.max(np.abs(ef))), es

    err, err_stage = 0.0, 0.0
    mixed_states = []
    for d, (sim_p, _) in enumerate(decks):
        first, nsp, _, _ = b.deck_info(d)
        e, es = field_err(sim_p, [b.points(first + k) for k in range(nsp)])
        err = max(err, e)
        mixed_states.append(es)
    b.close()
    decks2 = c4_decks(args.decks)
    for sim_p, _ in decks2:
        sim_p.remesh_period = 2
    v = ms.Simulation.batch(ctx, decks2, ms.PRECISION_MIXED)
    v.step(3)  # two full-field evaluations governed by measured errors
    for d, (sim_p, _) in enumerate(decks2):
        first, nsp, _, _ = v.deck_info(d)
        err_stage = max(err_stage, field_err(sim_p, [v.points(first + k) for k in range(nsp)])[0])
    v.close()
    f = ms.Simulation.batch(ctx, decks, ms.PRECISION_F64)
    f.step(args.warmup + args.steps)
    drift = 0.0
    for d in range(f.n_decks):
        first, nsp, _, _ = f.deck_info(d)
        ef = np.concatenate([f.points(first + k)[4] for k in range(nsp)])
        drift = max(drift, float(np.max(np.abs(mixed_states[d] - ef)) / np.max(np.abs(ef))))
    f.close()
    # plain FP64, one deck at a time
    t_plain = 0.0
    for sim_p, sps in decks:
        s = ms.Simulation.batch(ctx, [(sim_p, sps)], ms.PRECISION_F64)
        s.step()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            s.step()
        t_plain += time.perf_counter() - t1
        s.close()
    plain = npts * args.steps / t_plain
    lat = np.array(lat) * 1e3
    result = {
        "metric": "batched RK4 point-steps/s",
        "value": value,
        "unit": "point-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": f"{args.decks} independent two-species Landau decks in one batch "
                               "(amp log-uniform [0.01,0.5], k uniform [0.3,0.6], ion mass "
                               "ratio in {25,100,1836}, 16x32..128x256 panels per species, "
                               "eps 0.05, dt 0.1, remesh every step), mixed precision, "
                               "graph-captured step",
                   "decks": args.decks, "points": int(npts)},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "plain_f64_one_by_one": {"value": plain, "unit": "point-steps/s",
                                 "speedup_of_batched_mixed": value / plain},
        "remesh_algorithmic_bytes_per_step": remesh_bytes,
        "rel_e_err_vs_f64": err,
        "stage_field_rel_e_err_vs_f64": err_stage,
        "rel_e_diff_vs_f64_run": drift,
    }
    ctx.close()
    print(json.dumps(result))

def bench_c5(args):
    """Serving: requests arrive as a Poisson process and are stepped with continuous
    batching (mixed precision); per-step latency percentiles and aggregate rate."""
    import multi_species_amd as ms

    reqs = c5_requests(args.requests)
    ctx = ms.Context(0)
    srv = ms.Server(ctx, hbm_budget_bytes=200e9, max_active=0, precision=ms.PRECISION_MIXED)
    ids = []
    t0 = time.perf_counter()
    for t_arr, _, deck in reqs:
        delay = t_arr - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(*deck, args.req_steps))
    for rid in ids:
        srv.wait(rid, 1200)
    wall = time.perf_counter() - t0
    st = srv.stats()
    finals = {rid: [srv.result(rid, k) for k in range(2)] for rid in ids}
    srv.close()
    # E error of the served field (mixed, governed) on the final states of the
    # smallest three decks: the served E against the FP64 field of the same
    # points; and the one-call mixed entry point at its minimum band and with
    # the error target 4e-9 on the same points
    errs, errs_min, errs_chk = [], [], []
    order = sorted(ids, key=lambda r: sum(finals[r][k][0].size for k in range(2)))
    for rid in order[:3]:
        x = np.concatenate([finals[rid][k][0] for k in range(2)])
        w = np.concatenate([finals[rid][k][3] for k in range(2)])
        es = np.concatenate([finals[rid][k][4] for k in range(2)])
        L = reqs[ids.index(rid)][2][0].x_max
        ef = ctx.emq_periodic_f64(x, x, w, L, 0.05)
        rel = lambda e: float(np.max(np.abs(e - ef)) / np.max(np.abs(ef)))
        errs.append(rel(es))
        ctx.set_mixed_target(0.0)
        errs_min.append(rel(ctx.emq_periodic_mixed(x, x, w, L, 0.05, same=True)))
        ctx.set_mixed_target(4e-9)
        errs_chk.append(rel(ctx.emq_periodic_mixed(x, x, w, L, 0.05, same=True)))
    ctx.set_mixed_target(0.0)
    # the same requests in plain FP64, one deck at a time (3 timed steps per deck
    # after one untimed: the rate at which a one-deck-at-a-time FP64 loop runs them)
    plain_pts, plain_s = 0, 0.0
    for _, _, deck in reqs:
        s = ms.Simulation.batch(ctx, [deck], ms.PRECISION_F64)
        n = sum(s.counts(k)[0] for k in range(len(deck[1])))
        s.step()
        t1 = time.perf_counter()
        s.step(3)
        plain_s += (time.perf_counter() - t1) / 3
        plain_pts += n
        s.close()
    plain = plain_pts / plain_s
    ctx.close()
    result = {
        "metric": "served RK4 point-steps/s",
        "value": st["point_steps_per_s"],
        "unit": "point-steps/s",
        "n_gpus": 1, "steps": int(st["steps"]), "warmup": 0,
        "ms_per_step": st["step_p50_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": f"{args.requests} requests (Poisson arrivals at 20/s) of seeded "
                               "two-species decks (weak/strong Landau or bump-on-tail), "
                               "5k..500k base points log-uniform, adaptive refinement up to 2 "
                               f"levels, {args.req_steps} steps each at dt=0.1, HBM budget "
                               "200 GB, continuous batching, mixed precision",
                   "requests": args.requests},
        "latency_ms": {"p50": st["step_p50_ms"], "p99": st["step_p99_ms"]},
        "served": {k: st[k] for k in ("submitted", "completed", "failed", "steps",
                                      "point_steps", "busy_s", "span_s")},
        "wall_s": wall,
        "rel_e_err_vs_f64": max(errs) if errs else None,
        "plain_f64_one_by_one": {"value": plain, "unit": "point-steps/s",
                                 "speedup_of_served_mixed": st["point_steps_per_s"] / plain,
                                 "sample": "3 timed steps of each request's deck in plain FP64"},
        "entry_point_rel_e_err_vs_f64": {"min_band": max(errs_min) if errs_min else None,
                                         "target_4e-9": max(errs_chk) if errs_chk else None},
    }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
