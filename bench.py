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
host's cores (all threads, bound close, and one thread).

Beside the field the line carries the step-loop half of the metric:
"mixed" (the mixed-precision form of the same evaluation, in N^2
interaction-equivalents/s, with its VALU issue fraction), "plain_step_loop"
(the Landau deck's FP64 step loop against the oracle's OpenMP step loop on
the host), "batched" (64 decks stepped as one mixed-precision batch against
the same decks one by one in FP64: point-steps/s, p50/p99 step latency, E
error against FP64 at the same points) and "served" (a Poisson request
stream with continuous batching). --workload runs one of them on its own;
--workload remesh times the batched remesh at 2^24 grid points (64 species)
in algorithmic GB/s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

# host-core baselines: threads bound to consecutive cores (read by the OpenMP
# runtime when it loads, so before numpy, torch and the oracle)
os.environ.setdefault("OMP_PROC_BIND", "close")

import numpy as np  # noqa: E402

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


MIXED_VALU_FILE = os.path.join(ROOT, "profiles", "mixed_field_n1048576_valu_counters.md")
REMESH_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "c4_remesh_hbm.md")
REMESH24_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "remesh_2p24_hbm.md")
HBM_PEAK_GBPS = 8000.0   # MI355X HBM3E
SIMDS = 256 * 4          # MI355X: 256 CUs of 4 SIMDs
CLOCK_HZ = 2.4e9         # peak engine clock


def read_profile_value(path, key):
    """A `key: value` line of a committed profile summary (None if absent)."""
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return float(line.split(":")[1].split()[0])
    except OSError:
        pass
    return None


def read_traffic():
    return read_profile_value(TRAFFIC_FILE, "traffic_bytes_per_launch")


def host_info():
    """(CPU model, OpenMP threads) of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


class omp_threads:
    """Sets the OpenMP thread count of the host baselines (libgomp, which the
    oracle and the reference build link) for a with-block."""

    def __init__(self, n: int):
        import ctypes

        self.n = n
        self.lib = ctypes.CDLL("libgomp.so.1")
        self.lib.omp_get_max_threads.restype = ctypes.c_int

    def __enter__(self):
        self.prev = self.lib.omp_get_max_threads()
        self.lib.omp_set_num_threads(self.n)
        return self

    def __exit__(self, *exc):
        self.lib.omp_set_num_threads(self.prev)


def cpu_baseline(x, w, eps, budget_s: float = 10.0):
    """Reference E_MQ_DirectSum on a target sample x all N sources, on every
    host thread and on one."""
    import oracle

    kind = "reference"
    lib = oracle.ref_lib(required=False)
    fn = oracle.ref_direct_sum if lib is not None else oracle.direct_sum
    if lib is None:
        kind = "port"
    n = x.size
    model, cores = host_info()
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
    m1 = max(64, (m // cores) // 64 * 64)
    with omp_threads(1):
        t0 = time.perf_counter()
        fn(x[:m1], x, w, L0, eps)
        dt1 = time.perf_counter() - t0
    return {
        "value": m * n / dt,
        "unit": "pair-interactions/s",
        "cores": cores,
        "kind": kind,
        "cpu_model": model,
        "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
        "value_1_thread": m1 * n / dt1,
        "sample": f"{m} targets x {n} sources of the same deck (eps={eps}), "
                  f"{dt:.1f} s on {cores} host threads; {m1} targets on 1 thread, {dt1:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--eps", type=float, default=0.05)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--refinement", choices=["device", "host"], default="device",
                    help="where adaptive remeshes refine their trees (c3)")
    ap.add_argument("--workload", choices=["field", "open", "treecode", "c1", "c3", "c4", "c5",
                                           "remesh"],
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
                         "N=2^20 deck; remesh: the batched remesh at 2^24 target points "
                         "(32 decks, 513x513 points per species)")
    ap.add_argument("--precision", choices=["f64", "mixed"], default="f64",
                    help="c1: field precision of the step loop")
    ap.add_argument("--decks", type=int, default=64, help="c4: decks in the batch")
    ap.add_argument("--no-compare", action="store_true",
                    help="c4: only the batched steps (no FP64 comparison runs, no error checks; "
                         "for profiling the batched step alone)")
    ap.add_argument("--requests", type=int, default=40, help="c5: requests served")
    ap.add_argument("--req-steps", type=int, default=50, help="c5: steps per request")
    ap.add_argument("--blocks", default="c1,c4,c5",
                    help="field workload: the step-loop blocks added to the line "
                         "(comma-separated from c1,c4,c5; 'none' for the field alone)")
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
    if args.workload == "remesh":
        return bench_remesh(args)

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
        "metric": "E-field pair-interactions/s; batched RK4 point-steps/s + p99 step latency "
                  "@1e-8 E err",
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
    valu = read_profile_value(MIXED_VALU_FILE, "valu_instructions_per_evaluation")
    result["mixed"] = {
        # the mixed form sums ranges of sources through expansions: its rate is
        # N^2 interaction-equivalents per evaluation time, not pair evaluations
        "value": pairs * args.steps / wall_m,
        "unit": "pair-interaction-equivalents/s (N^2 per evaluation)",
        "ms_per_step": wall_m / args.steps * 1e3,
        "gpu_ms": mixed_ms,
        "roofline": {
            "bound": "valu issue",
            # wave64 VALU instructions of the whole evaluation (profile) x 4 cycles
            # each, over the evaluation's HIP-event time on every SIMD at peak clock
            "achieved": (valu * 4 / (mixed_ms * 1e-3 * SIMDS * CLOCK_HZ)) if valu else None,
            "peak": 1.0, "unit": "fraction of VALU issue slots",
            "frac": (valu * 4 / (mixed_ms * 1e-3 * SIMDS * CLOCK_HZ)) if valu else None,
            "valu_instructions_per_evaluation": valu,
            "source": os.path.relpath(MIXED_VALU_FILE, ROOT),
        },
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
    del tx, tw, te, tm
    torch.cuda.empty_cache()
    # the step-loop half of the metric: the plain FP64 step loop of the Landau
    # deck against the host cores, the batched mixed step of 64 decks against the
    # same decks one by one in FP64 (point-steps/s, p50/p99 step latency, E error),
    # and a served request stream
    blocks = [b for b in args.blocks.split(",") if b and b != "none"]
    if "c1" in blocks:
        result["plain_step_loop"] = c1_block(ctx, 200, 5, "f64", not args.no_cpu_baseline)
    if "c4" in blocks:
        result["batched"] = c4_block(ctx, 64, 10, 2)
    if "c5" in blocks:
        result["served"] = c5_block(ctx, 40, 20)
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


C1_WORKLOAD = ("two-species weak Landau damping deck (electrons q=-1 m=1 amp 0.01 k=0.5; ions "
               "q=1 m=1836 vth=1/sqrt(1836)), 32x64 quadratic panels per species (65x129 "
               "points), eps=0.05, direct sum, RK4 dt=0.1, remesh every step")


def c1_block(ctx, steps: int, warmup: int, precision: str = "f64", cpu: bool = True,
             cpu_budget_s: float = 8.0):
    """The whole step loop of the two-species weak-Landau deck on the device
    (one step = RK4 with three field evaluations, remesh, reduced-x field),
    and the oracle's OpenMP restatement of the reference step loop on the host
    cores beside it."""
    import multi_species_amd as ms

    sim_p, sps = c1_deck()
    prec = ms.PRECISION_MIXED if precision == "mixed" else ms.PRECISION_F64
    sim = ms.Simulation.batch(ctx, [(sim_p, sps)], prec)
    npts = sum(sim.counts(k)[0] for k in range(2))
    for _ in range(warmup):
        sim.step()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        sim.step()
        lat.append(time.perf_counter() - t1)
    wall = time.perf_counter() - t0
    sim.close()
    lat = np.array(lat) * 1e3
    out = {
        "value": npts * steps / wall, "unit": "point-steps/s", "steps": steps,
        "ms_per_step": wall / steps * 1e3, "points": npts, "precision": precision,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "workload": C1_WORKLOAD,
    }
    if cpu:
        import oracle

        model, cores = host_info()
        deck = oracle.landau_deck(initial_height=5, p_height=1)
        osim = oracle.OracleSim(deck)
        k = 0
        t0 = time.perf_counter()
        while k < 3 or time.perf_counter() - t0 < cpu_budget_s:
            osim.step()
            k += 1
        dt = time.perf_counter() - t0
        with omp_threads(1):
            t1 = time.perf_counter()
            osim.step()
            dt1 = time.perf_counter() - t1
        cpu_rate = npts * k / dt
        out["cpu_baseline"] = {
            "value": cpu_rate, "unit": "point-steps/s", "cores": cores, "kind": "port",
            "cpu_model": model, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "value_1_thread": npts / dt1,
            "sample": f"{k} steps of the same deck with the oracle's OpenMP FP64 restatement of "
                      f"the reference step loop, {dt:.1f} s on {cores} host threads; one more "
                      f"step on 1 thread, {dt1:.1f} s"}
        out["speedup_vs_cpu"] = out["value"] / cpu_rate
    return out


def bench_c1(args):
    """The c1 block as its own bench line."""
    import multi_species_amd as ms

    ctx = ms.Context(0)
    blk = c1_block(ctx, args.steps, args.warmup, args.precision, not args.no_cpu_baseline)
    ctx.close()
    result = {
        "metric": "RK4 point-steps/s", "value": blk["value"], "unit": "point-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": blk["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "f64" else "f64+f32", "data": "synthetic",
        "config": {"workload": C1_WORKLOAD, "points": blk["points"],
                   "precision": args.precision},
        "latency_ms": blk["latency_ms"],
    }
    for k in ("cpu_baseline", "speedup_vs_cpu"):
        if k in blk:
            result[k] = blk[k]
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
                               "degree 8, max_source 2000, host arrays (sort, tree, lists and sums on the device)",
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
    s.set_refinement(on_device=args.refinement == "device")
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
        "refinement": args.refinement,
    }))


C4_WORKLOAD = ("independent two-species Landau decks in one batch (amp log-uniform "
               "[0.01,0.5], k uniform [0.3,0.6], ion mass ratio in {25,100,1836}, 16x32..128x256 "
               "panels per species, eps 0.05, dt 0.1, remesh every step), mixed precision, "
               "graph-captured step")


DIAG_KEYS = ("charge", "momentum", "kinetic_energy", "electric_energy")


def mixed_diag_errors(ctx, ndecks: int, steps: int):
    """The same decks stepped as one batch in mixed precision and in FP64: after
    every step, each species' charge, momentum, kinetic and electric energy of the
    mixed run against the FP64 run, relative to the diagnostic's own scale (the sum
    of its terms' magnitudes: sum |q_w|, sum |p q_w| / |q|, sum p^2 |q_w| / (2 |q/m|),
    and the electric energy itself, a sum of positive terms). Max over species and
    steps; the electric-energy entry covers its whole history."""
    import multi_species_amd as ms

    decks = c4_decks(ndecks)
    sims = [ms.Simulation.batch(ctx, c4_decks(ndecks), prec)
            for prec in (ms.PRECISION_MIXED, ms.PRECISION_F64)]
    params = []
    for d, (_, sps) in enumerate(decks):
        first, nsp, _, _ = sims[1].deck_info(d)
        params += [(first + j, sps[j].q, sps[j].q / sps[j].m) for j in range(nsp)]
    worst = dict.fromkeys(DIAG_KEYS, 0.0)
    for _ in range(steps):
        for s in sims:
            s.step()
        for k, q, qm in params:
            dm, df = sims[0].diagnostics(k), sims[1].diagnostics(k)
            _, p, _, qw, _ = sims[1].points(k)
            scale = (np.sum(np.abs(qw)), np.sum(np.abs(p * qw)) / abs(q),
                     np.sum(p * p * np.abs(qw)) / (2 * abs(qm)), abs(df[3]))
            for i, key in enumerate(DIAG_KEYS):
                if scale[i] > 0:
                    worst[key] = max(worst[key], float(abs(dm[i] - df[i]) / scale[i]))
    for s in sims:
        s.close()
    return worst


def c4_block(ctx, ndecks: int, steps: int, warmup: int, compare: bool = True,
             plain_steps: int = 2, diag_steps: int = 5):
    """ndecks decks as one batch (mixed precision, graph-captured step) against the
    same decks stepped one by one in plain FP64; per-step latency of the batch, the
    mixed E against the FP64 field of the same points, and the mixed diagnostics
    against the same decks stepped in FP64 (mixed_diag_errors)."""
    import multi_species_amd as ms

    decks = c4_decks(ndecks)
    b = ms.Simulation.batch(ctx, decks, ms.PRECISION_MIXED)
    npts = sum(b.deck_info(d)[3] for d in range(b.n_decks))
    for _ in range(warmup):
        b.step()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        b.step()
        lat.append(time.perf_counter() - t1)
    wall = time.perf_counter() - t0
    value = npts * steps / wall
    # algorithmic bytes of one remesh of every species (target x, p in and f out;
    # old x, p, f; nine point ids and child/flag words per old panel)
    remesh_bytes = 0
    for k in range(sum(b.deck_info(d)[1] for d in range(b.n_decks))):
        n_k, p_k = b.counts(k)
        remesh_bytes += 24 * n_k + 24 * n_k + 48 * p_k
    lat = np.array(lat) * 1e3
    # the remesh's device work timed live with HIP events (three eager steps
    # after the timed ones), against its algorithmic bytes
    b.set_timing(True)
    rms = []
    for _ in range(3):
        b.step()
        rms.append(b.remesh_ms())
    b.set_timing(False)
    remesh_ms = float(np.median(rms))
    traffic = read_profile_value(REMESH_TRAFFIC_FILE, "remesh_counted_traffic_bytes_per_step")
    out = {
        "value": value, "unit": "point-steps/s", "steps": steps, "decks": ndecks,
        "points": int(npts), "ms_per_step": wall / steps * 1e3,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "remesh_algorithmic_bytes_per_step": remesh_bytes,
        "remesh_roofline": {
            "bound": "hbm", "achieved": remesh_bytes / (remesh_ms * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": remesh_bytes / (remesh_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS,
            "traffic": traffic, "ms": remesh_ms,
            "source": "HIP events around the remesh kernels of 3 eager steps; traffic from "
                      + os.path.relpath(REMESH_TRAFFIC_FILE, ROOT),
        },
        "workload": C4_WORKLOAD,
    }
    if not compare:
        b.close()
        return out

    # E error of the mixed batch against the FP64 field of the same points,
    # every deck: the E of the last step (remesh field), and the full-field
    # evaluation at moved points (the positions RK stages see) after one step
    # of the same decks with remesh every second step
    def field_err(sim_p, pts):
        x = np.concatenate([p[0] for p in pts])
        w = np.concatenate([p[3] for p in pts])
        es = np.concatenate([p[4] for p in pts])
        ef = ctx.emq_periodic_f64(x, x, w, sim_p.x_max - sim_p.x_min, sim_p.greens_epsilon)
        return float(np.max(np.abs(es - ef)) / np.max(np.abs(ef)))

    err, err_stage = 0.0, 0.0
    for d, (sim_p, _) in enumerate(decks):
        first, nsp, _, _ = b.deck_info(d)
        err = max(err, field_err(sim_p, [b.points(first + k) for k in range(nsp)]))
    b.close()
    decks2 = c4_decks(ndecks)
    for sim_p, _ in decks2:
        sim_p.remesh_period = 2
    v = ms.Simulation.batch(ctx, decks2, ms.PRECISION_MIXED)
    v.step(1)
    for d, (sim_p, _) in enumerate(decks2):
        first, nsp, _, _ = v.deck_info(d)
        err_stage = max(err_stage, field_err(sim_p, [v.points(first + k) for k in range(nsp)]))
    v.close()
    # plain FP64, one deck at a time (one untimed step each)
    t_plain = 0.0
    for sim_p, sps in decks:
        s = ms.Simulation.batch(ctx, [(sim_p, sps)], ms.PRECISION_F64)
        s.step()
        t1 = time.perf_counter()
        s.step(plain_steps)
        t_plain += (time.perf_counter() - t1) / plain_steps
        s.close()
    plain = npts / t_plain
    out["plain_f64_one_by_one"] = {"value": plain, "unit": "point-steps/s",
                                   "speedup_of_batched_mixed": value / plain,
                                   "sample": f"{plain_steps} timed steps of each deck alone"}
    out["rel_e_err_vs_f64"] = {"remesh_field": err, "full_field_at_moved_points": err_stage,
                               "bound": 1e-8}
    out["rel_diag_err_vs_f64"] = dict(mixed_diag_errors(ctx, ndecks, diag_steps),
                                      steps=diag_steps)
    return out


REMESH24_WORKLOAD = ("remesh of 2^24 grid points in one batched launch per kernel: 32 "
                     "two-species Landau decks (amp log-uniform [0.01,0.5], ion mass ratio in "
                     "{25,100,1836}), 256x256 quadratic panels (513x513 points) per species, "
                     "64 species, 1.68e7 target points; mixed-precision field, remesh every "
                     "step")


def remesh24_decks(n: int = 32, seed: int = 6):
    """The decks of the 2^24-point remesh: every species on a 513x513 grid (grid
    lines stay within the 1024 points one workgroup of the leaf search takes)."""
    import multi_species_amd as ms

    rng = np.random.default_rng(seed)
    decks = []
    for _ in range(n):
        amp = float(np.exp(rng.uniform(np.log(0.01), np.log(0.5))))
        mass = float(rng.choice([25.0, 100.0, 1836.0]))
        sim = ms.sim_params(x_min=0.0, x_max=L0, bcs=0, quadrature=0, greens_epsilon=0.05,
                            use_treecode=0, dt=0.1, remesh_period=1)
        e = ms.species_params(q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, initial_height=8,
                              p_height=0, ics_type=1, normalized_wavenumber=1.0, amp=amp,
                              pth=1.0)
        i = ms.species_params(q=1.0, m=mass, p_min=-6.0, p_max=6.0, initial_height=8,
                              p_height=0, ics_type=1, normalized_wavenumber=1.0, amp=0.0,
                              pth=1.0 / np.sqrt(mass))
        decks.append((sim, [e, i]))
    return decks


def bench_remesh(args):
    """The remesh at 2^24 target points (32 decks, 64 species, one launch per
    kernel): the remesh kernels' time by HIP events per step, against the
    algorithmic bytes of one remesh of every species."""
    import multi_species_amd as ms

    ctx = ms.Context(0)
    decks = remesh24_decks()
    b = ms.Simulation.batch(ctx, decks, ms.PRECISION_MIXED)
    npts = sum(b.deck_info(d)[3] for d in range(b.n_decks))
    nsp = sum(b.deck_info(d)[1] for d in range(b.n_decks))
    remesh_bytes = 0
    for k in range(nsp):
        n_k, p_k = b.counts(k)
        remesh_bytes += 24 * n_k + 24 * n_k + 48 * p_k
    for _ in range(args.warmup):
        b.step()
    b.set_timing(True)
    rms, lat = [], []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        b.step()
        lat.append(time.perf_counter() - t1)
        rms.append(b.remesh_ms())
    b.set_timing(False)
    b.close()
    ctx.close()
    remesh_ms = float(np.median(rms))
    gbps = remesh_bytes / (remesh_ms * 1e-3) / 1e9
    traffic = read_profile_value(REMESH24_TRAFFIC_FILE, "remesh_counted_traffic_bytes_per_step")
    result = {
        "metric": "remesh algorithmic GB/s", "value": gbps, "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": remesh_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": REMESH24_WORKLOAD, "points": int(npts), "species": int(nsp)},
        "remesh_algorithmic_bytes_per_step": remesh_bytes,
        "roofline": {"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": gbps / HBM_PEAK_GBPS, "traffic": traffic,
                     "source": "HIP events around the remesh kernels of each timed step; "
                               "traffic from " + os.path.relpath(REMESH24_TRAFFIC_FILE, ROOT)},
        "remesh_ms": {"min": float(np.min(rms)), "median": remesh_ms, "max": float(np.max(rms))},
        "step_ms_with_event_timing": float(np.median(lat) * 1e3),
    }
    print(json.dumps(result))


def bench_c4(args):
    """The c4 block as its own bench line."""
    import multi_species_amd as ms

    ctx = ms.Context(0)
    blk = c4_block(ctx, args.decks, args.steps, args.warmup, not args.no_compare)
    ctx.close()
    result = {
        "metric": "batched RK4 point-steps/s", "value": blk["value"], "unit": "point-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": blk["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": f"{args.decks} " + C4_WORKLOAD, "decks": args.decks,
                   "points": blk["points"]},
    }
    result.update({k: v for k, v in blk.items() if k not in ("value", "unit", "steps",
                                                             "ms_per_step", "workload",
                                                             "points", "decks")})
    print(json.dumps(result))


def c5_block(ctx, requests: int, req_steps: int, plain_steps: int = 3):
    """Serving: requests arrive as a Poisson process and are stepped with continuous
    batching (mixed precision, error governed at 4e-9); per-step latency percentiles
    and aggregate rate; every request's final E against the FP64 field of the same
    points; the same requests one by one in plain FP64."""
    import multi_species_amd as ms

    reqs = c5_requests(requests)
    srv = ms.Server(ctx, hbm_budget_bytes=200e9, max_active=0, precision=ms.PRECISION_MIXED)
    ids = []
    t0 = time.perf_counter()
    for t_arr, _, deck in reqs:
        delay = t_arr - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(*deck, req_steps))
    for rid in ids:
        srv.wait(rid, 1200)
    wall = time.perf_counter() - t0
    st = srv.stats()
    finals = {rid: [srv.result(rid, k) for k in range(2)] for rid in ids}
    srv.close()
    errs = []
    for rid, (_, _, deck) in zip(ids, reqs):
        x = np.concatenate([finals[rid][k][0] for k in range(2)])
        w = np.concatenate([finals[rid][k][3] for k in range(2)])
        es = np.concatenate([finals[rid][k][4] for k in range(2)])
        sim_p = deck[0]
        ef = ctx.emq_periodic_f64(x, x, w, sim_p.x_max - sim_p.x_min, sim_p.greens_epsilon)
        errs.append(float(np.max(np.abs(es - ef)) / np.max(np.abs(ef))))
    sizes = [int(sum(finals[r][k][0].size for k in range(2))) for r in ids]
    plain_pts, plain_s = 0, 0.0
    for _, _, deck in reqs:
        s = ms.Simulation.batch(ctx, [deck], ms.PRECISION_F64)
        n = sum(s.counts(k)[0] for k in range(len(deck[1])))
        s.step()
        t1 = time.perf_counter()
        s.step(plain_steps)
        plain_s += (time.perf_counter() - t1) / plain_steps
        plain_pts += n
        s.close()
    plain = plain_pts / plain_s
    return {
        "value": st["point_steps_per_s"], "unit": "point-steps/s",
        "latency_ms": {"p50": st["step_p50_ms"], "p99": st["step_p99_ms"]},
        "served": {k: st[k] for k in ("submitted", "completed", "failed", "steps",
                                      "steps_replayed", "point_steps", "busy_s", "span_s")},
        "graph_replays": ("steps in which no resident deck remeshes adaptively replay a captured "
                          "graph; every request here remeshes adaptively at every step (its "
                          "tree, and with it the arrays' sizes, changes), so those steps run "
                          "eagerly"),
        "wall_s": wall, "requests": requests, "steps_per_request": req_steps,
        "points_per_request": {"min": min(sizes), "max": max(sizes)},
        "kinds": sorted({k for _, k, _ in reqs}),
        "rel_e_err_vs_f64": {"max_over_requests": max(errs), "bound": 1e-8},
        "plain_f64_one_by_one": {"value": plain, "unit": "point-steps/s",
                                 "speedup_of_served_mixed": st["point_steps_per_s"] / plain,
                                 "sample": f"{plain_steps} timed steps of each request's deck"},
        "workload": (f"{requests} requests (Poisson arrivals at 20/s) of seeded two-species "
                     "decks (weak/strong Landau or bump-on-tail), 5k..500k base points "
                     "log-uniform, adaptive refinement up to 2 levels, "
                     f"{req_steps} steps each at dt=0.1, HBM budget 200 GB, continuous "
                     "batching, mixed precision"),
    }


def bench_c5(args):
    """The c5 block as its own bench line."""
    import multi_species_amd as ms

    ctx = ms.Context(0)
    blk = c5_block(ctx, args.requests, args.req_steps)
    ctx.close()
    result = {
        "metric": "served RK4 point-steps/s", "value": blk["value"], "unit": "point-steps/s",
        "n_gpus": 1, "steps": int(blk["served"]["steps"]), "warmup": 0,
        "ms_per_step": blk["latency_ms"]["p50"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": blk["workload"], "requests": args.requests},
    }
    result.update({k: v for k, v in blk.items() if k not in ("value", "unit", "workload")})
    print(json.dumps(result))


if __name__ == "__main__":
    main()
