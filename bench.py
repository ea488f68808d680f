#!/usr/bin/env python3
"""Benchmark: periodic regularized E-field evaluation at N = 2^20 (FP64).

Workload (the field-evaluation-alone-at-scale deck): N = 2^20 targets =
sources, x uniform in [0, 4 pi) plus a 0.05 sin perturbation, weights
q f(v) dA with f Maxwellian in v ~ U[-6, 6], dA = 1e-4, half electron- and
half ion-signed, eps = 0.05. One step = one full direct-sum evaluation
(N^2 pair interactions) through fs_emq_periodic_f64_dev on device-resident
inputs. Prints one JSON line.

The roofline block uses HIP events (torch.cuda.Event on the stream the
library launches on) around the timed launches; the kernel is VALU-bound
(no MFMA: not a contraction), so it is priced at 12 FLOP per pair against
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

L0 = 4 * np.pi
FLOP_PER_PAIR = 12
FP64_PEAK_TFLOPS = 78.6
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "direct_f64_n1048576_traffic.txt")


def c2_inputs(n: int, seed: int = 0):
    rng = np.random.default_rng(seed)
    u = rng.uniform(0, L0, n)
    x = np.mod(u + 0.05 * np.sin(2 * np.pi / L0 * u), L0)
    v = rng.uniform(-6, 6, n)
    q = np.where(np.arange(n) < n // 2, -1.0, 1.0)
    w = q * np.exp(-v * v / 2) / np.sqrt(2 * np.pi) * 1e-4
    return x, w


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
    achieved_tflops = FLOP_PER_PAIR * pairs / (kernel_ms * 1e-3) / 1e12
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
            "flop_per_pair": FLOP_PER_PAIR,
        },
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(x, w, args.eps)
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
