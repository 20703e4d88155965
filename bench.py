#!/usr/bin/env python3
"""Benchmark of the size-distribution ensemble path on one MI355X.

``python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload spheres|lognormals]``

A step is one pass of the hot path over one batch of synthetic input already resident in
HBM. The default workload is 2^20 independent spheres (x log-uniform on [0.1, 1000],
Re m uniform on [1.30, 1.70], Im m log-uniform on [1e-8, 1e-1]); the metric counts
(particle, order) Mie terms per second with N = ceil(x + 6 x^(1/3) + 3) terms per sphere.
Prints one JSON line with the roofline of the dominant kernel (timed with HIP events on the
library's stream) and a CPU baseline (the FP64 oracle on the host's cores).
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

METRIC = "Mie (particle·order) terms/s + ensembles/s, fused vs plain FP64 GPU at 1e-6 rel"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of mie_batch on this workload: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 from
# two rocprofv3 --pmc runs, scaled by what the counters read for a 1 GiB device copy
# (profiles/mie_batch_spheres_2p20.md).
MIE_BATCH_TRAFFIC_BYTES = 1.207e9


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_threads() -> int:
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(16, n or 1))


def bench_spheres(args, torch):
    from scatrs_amd import _native, workloads
    from scatrs_amd.mie import efficiencies_device

    x, m = workloads.spheres(1 << 20)
    terms, flops = workloads.mie_cost(x, m)
    dev = torch.device("cuda", 0)
    xt = torch.from_numpy(x).to(dev)
    mt = torch.from_numpy(m).to(dev)
    ctx = _native.context(0)
    log(f"spheres: {x.size}, terms {terms}, algorithmic flops {flops:.4e}")
    for _ in range(args.warmup):
        efficiencies_device(xt, mt)
    torch.cuda.synchronize()
    ctx.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        values, _, got_terms = efficiencies_device(xt, mt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ktimes = ctx.kernel_times()
    ctx.set_profiling(False)
    assert got_terms == terms, (got_terms, terms)
    elapsed = t1 - t0
    ms_step = 1e3 * elapsed / args.steps
    value = terms * args.steps / elapsed
    k_ms = float(np.mean(ktimes)) if ktimes else float("nan")
    achieved = flops / (k_ms * 1e-3) / 1e12
    log(f"step {ms_step:.3f} ms, kernel mie_batch {k_ms:.3f} ms avg over {len(ktimes)}")
    # a parity spot check of the timed outputs against the oracle (not timed)
    check = None
    if not args.no_check:
        import oracle

        idx = np.random.default_rng(1).choice(x.size, 2048, replace=False)
        ref, _ = oracle.mie_efficiencies_batch(x[idx], m[idx], threads=cpu_threads())
        got = values.cpu().numpy()[:, idx]
        rel = np.abs(got[[0, 1, 4]] - ref[[0, 1, 4]]) / np.abs(ref[[0, 1, 4]])
        check = float(rel.max())
        log(f"max rel err vs oracle (Q_ext, Q_sca, g) on 2048 samples: {check:.3e}")
    cpu = None
    if not args.no_cpu_baseline:
        import oracle

        th = cpu_threads()
        oracle.mie_efficiencies_batch(x[:4096], m[:4096], threads=th)  # warm
        c0 = time.perf_counter()
        oracle.mie_efficiencies_batch(x, m, threads=th)
        c1 = time.perf_counter()
        cpu = {"value": terms / (c1 - c0), "unit": "terms/s", "cores": th, "kind": "port",
               "sample": f"all 2^20 spheres of the workload, one pass ({c1 - c0:.2f} s)"}
        log(f"cpu baseline: {cpu['value']:.4e} terms/s on {th} threads")
    roof = {"bound": "mfma", "pipe": "fp64 (vector; MI355X's FP64 vector and matrix peaks "
            "are both 78.6 TFLOP/s)", "kernel": "mie_batch",
            "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": args.traffic,
            "kernel_ms": k_ms, "flops_per_launch": flops}
    return {
        "metric": METRIC, "value": value, "unit": "terms/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2^20 batched single spheres, x log-uniform [0.1, 1000], "
                   "Re m U[1.30, 1.70], Im m log-uniform [1e-8, 1e-1]; outputs Q_ext, Q_sca, g "
                   "(all five efficiencies computed)", "spheres": 1 << 20, "terms": terms,
                   "seed": workloads.SEED},
        "roofline": roof, "cpu_baseline": cpu, "max_rel_err_vs_oracle": check,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="spheres", choices=["spheres"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--traffic", type=float, default=MIE_BATCH_TRAFFIC_BYTES,
                    help="HBM bytes per launch of the dominant kernel, from rocprofv3 --pmc")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="first copy this many MiB device-to-device (a PMC byte-count reference)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build's path runs on a single MI355X and no "
              "figure at several GPUs is produced.",
              file=sys.stderr)
        return 2
    import torch

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 1
    if args.copy_calibration:
        n = args.copy_calibration << 20
        a = torch.empty(n // 4, dtype=torch.float32, device="cuda")
        a.fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calibration} MiB x 3")
    res = bench_spheres(args, torch)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
