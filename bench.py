#!/usr/bin/env python3
"""Benchmark of the size-distribution ensemble path on one MI355X.

``python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload spheres|lognormals]``

A step is one pass of the hot path over one batch of synthetic input already resident in
HBM. The headline workload is 2^20 independent spheres (x log-uniform on [0.1, 1000],
Re m uniform on [1.30, 1.70], Im m log-uniform on [1e-8, 1e-1]); `value` counts
(particle, order) Mie terms per second with N = ceil(x + 6 x^(1/3) + 3) terms per sphere.
The same line carries the ensembles/s of the ensemble workloads (4096 lognormals x 64
wavelengths with the matrix-core and the plain FP64 angular kernels; the adjoint workload; the
layered workload), each with its CPU baseline (the FP64 oracle on the host's cores), and the
rooflines of the dominant kernels (timed with HIP events on the library's stream).
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
FP16_PEAK_TFLOPS = 2500.0  # dense
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of mie_batch on this workload: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 from
# two rocprofv3 --pmc runs, scaled by what the counters read for a 1 GiB device copy
# (profiles/mie_batch_spheres_2p20.md).
MIE_BATCH_TRAFFIC_BYTES = 1.602e9
# mie_batch on the same workload under rocprofv3 --kernel-trace --stats
# (profiles/mie_batch_spheres_2p20.md): the average duration of its launches.
MIE_BATCH_ROCPROF = {"kernel_ms": 0.9629, "source": "profiles/mie_batch_spheres_2p20.md"}


# The ensemble legs after the sphere batch, in the order they run.
LEGS = ("c1", "ensembles", "adjoint", "layered", "distributions", "adaptive")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_threads() -> int:
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(16, n or 1))


def host_info() -> dict:
    """The host the CPU baselines ran on: logical CPUs, those this process may use, the model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"nproc": os.cpu_count(), "usable": usable, "cpu_model": model}


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
    tuned = _tune(args, torch, ctx, lambda: efficiencies_device(xt, mt), ["sphere_waves"])
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
        # one thread on every 16th sphere (the same mix of sizes)
        sub = slice(0, None, 16)
        sub_terms, _ = workloads.mie_cost(x[sub], m[sub])
        c2 = time.perf_counter()
        oracle.mie_efficiencies_batch(x[sub], m[sub], threads=1)
        c3 = time.perf_counter()
        cpu = {"value": terms / (c1 - c0), "unit": "terms/s", "cores": th, "kind": "port",
               "sample": f"all 2^20 spheres of the workload, one pass ({c1 - c0:.2f} s); one "
                         f"thread: every 16th sphere ({c3 - c2:.2f} s)",
               "one_thread": sub_terms / (c3 - c2), "host": host_info()}
        log(f"cpu baseline: {cpu['value']:.4e} terms/s on {th} threads, "
            f"{cpu['one_thread']:.4e} on 1")
    roof = {"bound": "fp64-valu", "pipe": "FP64 vector ALU (MI355X's FP64 vector peak, "
            "78.6 TFLOP/s)", "kernel": "mie_batch",
            "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": args.traffic,
            "kernel_ms": k_ms, "flops_per_launch": flops,
            "rocprof": dict(MIE_BATCH_ROCPROF,
                            frac=flops / (MIE_BATCH_ROCPROF["kernel_ms"] * 1e-3) / 1e12
                            / FP64_PEAK_TFLOPS)}
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
        "tuning": tuned,
    }


def _tune(args, torch, ctx, run, params, repeats=5):
    """The project's autotuner (scatrs_amd.tune) on this workload: the chosen launch parameters
    stay applied on the context; returns them with the candidates' times."""
    from scatrs_amd import tune

    if args.no_tune:
        return None
    res = tune.tune(run, lambda name, v: ctx.set_tuning(tune.KEYS[name], v), params,
                    repeats=repeats, sync=torch.cuda.synchronize)
    log("tuned " + ", ".join(f"{k} = {v}" for k, v in res["best"].items()) + " (" +
        "; ".join(f"{k}: " + ", ".join(f"{c} {1e3 * t:.3f} ms" for c, t in row.items())
                  for k, row in res["timings"].items()) + ")")
    return {"best": res["best"],
            "ms": {k: {str(c): 1e3 * t for c, t in row.items()} for k, row in res["timings"].items()}}


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _timed(torch, ctx, fn, steps):
    """Mean wall seconds per call of fn over `steps` calls (after a synchronize) and the library's
    marked kernel regions over them."""
    torch.cuda.synchronize()
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    return (t1 - t0) / steps, kt, out


def bench_c1(args, torch):
    """The reference's CPU-runnable case: a sulfate lognormal (r_g = 0.08 um, sigma_g = 1.6,
    m = 1.43 + 1e-8 i) at 0.35, 0.55 and 1.0 um, 128 radius nodes, Greek to 32 on a 64-cosine
    rule (exact for its N <= 33). Latency-bound: ensembles/s of back-to-back eager calls and of
    replays of the call captured as a HIP graph, and the oracle on the host."""
    from scatrs_amd import _native, workloads
    from scatrs_amd.ensemble import EnsembleGraph, ensemble_fixed_device

    kind, p0, p1, wl, m = workloads.sulfate()
    t = [_dev(torch, v) for v in (kind, p0, p1, wl, m)]
    kw = dict(n_radius=128, n_angles=64, order=32, want_phase=False)
    n = args.c1_repeats
    ctx = _native.context(0)
    tuned = _tune(args, torch, ctx, lambda: ensemble_fixed_device(*t, **kw), ["split_blocks"],
                  repeats=20)
    for _ in range(5):
        sc, gr, _, _ = ensemble_fixed_device(*t, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        sc, gr, _, _ = ensemble_fixed_device(*t, **kw)
    torch.cuda.synchronize()
    eager = (time.perf_counter() - t0) / n
    g = EnsembleGraph(*t, **kw)
    for _ in range(5):
        g.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        g.launch()
    torch.cuda.synchronize()
    replay = (time.perf_counter() - t0) / n
    g.status()
    same = bool(torch.equal(g.scalars, sc) and torch.equal(g.greek, gr))
    g.close()
    log(f"C1: eager {1e6 * eager:.1f} us/call, graph replay {1e6 * replay:.1f} us/call "
        f"({3 / replay:.0f} ensembles/s), replay equals eager: {same}")
    res = {"workload": "sulfate lognormal r_g 0.08 um, sigma_g 1.6, m 1.43+1e-8i at 0.35, 0.55, "
                       "1.0 um; 128 radius nodes, Greek to 32 (64-cosine rule)",
           "ensembles": 3, "repeats": n, "eager_us_per_call": 1e6 * eager,
           "graph_us_per_call": 1e6 * replay, "ensembles_per_s": 3 / replay,
           "eager_ensembles_per_s": 3 / eager, "graph_equals_eager": same, "tuning": tuned}
    if not args.no_check:
        import oracle

        sch, grh = sc.cpu().numpy(), gr.cpu().numpy()
        es, eg = 0.0, 0.0
        for e in range(3):
            ref = oracle.ensemble_fixed(0, p0[e], p1[e], wl[e], m[e], n_radius=128, n_angles=64,
                                        order=32)
            k = [0, 1, 4, 5]
            es = max(es, float(np.max(np.abs(sch[e][k] - ref["scalars"][k]) /
                                      np.abs(ref["scalars"][k]))))
            eg = max(eg, float(np.max(np.abs(grh[e] - ref["greek"])) /
                               np.max(np.abs(ref["greek"][0]))))
        res["max_rel_err_vs_oracle"] = {"c_ext_c_sca_ssa_g": es, "greek": eg}
    if not args.no_cpu_baseline:
        import oracle

        base = {}
        for th in sorted({1, cpu_threads()}):
            reps, c0 = 0, time.perf_counter()
            while time.perf_counter() - c0 < 2.0:
                oracle.ensemble_batch(kind, p0, p1, wl, m, 128, 64, 32, threads=th)
                reps += 1
            base[th] = 3 * reps / (time.perf_counter() - c0)
        th = max(base)
        res["cpu_baseline"] = {"value": base[th], "unit": "ensembles/s", "cores": th,
                               "kind": "port", "one_thread": base[1], "host": host_info(),
                               "sample": "the 3 ensembles, repeated for 2 s per thread count"}
        log(f"C1 cpu baseline: {base[th]:.1f} ensembles/s on {th} threads, {base[1]:.1f} on 1")
    return res


def bench_ensembles(args, torch):
    """4096 lognormals x 64 wavelengths, 512 radius nodes, 1024 angles, Greek to 256 and the
    phase matrix: ensembles/s with the FP64 matrix-core angular kernel (fused) and with the
    plain FP64 vector kernel, the angular kernel's roofline, and the oracle on a sample."""
    from scatrs_amd import _native, workloads
    from scatrs_amd.ensemble import ensemble_fixed_device

    kind, p0, p1, wl, m = workloads.lognormals()
    if args.ens_dists < 4096:
        keep = args.ens_dists * 64
        kind, p0, p1, wl, m = kind[:keep], p0[:keep], p1[:keep], wl[:keep], m[:keep]
    nr, na, order = 512, 1024, 256
    h = (na + 1) // 2
    ctx = _native.context(0)
    t = [_dev(torch, v) for v in (kind, p0, p1, wl, m)]
    fn = lambda: ensemble_fixed_device(*t, n_radius=nr, n_angles=na, order=order)  # noqa: E731
    log(f"ensembles: {kind.size} (4096 lognormals x 64 wavelengths), warm-up")
    # The fused path: fp16-split matrix-core angular products (kernel 2).
    ctx.set_angular_kernel(2)
    try:
        tuned = _tune(args, torch, ctx, fn, ["chunk_share", "node_waves"], repeats=1)
        fn()
        sec, kt, out = _timed(torch, ctx, fn, args.ens_steps)
    finally:
        ctx.set_angular_kernel(0)
    terms = out[3]
    node_ms = float(np.sum(kt[0::2])) / args.ens_steps
    ang_ms = float(np.sum(kt[1::2])) / args.ens_steps
    flops = 16.0 * h * terms  # two real GEMMs x 2 tables x 4 columns per (node, order, cosine)
    log(f"fused (fp16-split angular): {sec:.3f} s/step, nodes {node_ms:.0f} ms, angular "
        f"{ang_ms:.0f} ms")
    res = {"workload": "4096 lognormals (r_g log-uniform [0.01, 5] um, sigma_g U[1.2, 2.5], m as "
                       "the sphere batch) x 64 wavelengths log-spaced 0.3-2.5 um; 512 radius nodes, "
                       "1024 angles, Greek to 256, F11..F44",
           "ensembles": int(kind.size), "terms": int(terms), "tuning": tuned,
           "ensembles_per_s": kind.size / sec, "s_per_step": sec, "steps": args.ens_steps,
           "node_ms_per_step": node_ms, "angular_ms_per_step": ang_ms}
    ach = flops / (ang_ms * 1e-3) / 1e12
    # Three fp16 products per FP64-equivalent product: the fp16 dense peak over 3.
    peak16 = FP16_PEAK_TFLOPS / 3.0
    res["roofline_angular"] = {"bound": "mfma",
                               "pipe": "fp16 matrix cores (v_mfma_f32_16x16x32_f16), 3 products "
                                       "per split product; peak = 2500/3 TFLOP/s",
                               "kernel": "ens_angular_f16", "achieved": ach, "peak": peak16,
                               "unit": "TFLOP/s", "frac": ach / peak16, "flops_per_step": flops}
    ctx.set_angular_kernel(0)
    sec64, kt64, out64 = _timed(torch, ctx, fn, 1)
    ang64 = float(np.sum(kt64[1::2]))
    res["fp64_matrix_cores"] = {"ensembles_per_s": kind.size / sec64, "s_per_step": sec64,
                                "angular_ms_per_step": ang64,
                                "angular_tflops": flops / (ang64 * 1e-3) / 1e12,
                                "angular_frac_of_fp64_peak": flops / (ang64 * 1e-3) / 1e12 /
                                FP64_PEAK_TFLOPS}
    res["speedup_vs_fp64_matrix_cores"] = sec64 / sec
    log(f"FP64 matrix cores: {sec64:.3f} s/step, angular {ang64:.0f} ms; fused is "
        f"{sec64 / sec:.2f}x")
    # alpha1 of every ensemble, fp16-split against FP64 matrix cores (on the device)
    d_all = float(((out[1][:, 0] - out64[1][:, 0]).abs().amax(dim=1) /
                   out64[1][:, 0].abs().amax(dim=1)).max().item())
    del out64
    log(f"alpha1, fp16-split vs FP64 matrix cores over all {kind.size} ensembles: {d_all:.2e}")
    if not args.no_plain:
        ctx.set_angular_kernel(1)
        try:
            sec_p, kt_p, _ = _timed(torch, ctx, fn, 1)
        finally:
            ctx.set_angular_kernel(0)
        ang_p = float(np.sum(kt_p[1::2]))
        res["plain_fp64"] = {"ensembles_per_s": kind.size / sec_p, "s_per_step": sec_p,
                             "angular_ms_per_step": ang_p,
                             "angular_tflops": flops / (ang_p * 1e-3) / 1e12,
                             "angular_frac_of_fp64_peak": flops / (ang_p * 1e-3) / 1e12 /
                             FP64_PEAK_TFLOPS}
        res["speedup_vs_plain"] = sec_p / sec
        log(f"plain FP64: {sec_p:.3f} s/step, fused speedup {sec_p / sec:.2f}x")
    if not args.no_check:
        import oracle

        idx = np.random.default_rng(2).choice(kind.size, 4, replace=False)
        ti = torch.as_tensor(idx, device=out[0].device)
        sc, gr = out[0][ti].cpu().numpy(), out[1][ti].cpu().numpy()
        err_s, err_g = 0.0, 0.0
        for r, e in enumerate(idx):
            ref = oracle.ensemble_fixed(0, p0[e], p1[e], wl[e], m[e], n_radius=nr, n_angles=na,
                                        order=order)
            rs = ref["scalars"]
            k = [0, 1, 4, 5]  # c_ext, c_sca, ssa, g
            err_s = max(err_s, float(np.max(np.abs(sc[r][k] - rs[k]) / np.abs(rs[k]))))
            err_g = max(err_g, float(np.max(np.abs(gr[r][0] - ref["greek"][0])) /
                                     np.max(np.abs(ref["greek"][0]))))
        res["max_rel_err_vs_oracle"] = {"c_ext_c_sca_ssa_g": err_s, "alpha1": err_g,
                                        "sample": "4 ensembles",
                                        "alpha1_fp16_split_vs_fp64_all_ensembles": d_all}
        log(f"vs oracle on 4 ensembles: scalars {err_s:.2e}, alpha1 {err_g:.2e}")
    if not args.no_cpu_baseline:
        import oracle

        th = cpu_threads()
        sel = np.arange(0, kind.size, kind.size // args.ens_cpu_sample)[:args.ens_cpu_sample]
        c0 = time.perf_counter()
        oracle.ensemble_batch(kind[sel], p0[sel], p1[sel], wl[sel], m[sel], nr, na, order,
                              threads=th)
        c1 = time.perf_counter()
        one = sel[::4]
        c2 = time.perf_counter()
        oracle.ensemble_batch(kind[one], p0[one], p1[one], wl[one], m[one], nr, na, order,
                              threads=1)
        c3 = time.perf_counter()
        res["cpu_baseline"] = {"value": sel.size / (c1 - c0), "unit": "ensembles/s",
                               "cores": th, "kind": "port",
                               "sample": f"{sel.size} of the ensembles, evenly spaced "
                                         f"({c1 - c0:.1f} s); one thread: every 4th of them "
                                         f"({c3 - c2:.1f} s)",
                               "one_thread": one.size / (c3 - c2), "host": host_info()}
        log(f"cpu baseline: {sel.size / (c1 - c0):.2f} ensembles/s on {th} threads, "
            f"{one.size / (c3 - c2):.2f} on 1")
    return res


def bench_adjoint(args, torch):
    """The adjoint workload: 1024 of the lognormals x 32 wavelengths, forward plus the reverse
    pass of a loss on ssa, g and alpha1[0:64] (one step = forward + VJP)."""
    from scatrs_amd import _native, workloads
    from scatrs_amd.ensemble import ensemble_fixed_device, ensemble_fixed_vjp_device

    (kind, p0, p1, wl, m), tg = workloads.adjoint()
    nr, na, order = 512, 1024, 256
    ctx = _native.context(0)
    t = [_dev(torch, v) for v in (kind, p0, p1, wl, m)]
    tt = {k: _dev(torch, v) for k, v in tg.items()}

    def step():
        sc, gr, _, _ = ensemble_fixed_device(*t, n_radius=nr, n_angles=na, order=order,
                                             want_phase=False)
        _, bs, bg = workloads.adjoint_loss(sc, gr, tt)
        g, _ = ensemble_fixed_vjp_device(*t, bar_scalars=bs, bar_greek=bg, n_radius=nr,
                                         n_angles=na, order=order)
        return g

    def run(kernel):
        ctx.set_angular_kernel(kernel)
        try:
            step()
            return _timed(torch, ctx, step, args.ens_steps)
        finally:
            ctx.set_angular_kernel(0)

    log(f"adjoint: {kind.size} ensembles (1024 lognormals x 32 wavelengths), warm-up")
    # Fused: the forward tables (forward call and reverse pass) from the fp16-split kernel; the
    # reverse products in FP64 on the matrix cores.
    sec, _, g = run(2)
    log(f"adjoint (fp16-split forward tables): {sec:.3f} s per forward + VJP step")
    sec64, _, g64 = run(0)
    log(f"adjoint (FP64 matrix cores throughout): {sec64:.3f} s per forward + VJP step")
    gs = g64.abs().amax(dim=0).clamp_min(1e-300)
    gdiff = float(((g - g64).abs().amax(dim=0) / gs).max().item())
    res = {"workload": "first 1024 of the 4096 lognormals x every other wavelength (32); "
                       "512 radius nodes, 1024 angles, Greek to 256; loss sum w[(ssa-t)^2 + "
                       "(g-t)^2 + sum_{l<64} (alpha1_l-t_l)^2]; step = forward + VJP w.r.t. "
                       "r_g, sigma_g, Re m, Im m",
           "ensembles": int(kind.size), "ensembles_per_s": kind.size / sec, "s_per_step": sec,
           "angular": "fp16-split forward tables, FP64 matrix-core reverse products",
           "fp64_matrix_cores": {"ensembles_per_s": kind.size / sec64, "s_per_step": sec64},
           "speedup_vs_fp64_matrix_cores": sec64 / sec,
           "grad_max_diff_vs_fp64": gdiff,
           "finite": bool(torch.isfinite(g).all().item())}
    # The reverse products' roofline: one reverse pass alone, its marked regions per chunk being
    # the forward nodes and tables it recomputes, the reverse angular products, then the
    # reverse node solves; 32 h N flops per node (the forward S1, S2 again and the transposed
    # products) on the FP64 matrix cores.
    sc, gr, _, terms = ensemble_fixed_device(*t, n_radius=nr, n_angles=na, order=order,
                                             want_phase=False)
    _, bs, bg = workloads.adjoint_loss(sc, gr, tt)
    ctx.set_angular_kernel(2)
    try:
        _, kt_r, _ = _timed(torch, ctx, lambda: ensemble_fixed_vjp_device(
            *t, bar_scalars=bs, bar_greek=bg, n_radius=nr, n_angles=na, order=order), 1)
    finally:
        ctx.set_angular_kernel(0)
    rev_ms = float(np.sum(kt_r[2::4]))
    h = (na + 1) // 2
    ach = 32.0 * h * terms / (rev_ms * 1e-3) / 1e12
    res["roofline_reverse"] = {"bound": "mfma", "pipe": "FP64 matrix cores "
                               "(v_mfma_f64_16x16x4f64)", "kernel": "ens_angular_vjp",
                               "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                               "frac": ach / FP64_PEAK_TFLOPS, "reverse_ms_per_step": rev_ms,
                               "flops_per_step": 32.0 * h * terms}
    log(f"reverse products: {rev_ms:.0f} ms, {ach:.1f} TFLOP/s ({ach / FP64_PEAK_TFLOPS:.3f} "
        f"of FP64 peak)")
    if not args.no_cpu_baseline:
        import oracle

        n = args.adj_cpu_sample
        sel = np.arange(0, kind.size, kind.size // n)[:n]
        c0 = time.perf_counter()
        for e in sel:
            oracle.ensemble_fixed_jacobian(0, p0[e], p1[e], wl[e], m[e], None, nr, na, order)
        c1 = time.perf_counter()
        res["cpu_baseline"] = {"value": n / (c1 - c0), "unit": "ensembles/s", "cores": 1,
                               "kind": "port", "host": host_info(),
                               "sample": f"forward-mode Jacobian (all 4 parameters) of {n} evenly "
                                         f"spaced ensembles, one thread ({c1 - c0:.1f} s)"}
    return res


def bench_layered(args, torch):
    """Gamma distributions of a core and 1-4 shells (2048 x 32 wavelengths), 256 radius nodes,
    512 angles, Greek to 128: forward + VJP for every layer and radius fraction."""
    from scatrs_amd import _native, workloads
    from scatrs_amd.ensemble import ensemble_fixed_device, ensemble_fixed_vjp_device

    nr, na, order = 256, 512, 128
    ctx = _native.context(0)
    total_s, total64, total_e, total_terms = 0.0, 0.0, 0, 0
    per_l = {}
    for L, (kind, p0, p1, wl, fr, m) in sorted(workloads.layered().items()):
        t = [_dev(torch, v) for v in (kind, p0, p1, wl, m)]
        tfr = _dev(torch, fr)
        rng = np.random.default_rng(L)
        bs = _dev(torch, rng.normal(size=(kind.size, 11)))
        bg = _dev(torch, rng.normal(size=(kind.size, 6, order + 1)))

        def step():
            sc, gr, _, terms = ensemble_fixed_device(*t, tfr, n_radius=nr, n_angles=na,
                                                     order=order, want_phase=False)
            g, _ = ensemble_fixed_vjp_device(*t, tfr, bar_scalars=bs, bar_greek=bg, n_radius=nr,
                                             n_angles=na, order=order)
            return terms, sc, gr, g

        secs, outs = {}, {}
        for kernel in (2, 0):  # fp16-split forward tables, then FP64 throughout
            ctx.set_angular_kernel(kernel)
            try:
                step()
                secs[kernel], _, (terms, *outs[kernel]) = _timed(torch, ctx, step, 1)
            finally:
                ctx.set_angular_kernel(0)
        # The timed outputs: finite everywhere, and one sampled row against the oracle (forward
        # and the VJP against the oracle's Jacobian contracted with the same cotangents).
        sc, gr, g = outs[2]
        finite = all(bool(torch.isfinite(v).all().item()) for v in (sc, gr, g, *outs[0]))
        assert finite, f"non-finite layered outputs at L={L}"
        check = None
        if not args.no_check:
            import oracle

            e = int(np.random.default_rng(60 + L).integers(kind.size))
            ref = oracle.ensemble_fixed(1, p0[e], p1[e], wl[e], m[e], fractions=fr[e],
                                        n_radius=nr, n_angles=na, order=order)
            s_e, g_e = sc[e].cpu().numpy(), gr[e].cpu().numpy()
            k = [0, 1, 4, 5]
            J = oracle.ensemble_fixed_jacobian(1, p0[e], p1[e], wl[e], m[e], fr[e], nr, na,
                                               order)
            bs_e, bg_e = bs[e].cpu().numpy(), bg[e].cpu().numpy()
            terms_j = np.concatenate([J["d_scalars"] * bs_e[None, :],
                                      (J["d_greek"] * bg_e[None]).reshape(
                                          J["d_greek"].shape[0], -1)], axis=1)
            scale = np.abs(terms_j).sum(axis=1)
            gerr = np.abs(g[e].cpu().numpy() - terms_j.sum(axis=1)) / np.where(scale > 0, scale,
                                                                              1.0)
            check = {"row": e,
                     "c_ext_c_sca_ssa_g": float(np.max(np.abs(s_e[k] - ref["scalars"][k]) /
                                                       np.abs(ref["scalars"][k]))),
                     "alpha1": float(np.max(np.abs(g_e[0] - ref["greek"][0])) /
                                     np.max(np.abs(ref["greek"][0]))),
                     "vjp_vs_oracle_jacobian": float(gerr.max())}
            assert check["c_ext_c_sca_ssa_g"] < 1e-6 and check["alpha1"] < 1e-6, check
            assert check["vjp_vs_oracle_jacobian"] < 1e-5, check
            log(f"layered L={L} row {e} vs oracle: {check}")
        del outs, sc, gr, g
        sec = secs[2]
        per_l[str(L)] = {"ensembles": int(kind.size), "s": sec, "s_fp64_matrix_cores": secs[0],
                         "terms": int(terms), "finite": finite, "check_vs_oracle": check}
        total_s += sec
        total64 += secs[0]
        total_e += kind.size
        total_terms += terms
        log(f"layered L={L}: {kind.size} ensembles, {sec:.3f} s forward + VJP "
            f"({secs[0]:.3f} s in FP64 throughout)")
    return {"workload": "2048 gamma distributions (r_eff log-uniform [0.05, 3] um, v_eff U[0.05, "
                        "0.4]) of a core and 1-4 shells (fractions Dirichlet(2), Re m U[1.3, 2.0], "
                        "Im m log-uniform [1e-6, 1]) x 32 wavelengths; 256 radius nodes, 512 "
                        "angles, Greek to 128; step = forward + VJP (all layer indices and "
                        "radius fractions)",
            "ensembles": total_e, "terms": int(total_terms), "s_per_step": total_s,
            "ensembles_per_s": total_e / total_s,
            "angular": "fp16-split forward tables, FP64 matrix-core reverse products",
            "fp64_matrix_cores": {"s_per_step": total64, "ensembles_per_s": total_e / total64},
            "per_layer_count": per_l}


def bench_distributions(args, torch):
    """The other distribution kinds and mixtures: modified gamma, power-law and monodisperse
    spheres (1536 x 32 wavelengths; 256 radius nodes, 512 angles, Greek to 128), forward +
    VJP; and 1024 bimodal lognormals x 32 wavelengths, forward: both modes on the device with
    normalized outputs off, mixed by number fraction (mix_parts)."""
    from scatrs_amd import _native, workloads
    from scatrs_amd.ensemble import ensemble_fixed_device, ensemble_fixed_vjp_device, mix_parts

    nr, na, order = 256, 512, 128
    ctx = _native.context(0)
    kind, p0, p1, p2, wl, m = workloads.distributions()
    t = [_dev(torch, v) for v in (kind, p0, p1, wl, m)]
    tp2 = _dev(torch, p2)
    rng = np.random.default_rng(5)
    bs = rng.normal(size=(kind.size, 11))
    bs[:, 7] = 0.0
    bs, bg = _dev(torch, bs), _dev(torch, rng.normal(size=(kind.size, 6, order + 1)))

    def step():
        _, _, _, terms = ensemble_fixed_device(*t, p2=tp2, n_radius=nr, n_angles=na,
                                               order=order, want_phase=False)
        ensemble_fixed_vjp_device(*t, p2=tp2, bar_scalars=bs, bar_greek=bg, n_radius=nr,
                                  n_angles=na, order=order)
        return terms

    secs = {}
    for kernel in (2, 0):
        ctx.set_angular_kernel(kernel)
        try:
            step()
            secs[kernel], _, terms = _timed(torch, ctx, step, 1)
        finally:
            ctx.set_angular_kernel(0)
    log(f"distributions: {kind.size} ensembles, {secs[2]:.3f} s forward + VJP "
        f"({secs[0]:.3f} s in FP64 throughout)")
    (kk, q0, q1, w2, mm), frac = workloads.bimodal()
    tb = [_dev(torch, v) for v in (kk, q0, q1, w2, mm)]
    f = _dev(torch, frac)
    cnt = kk.size // 2

    def mstep():
        sc, gr, ph, _ = ensemble_fixed_device(*tb, n_radius=nr, n_angles=na, order=order,
                                              normalized=False)
        return mix_parts(sc.reshape(2, cnt, 11), gr.reshape(2, cnt, 6, order + 1),
                         ph.reshape(2, cnt, 6, na), f, True)

    ctx.set_angular_kernel(2)
    try:
        mstep()
        msec, _, _ = _timed(torch, ctx, mstep, 1)
    finally:
        ctx.set_angular_kernel(0)
    log(f"bimodal mixtures: {cnt} mixtures ({kk.size} parts), {msec:.3f} s forward")
    return {"workload": "modified gamma (alpha U[1, 8], gamma U[0.5, 1.5], mode radius "
                        "log-uniform [0.05, 2] um), power law (exponent U[2, 4.5], r_min "
                        "log-uniform [0.01, 0.1] um, r_max [1, 5] um) and monodisperse (radius "
                        "log-uniform [0.05, 5] um) spheres, a third each, x 32 wavelengths; 256 "
                        "radius nodes, 512 angles, Greek to 128; step = forward + VJP "
                        "(fp16-split forward tables)",
            "ensembles": int(kind.size), "terms": int(terms), "s_per_step": secs[2],
            "ensembles_per_s": kind.size / secs[2],
            "fp64_matrix_cores": {"s_per_step": secs[0], "ensembles_per_s": kind.size / secs[0]},
            "mixtures": {"workload": "1024 bimodal lognormals (fine r_g log-uniform [0.02, 0.2] "
                                     "um, coarse [0.5, 3] um, coarse number fraction "
                                     "log-uniform [1e-4, 0.1]) x 32 wavelengths, forward, "
                                     "fp16-split tables",
                         "mixtures": cnt, "s_per_step": msec, "mixtures_per_s": cnt / msec}}


def bench_adaptive(args, torch):
    """The adaptive engine (scatrs_amd.adaptive) on the reference's own published ensemble
    workloads, best of 3 calls each as the reference times them: a lognormal (r_g = 1 um,
    sigma_g = 2, m = 1.33 + 1e-4 i) at rtol 1e-6 per wavelength, without and with the Greek
    coefficients to 32 (docs/theory/ensembles.md:431-438); W1, the same lognormal at 0.3, 0.5
    and 1 um in one call, Greek to 63, rtol 3e-4; W2, 50 lognormals x 20 wavelengths, m = 1.43
    + 1e-8 i, Greek to 63, rtol 3e-4 (docs/validation/sasktran2.md:41-62). The published
    numbers are the reference's on an Apple M4 Pro, all 12 cores. The Greek coefficients are
    projected exactly, as the reference's are: each pass-2 row on N_max + (L+1)/2 cosines."""
    from scatrs_amd import _native
    from scatrs_amd.ensemble import Ensemble
    from scatrs_amd.particles import Sphere
    from scatrs_amd.size import LogNormal

    ctx = _native.context(0)

    def best(ens, wl, **kw):
        ens.compute(wl, **kw)
        times = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = ens.compute(wl, **kw)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        return min(times), r

    e = Ensemble(Sphere(1.33 + 1e-4j), LogNormal(1.0, 2.0))
    rg = np.logspace(np.log10(0.05), 0.0, 10)
    RG, SG = np.meshgrid(rg, np.array([1.2, 1.4, 1.6, 1.8, 2.0]), indexing="ij")
    e2 = Ensemble(Sphere(1.43 + 1e-8j), LogNormal(RG.reshape(-1), SG.reshape(-1)))
    cases = []
    for wl, pub_s, pub_g in ((1.0, 0.037, 0.093), (0.5, 0.11, 0.31), (0.3, 0.20, 0.90)):
        cases.append((f"lognormal lambda={wl} rtol 1e-6", e, np.array([wl]),
                      dict(rtol=1e-6, phase=False, n_angles=None), pub_s))
        cases.append((f"lognormal lambda={wl} rtol 1e-6 Greek to 32", e, np.array([wl]),
                      dict(rtol=1e-6, greek=32, phase=False, n_angles=None), pub_g))
    cases.append(("W1: lognormal at 0.3, 0.5, 1 um, Greek to 63, rtol 3e-4", e,
                  np.array([0.3, 0.5, 1.0]), dict(rtol=3e-4, greek=63, phase=False,
                                                  n_angles=None), 0.0838))
    cases.append(("W2: 50 lognormals x 20 wavelengths, Greek to 63, rtol 3e-4", e2,
                  np.linspace(0.28, 1.02, 20), dict(rtol=3e-4, greek=63, phase=False,
                                                    n_angles=None), 0.83))
    out = []
    for name, ens, wl, kw, pub in cases:
        row = {"workload": name, "published_s_m4pro_12_cores": pub}
        for kernel, key in ((0, "fp64"), (2, "fp16_split")):
            ctx.set_angular_kernel(kernel)
            try:
                sec, r = best(ens, wl, **kw)
            finally:
                ctx.set_angular_kernel(0)
            row[key] = {"s": sec, "speedup_vs_published": pub / sec}
        row["panels"] = int(np.sum(r.panels))
        row["converged"] = bool(np.all(r.converged))
        log(f"adaptive {name}: {row['fp64']['s']:.3f} s FP64 ({pub / row['fp64']['s']:.2f}x the "
            f"published {pub} s), {row['fp16_split']['s']:.3f} s fp16-split, {row['panels']} "
            f"panels, converged {row['converged']}")
        out.append(row)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="spheres", choices=["spheres"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the library's default launch parameters (no autotuning)")
    ap.add_argument("--traffic", type=float, default=MIE_BATCH_TRAFFIC_BYTES,
                    help="HBM bytes per launch of the dominant kernel, from rocprofv3 --pmc")
    ap.add_argument("--no-ensembles", action="store_true",
                    help="only the sphere batch (skip the ensemble, adjoint and layered lines)")
    ap.add_argument("--legs", default=",".join(LEGS),
                    help="comma-separated ensemble legs to run after the sphere batch, from "
                         + ", ".join(LEGS) + " (or 'none')")
    ap.add_argument("--ens-dists", type=int, default=4096,
                    help="distributions of the 4096 x 64 lognormal workload to run (a prefix; "
                         "for profiling runs)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip timing the plain FP64 angular kernel")
    ap.add_argument("--ens-steps", type=int, default=1)
    ap.add_argument("--c1-repeats", type=int, default=2000,
                    help="calls timed per mode of the sulfate case")
    ap.add_argument("--ens-cpu-sample", type=int, default=32)
    ap.add_argument("--adj-cpu-sample", type=int, default=4)
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
    legs = [] if args.no_ensembles or args.legs == "none" else args.legs.split(",")
    for leg in legs:
        if leg not in LEGS:
            print(f"unknown leg {leg!r}; the legs are {', '.join(LEGS)}", file=sys.stderr)
            return 2
    for leg in LEGS:
        if leg in legs:
            res[leg] = globals()["bench_" + leg](args, torch)
    print(json.dumps(res), flush=True)
    # Leave nothing behind: the device idle and the library's contexts released before the
    # interpreter exits (bench.py itself starts no other process or thread).
    from scatrs_amd import _native

    torch.cuda.synchronize()
    _native.release_contexts()
    return 0


if __name__ == "__main__":
    sys.exit(main())
