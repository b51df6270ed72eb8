"""Benchmark: mGP iterations/s of the ePlace loop on one MI355X.

Workload: the seeded adaptec1-shaped synthetic design (211k movable cells,
543 fixed macros, 221k nets, ~0.94M pins with a power-law degree tail to 500,
die 11,600 with rows of 12, 57% utilisation, ~159k fillers), 512x512 bins,
mGP from a clustered start.  A step is one mGP iteration (one Nesterov step
with its backtracking evaluations) of the device-resident engine; `value` is
FP32-path iterations/s over K timed iterations after W warm-up iterations.
The FP64 path is timed the same way and reported beside it.

Also reported: the roofline of the dominant kernel family, timed with HIP
events on the engine's stream over the timed region, and a CPU baseline (the
oracle's FP32 restatement of the reference CPU path, OpenMP on this host's
cores) on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-fp64] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GP iters/s @211k cells, 512² bins (FP32 vs FP64 path); jobs/h + p99 latency"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK = {32: 157.3, 64: 78.6}   # dense matrix TFLOP/s (FP32, FP64)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_run(design, prec, steps, warmup, profile_mask=0, max_iter=None):
    from superplace_amd import hip
    cfg = hip.GpConfig(precision=prec, gp_bin_dim=512, gp_max_iter=max_iter or (warmup + steps))
    t0 = time.time()
    eng = hip.GpEngine(design, cfg)
    setup = time.time() - t0
    if profile_mask:
        eng.set_profiling(profile_mask)
    res, rows = eng.run()
    n = len(rows)
    if n <= warmup:
        raise RuntimeError(f"stage ended after {n} iterations, before the warm-up finished")
    k = n - warmup
    t_ms = rows[-1]["elapsed_ms"] - rows[warmup - 1]["elapsed_ms"] if warmup > 0 else rows[-1]["elapsed_ms"]
    evals = rows[-1]["evaluations"] - (rows[warmup - 1]["evaluations"] if warmup > 0 else 0)
    stats = eng.kernel_stats()
    eng.close()
    return dict(iters=k, ms=t_ms, iters_per_s=1000.0 * k / t_ms, evals=evals, setup_s=setup,
                result=res, last=rows[-1], stats=stats)


def measured_traffic(family):
    """HBM bytes per launch of a kernel family from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/hbm_traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
        return table["families"][family]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(design, iters):
    """The oracle's restatement of the reference CPU path (FP32 arithmetic with
    FP64 fixed-chunk reductions, 8 density chunks, FFT-based DCTs in place of
    FFTW, OpenMP over the host's cores) on the first `iters` iterations of the
    same workload."""
    import oracle
    threads = os.cpu_count() or 1
    cfg = oracle.default_config(precision=32, gp_bin_dim=512, gp_max_iter=iters)
    oracle.set_fast_poisson(True)
    t0 = time.time()
    try:
        res, rows, _, _ = oracle.eplace_run(design, cfg, threads=threads)
    finally:
        oracle.set_fast_poisson(False)
    wall = time.time() - t0
    return dict(value=res.iterations / res.seconds_loop, unit="iters/s",
                cores=oracle.get_threads(), kind="port",
                sample=(f"first {res.iterations} mGP iterations of the same workload and start "
                        f"(FP32 with FP64 fixed-chunk reductions, 8 density chunks, FFT DCTs, "
                        f"OpenMP); loop {res.seconds_loop:.1f} s, wall incl. setup {wall:.1f} s"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(instances shard across GPUs without communication); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from superplace_amd import design as D
    from superplace_amd import hip

    t0 = time.time()
    d = D.config_adaptec1()
    log(f"design: {d.num_movable} cells, {d.num_nodes - d.num_movable} fixed, {d.num_fillers} "
        f"fillers, {d.num_pins} pins, {d.num_nets} nets ({time.time() - t0:.1f} s)")

    # 1) FP32 throughput: the graph-launched device loop, nothing else timed
    r32 = timed_run(d, 32, args.steps, args.warmup)
    log(f"FP32: {r32['iters']} timed iterations in {r32['ms']:.1f} ms -> "
        f"{r32['iters_per_s']:.1f} iters/s ({r32['evals']} evaluations); tau "
        f"{r32['last']['overflow']:.4f} HPWL {r32['last']['hpwl']:.6g}")

    # 2) all kernel families timed (slots launched directly, every kernel of a
    #    family carrying its HIP event pair) to find the dominant one
    probe = timed_run(d, 32, args.steps, args.warmup,
                      profile_mask=(1 << len(hip.KERNEL_FAMILIES)) - 1)
    fam_us = {f: 1e3 * st["ms"] / max(st["launches"], 1) for f, st in probe["stats"].items()}
    dominant = max(fam_us, key=fam_us.get)
    log("kernel families, average us per launch: " +
        ", ".join(f"{f} {v:.1f}" for f, v in fam_us.items()) + f"; dominant {dominant}")

    # 3) roofline of the dominant family over the same timed region, timed alone
    dom_bit = 1 << hip.KERNEL_FAMILIES.index(dominant)
    rr = timed_run(d, 32, args.steps, args.warmup, profile_mask=dom_bit)
    st = rr["stats"][dominant]
    avg_ms = st["ms"] / max(st["launches"], 1)
    mfma = st["flops"] > 0   # dense-transform Poisson; every other kernel is memory-bound
    if mfma:
        achieved = st["flops"] / max(st["launches"], 1) / (avg_ms * 1e-3) / 1e12
        roof = dict(bound="mfma", achieved=achieved, peak=MFMA_PEAK[32], unit="TFLOP/s",
                    frac=achieved / MFMA_PEAK[32], traffic=None)
    else:
        achieved = st["bytes"] / max(st["launches"], 1) / (avg_ms * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=achieved / HBM_PEAK_GBS, traffic=None)
    roof["traffic"] = measured_traffic(dominant)
    roof.update(kernel=dominant, avg_launch_us=avg_ms * 1e3, launches=st["launches"],
                algorithmic_per_launch=(st["flops"] if mfma else st["bytes"])
                / max(st["launches"], 1))
    log(f"roofline {dominant}: {achieved:.1f} {roof['unit']} over {st['launches']} launches, "
        f"{avg_ms * 1e3:.1f} us each")

    # 3) FP64 path
    r64 = None
    if not args.no_fp64:
        r64 = timed_run(d, 64, args.steps, args.warmup)
        log(f"FP64: {r64['iters_per_s']:.1f} iters/s; tau {r64['last']['overflow']:.4f} "
            f"HPWL {r64['last']['hpwl']:.6g}")

    # 4) CPU baseline (rank 0, N=1 only), bounded sample
    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline: {args.cpu_iters} oracle FP32 iterations on {os.cpu_count()} cores ...")
        cpu = cpu_baseline(d, args.cpu_iters)
        log(f"CPU baseline: {cpu['value']:.3f} iters/s")

    out = {
        "metric": METRIC,
        "value": r32["iters_per_s"],
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": r32["iters"],
        "warmup": args.warmup,
        "ms_per_step": r32["ms"] / r32["iters"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "synthetic adaptec1-shaped mGP, seed 1002",
                   "cells": d.num_movable, "fixed_macros": d.num_nodes - d.num_movable,
                   "fillers": d.num_fillers, "nets": d.num_nets, "pins": d.num_pins,
                   "bins": 512, "die": 11600, "utilisation": 0.57},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp64_iters_per_s": r64["iters_per_s"] if r64 else None,
        "fp32_over_fp64": (r32["iters_per_s"] / r64["iters_per_s"]) if r64 else None,
        "evals_per_iter": r32["evals"] / r32["iters"],
        "kernel_us_per_launch": {f: round(v, 2) for f, v in fam_us.items()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
