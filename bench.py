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

Many instances: a batch of 64 uneven synthetic designs (5k-80k cells, bins
128-512) advanced together by the library's scheduler against the same
designs run one after another, and a serving stream (Poisson arrivals,
10k-640k cells, bins 256-1024, mGP to overflow 0.10) with its p50/p99
time-to-converge and jobs/h, against the same arrivals served one at a time
on the FP64 path.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-fp64] [--no-cpu] [--no-serve]
       python bench.py --roofline-only FAMILY [--steps K] [--warmup W]   (that leg alone)
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
    sizes = dict(n=eng.n_opt, B=eng.dim * eng.dim, dim=eng.dim, P=design.num_pins,
                 E=design.num_nets, Nmov=design.num_movable,
                 Nfix=design.num_nodes - design.num_movable)
    eng.close()
    return dict(iters=k, ms=t_ms, iters_per_s=1000.0 * k / t_ms, evals=evals, setup_s=setup,
                result=res, last=rows[-1], stats=stats, sizes=sizes)


def algorithmic_work(z, fused_finalize):
    """Algorithmic bytes (and flops) per launch of each kernel family, from
    the per-kernel formulas of the hot-path measurement plan (n optimised
    nodes, B bins, P pins, E nets, FP32):
      scatter    17n + 8B            (x, y, w, h, flags in; int64 grid out)
      finalize   16B                 (int64 grid + static terms in, rho out)
      poisson    16B, 2 nx ny (4 nx + 3 ny) flops (DCT as dense products);
                 with the finalize fused into its first pass, 16B + 16B - 8B
                 (rho is neither written nor read back)
      wa_nets    16P + 4E + 8(Nmov + Nfix) + 8 Nmov
      node_grad  25n + 8B (gather) + 8P + 4 Nmov (node-pin pass)
      update     2n * 28 + 8n        (u, g, v, u_prev, g_prev in; u, v out; w, h)
      sort, reduce: no algorithmic bytes of their own (ordering and FP64
                 block sums of this implementation)."""
    n, B, P, E, nm, nf, dim = (z["n"], z["B"], z["P"], z["E"], z["Nmov"], z["Nfix"], z["dim"])
    dct_flops = 2.0 * dim * dim * (4 * dim + 3 * dim)
    return {
        "scatter": (17 * n + 8 * B, 0.0),
        "finalize": (16 * B, 0.0),
        "poisson": ((24 if fused_finalize else 16) * B, dct_flops),
        "wa_nets": (16 * P + 4 * E + 8 * (nm + nf) + 8 * nm, 0.0),
        "node_grad": (25 * n + 8 * B + 8 * P + 4 * nm, 0.0),
        "update": (64 * n, 0.0),
        "sort": (0.0, 0.0),
        "reduce": (0.0, 0.0),
    }


def iteration_roofline_us(z, evals_per_iter):
    """Roofline time of one mGP iteration: the algorithmic bytes of one
    evaluation at 8 TB/s plus the DCT flops at the FP32 matrix peak, times the
    evaluations per iteration."""
    w = algorithmic_work(z, False)
    # scatter, rho/tau, WA + HPWL, gather + node pass, Nesterov: the bytes;
    # the transforms: their flops
    byts = sum(w[f][0] for f in ("scatter", "finalize", "wa_nets", "node_grad", "update"))
    flops = w["poisson"][1]
    return evals_per_iter * (byts / (HBM_PEAK_GBS * 1e9) + flops / (MFMA_PEAK[32] * 1e12)) * 1e6


def roofline_run(d, family, steps, warmup):
    """The dominant family timed alone with HIP events over every launch of
    its kernels in one FP32 run (set-up evaluations, the loop slots and the
    delivery), the population a kernel trace of the same run counts."""
    from superplace_amd import hip
    bit = 1 << hip.KERNEL_FAMILIES.index(family)
    rr = timed_run(d, 32, steps, warmup, profile_mask=bit)
    st = rr["stats"][family]
    launches = max(st["launches"], 1)
    avg_s = st["ms"] / launches * 1e-3
    fused = family == "poisson" and rr["stats"]["finalize"]["launches"] == 0
    alg_bytes, alg_flops = algorithmic_work(rr["sizes"], fused)[family]
    model_bytes = st["bytes"] / launches   # the engine's own byte model (windows, records)
    if family == "poisson":
        achieved = alg_flops / avg_s / 1e12
        roof = dict(bound="mfma", achieved=achieved, peak=MFMA_PEAK[32], unit="TFLOP/s",
                    frac=achieved / MFMA_PEAK[32])
        algorithmic = alg_flops
    else:
        achieved = alg_bytes / avg_s / 1e9
        roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=achieved / HBM_PEAK_GBS)
        algorithmic = alg_bytes
    roof["traffic"] = measured_traffic(family)
    roof.update(kernel=family, avg_launch_us=avg_s * 1e6, launches=st["launches"],
                algorithmic_per_launch=algorithmic,
                algorithmic_bytes_per_launch=alg_bytes,
                overhead_bytes_per_launch=max(model_bytes - alg_bytes, 0.0))
    return roof


def auto_bin_dim(d, lo, hi):
    from superplace_amd import design as D
    return D.auto_bin_dim(d, lo, hi)


def job_designs(count, lo, hi, seed0, dim_lo, dim_hi):
    from superplace_amd import design as D
    return D.job_designs(count, lo, hi, seed0, dim_lo, dim_hi)


def batch_bench(count):
    """Uneven instances advanced together, against the same instances run one
    after another: by batched launches (sp_hip_gp_run_batch: every kernel of
    an iteration launched once for all the instances) and by the scheduler
    (each instance on its own stream); each job's iterations must equal its
    own run's."""
    from superplace_amd import hip, serve
    jobs = job_designs(count, 5000, 80000, 0, 128, 512)
    cfgs = [hip.GpConfig(gp_bin_dim=dim) for _, dim in jobs]
    designs = [d for d, _ in jobs]
    t0 = time.time()
    alone = []
    for d, cfg in zip(designs, cfgs):
        eng = hip.GpEngine(d, cfg)
        res, _ = eng.run(max_rows=1)
        eng.positions()
        eng.close()
        alone.append(res["iterations"])
    seq_s = time.time() - t0
    t0 = time.time()
    bres = hip.run_batch(designs, cfgs, positions=False)
    batched_s = time.time() - t0
    srv = serve.Server(hbm_budget=200 << 30, max_active=16)
    for d, cfg in zip(designs, cfgs):
        srv.submit(d, cfg, 0.0)
    t0 = time.time()
    srv.drain(timeout_s=600)
    sched_s = time.time() - t0
    infos = [srv.job(i, positions=False) for i in range(count)]
    srv.close()
    if ([r["iterations"] for r in bres] != alone or
            [i["result"]["iterations"] for i in infos] != alone or
            any(i["state"] != serve.DONE for i in infos)):
        raise RuntimeError("batched jobs differ from the same jobs run alone")
    iters = sum(alone)
    cells = [d.num_movable for d in designs]
    return dict(jobs=count, cells_min=min(cells), cells_max=max(cells), cells_total=sum(cells),
                bins=sorted(set(c.gp_bin_dim for c in cfgs)), iterations=iters,
                batched_s=batched_s, scheduler_s=sched_s, serial_s=seq_s,
                batched_iters_per_s=iters / batched_s, scheduler_iters_per_s=iters / sched_s,
                serial_iters_per_s=iters / seq_s, speedup=seq_s / batched_s,
                scheduler_speedup=seq_s / sched_s)


def serve_bench(count, rate):
    """Poisson arrivals at `rate` jobs/s served together (FP32), against the
    same arrivals served one at a time on the FP64 path; latency = finish -
    arrival (time to overflow 0.10, queueing included)."""
    import numpy as np
    from superplace_amd import design as D
    from superplace_amd import hip, serve
    jobs, arrivals = D.serving_stream(count, rate)

    def run(prec, slots):
        srv = serve.Server(hbm_budget=200 << 30, max_active=slots)
        for (d, dim), t in zip(jobs, arrivals):
            srv.submit(d, hip.GpConfig(gp_bin_dim=dim, precision=prec), float(t))
        srv.drain(timeout_s=900)
        infos = [srv.job(i, positions=False) for i in range(count)]
        srv.close()
        if any(i["state"] != serve.DONE for i in infos):
            raise RuntimeError("a served job failed")
        lat = np.array([i["finish_s"] - i["arrival_s"] for i in infos])
        span = max(i["finish_s"] for i in infos) - float(arrivals[0])
        return dict(p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                    jobs_per_hour=3600.0 * count / span,
                    converged=sum(i["result"]["converged"] for i in infos))

    mix = run(32, 16)
    serial = run(64, 1)
    cells = [d.num_movable for d, _ in jobs]
    return dict(jobs=count, arrival_rate_per_s=rate, cells_min=min(cells), cells_max=max(cells),
                pins_max=max(d.num_pins for d, _ in jobs), served_fp32=mix, serial_fp64=serial)


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


def cpu_baseline(design, iters, min_seconds=10.0, max_runs=8):
    """The oracle's restatement of the reference CPU path (FP32 arithmetic with
    FP64 fixed-chunk reductions, 8 density chunks, FFT-based DCTs in place of
    FFTW, OpenMP over the host's cores) on the first `iters` iterations of the
    same workload, run again from the same start until the loops add up to
    `min_seconds` (at most `max_runs` runs): iterations over loop seconds."""
    import oracle
    threads = os.cpu_count() or 1
    cfg = oracle.default_config(precision=32, gp_bin_dim=512, gp_max_iter=iters)
    oracle.set_fast_poisson(True)
    t0 = time.time()
    its, secs, runs = 0, 0.0, 0
    try:
        while runs < max_runs and (runs == 0 or secs < min_seconds):
            res, rows, _, _ = oracle.eplace_run(design, cfg, threads=threads)
            its += res.iterations
            secs += res.seconds_loop
            runs += 1
    finally:
        oracle.set_fast_poisson(False)
    wall = time.time() - t0
    return dict(value=its / secs, unit="iters/s",
                cores=oracle.get_threads(), kind="port",
                sample=(f"{runs} runs of the first {res.iterations} mGP iterations of the same "
                        f"workload and start (FP32 with FP64 fixed-chunk reductions, 8 density "
                        f"chunks, FFT DCTs, OpenMP); loops {secs:.1f} s, wall incl. setup "
                        f"{wall:.1f} s"))


def copy_calibration():
    """One device-to-device copy of 256 MiB (the runtime's copy kernel): the
    known traffic that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    from superplace_amd import hip
    hip.lib()   # the runtime the product library links
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipFree.argtypes = [C.c_void_p]
    n = 256 << 20
    a, b = C.c_void_p(), C.c_void_p()
    if rt.hipMalloc(C.byref(a), n) or rt.hipMalloc(C.byref(b), n):
        raise RuntimeError("hipMalloc failed")
    try:
        if rt.hipMemcpy(b, a, n, 3) or rt.hipDeviceSynchronize():   # device to device
            raise RuntimeError("copy failed")
    finally:
        rt.hipFree(a)
        rt.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=200)
    ap.add_argument("--no-serve", action="store_true", help="skip the batch and serving runs")
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--serve-jobs", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="arrivals per second")
    ap.add_argument("--roofline-only", default="", metavar="FAMILY",
                    help="run only the roofline leg of this kernel family")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(instances shard across GPUs without communication); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return
    from superplace_amd import design as D
    from superplace_amd import hip, tune

    t0 = time.time()
    d = D.config_adaptec1()
    if args.roofline_only:
        # the roofline leg alone (for a kernel trace of exactly its launches)
        roof = roofline_run(d, args.roofline_only, args.steps, args.warmup)
        print(json.dumps({"roofline": roof}), flush=True)
        return
    log(f"design: {d.num_movable} cells, {d.num_nodes - d.num_movable} fixed, {d.num_fillers} "
        f"fillers, {d.num_pins} pins, {d.num_nets} nets ({time.time() - t0:.1f} s)")

    # 1) FP32 throughput: the graph-launched device loop, nothing else timed
    r32 = timed_run(d, 32, args.steps, args.warmup)
    log(f"FP32: {r32['iters']} timed iterations in {r32['ms']:.1f} ms -> "
        f"{r32['iters_per_s']:.1f} iters/s ({r32['evals']} evaluations); tau "
        f"{r32['last']['overflow']:.4f} HPWL {r32['last']['hpwl']:.6g}")

    # 2) all kernel families timed (slots launched one at a time, every kernel
    #    of a family carrying its HIP event pair) to find the dominant one:
    #    the family with the most time in the run
    probe = timed_run(d, 32, args.steps, args.warmup,
                      profile_mask=(1 << len(hip.KERNEL_FAMILIES)) - 1)
    fam_us = {f: 1e3 * st["ms"] / st["launches"] for f, st in probe["stats"].items()
              if st["launches"] > 0}
    fam_total = {f: st["ms"] for f, st in probe["stats"].items() if st["launches"] > 0}
    dominant = max(fam_total, key=fam_total.get)
    log("kernel families, average us per launch: " +
        ", ".join(f"{f} {v:.1f}" for f, v in fam_us.items()) + f"; dominant {dominant}")

    # 3) roofline of the dominant family, timed alone
    roof = roofline_run(d, dominant, args.steps, args.warmup)
    log(f"roofline {dominant}: {roof['achieved']:.1f} {roof['unit']} over {roof['launches']} "
        f"launches, {roof['avg_launch_us']:.2f} us each")
    evals_per_iter = r32["evals"] / r32["iters"]
    t_roof = iteration_roofline_us(r32["sizes"], evals_per_iter)
    iteration = dict(roofline_us=t_roof, measured_us=1e3 * r32["ms"] / r32["iters"],
                     frac=t_roof / (1e3 * r32["ms"] / r32["iters"]))

    # 3) FP64 path
    r64 = None
    if not args.no_fp64:
        r64 = timed_run(d, 64, args.steps, args.warmup)
        log(f"FP64: {r64['iters_per_s']:.1f} iters/s; tau {r64['last']['overflow']:.4f} "
            f"HPWL {r64['last']['hpwl']:.6g}")

    # 4) many instances: a batch advanced together, and a serving stream
    batch = served = None
    if not args.no_serve:
        batch = batch_bench(args.batch_jobs)
        log(f"batch: {batch['jobs']} jobs ({batch['cells_min']}-{batch['cells_max']} cells), "
            f"{batch['iterations']} iterations in {batch['batched_s']:.2f} s batched, "
            f"{batch['scheduler_s']:.2f} s scheduled, {batch['serial_s']:.2f} s one at a time "
            f"({batch['speedup']:.2f}x, {batch['scheduler_speedup']:.2f}x)")
        served = serve_bench(args.serve_jobs, args.serve_rate)
        m, f = served["served_fp32"], served["serial_fp64"]
        log(f"serving: {served['jobs']} jobs at {args.serve_rate}/s: p50 {m['p50_s']:.3f} s, "
            f"p99 {m['p99_s']:.3f} s, {m['jobs_per_hour']:.0f} jobs/h; one at a time FP64: "
            f"p50 {f['p50_s']:.3f} s, p99 {f['p99_s']:.3f} s, {f['jobs_per_hour']:.0f} jobs/h")

    # 5) CPU baseline (rank 0, N=1 only), bounded sample
    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline: runs of up to {args.cpu_iters} oracle FP32 iterations on "
            f"{os.cpu_count()} cores, at least 10 s of loops ...")
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
        "evals_per_iter": evals_per_iter,
        "iteration_roofline": iteration,
        "kernel_us_per_launch": {f: round(v, 2) for f, v in fam_us.items()},
        "launch_shapes": tune.get_shapes(512),
        "batch": batch,
        "serving": served,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
