"""Launch-shape autotuner of the device loop (sp_hip_shapes).

For each bin-grid size it times the FP32 loop of a representative synthetic
design under candidate shapes and keeps the fastest, one knob at a time
(coordinate descent, two passes):

- col_pairs: the Poisson column pass sharing its launch with the WA blocks
  (0), or on its own with 1, 2, 4 or 8 column pairs per block;
- scatter_chunk: density records per scatter block (256 to 2048);
- sort_period: slots between the sort passes' launches;
- graph_slots: iterations per graph launch.

No shape changes a result bit (the scatter sums integers, the sort only
reorders records, every column pair is transformed alike, the graph only
groups iterations); every timed run is checked against the first one of its
size, so the tuning itself tests that claim.

    python -m superplace_amd.tune [--dims 128 256 512 1024] [--write]

prints the table, and with --write stores it in superplace_amd/shapes.json,
which the library's compiled-in table (kTunedShapes in csrc/engine.hip)
mirrors.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES_JSON = os.path.join(HERE, "shapes.json")
KNOBS = {
    "col_pairs": [0, 1, 2, 4, 8],
    "scatter_chunk": [256, 512, 1024, 2048],
    "sort_period": [1, 2, 4],
    "graph_slots": [4, 8, 16],
}
DEFAULT = {"scatter_chunk": 512, "sort_period": 2, "col_pairs": 0, "graph_slots": 8}


def set_shapes(dim: int, shapes: dict) -> None:
    from . import hip
    s = hip.SpHipShapes(**shapes)
    hip.check(hip.lib().sp_hip_shapes_set(int(dim), C.byref(s)))


def get_shapes(dim: int) -> dict:
    from . import hip
    s = hip.SpHipShapes()
    hip.check(hip.lib().sp_hip_shapes_get(int(dim), C.byref(s)))
    return s.as_dict()


def reset_shapes() -> None:
    from . import hip
    hip.check(hip.lib().sp_hip_shapes_reset())


def design_for(dim: int):
    """The benchmark's adaptec1-shaped design at 512 bins per side; else a
    seeded synthetic design of dim^2 cells (the automatic rule then picks
    dim), the config-1 degree distribution and 1.05 nets per cell."""
    from . import design as D
    if dim == 512:
        return D.config_adaptec1()
    cells = dim * dim
    d = D.synth_design(cells, max(8, cells // 400), int(cells * 1.05), 25.25 * np.sqrt(cells),
                       0.57, seed=7000 + dim, name=f"tune{dim}")
    if D.auto_bin_dim(d, 4, 1024) != dim:
        raise RuntimeError(f"tuning design for {dim} bins gets another grid size")
    return d


def time_loop(d, dim: int, shapes: dict, iters: int, repeats: int):
    """Seconds per iteration of `iters` loop iterations from the design's
    start (best of `repeats`), and the run's result signature; (None, None)
    where the instance does not take these shapes (a column pass whose LDS
    would not fit)."""
    from . import hip
    set_shapes(dim, shapes)
    best, sig = None, None
    for _ in range(repeats):
        eng = hip.GpEngine(d, hip.GpConfig(gp_bin_dim=dim, gp_max_iter=iters,
                                           target_overflow=1e-6, gp_stagnation_window=0))
        if eng.info["shapes"].as_dict() != shapes:
            eng.close()
            return None, None
        res, _ = eng.run(max_rows=1)
        x, y = eng.positions()
        eng.close()
        t = res["seconds_loop"] / max(res["iterations"], 1)
        best = t if best is None else min(best, t)
        s = (res["iterations"], res["evaluations"], res["hpwl"], res["overflow"],
             hash(x.tobytes()), hash(y.tobytes()))
        if sig is not None and s != sig:
            raise RuntimeError("a repeated run gave another result")
        sig = s
    return best, sig


def tune_dim(dim: int, iters: int = 60, repeats: int = 3, log=print) -> dict:
    d = design_for(dim)
    cur = dict(DEFAULT)
    ref_sig = None
    timings = {}

    def measure(shapes):
        nonlocal ref_sig
        key = tuple(sorted(shapes.items()))
        if key not in timings:
            t, sig = time_loop(d, dim, shapes, iters, repeats)
            if t is None:
                timings[key] = float("inf")
                log(f"  {dim}: {shapes} -> not taken at this size")
                return timings[key]
            if ref_sig is None:
                ref_sig = sig
            elif sig != ref_sig:
                raise RuntimeError(f"shapes {shapes} changed the result at {dim} bins")
            timings[key] = t
            log(f"  {dim}: {shapes} -> {1e6 * t:.1f} us/iteration")
        return timings[key]

    base = measure(cur)
    for _ in range(2):
        for knob, values in KNOBS.items():
            best_v, best_t = cur[knob], measure(cur)
            for v in values:
                t = measure(dict(cur, **{knob: v}))
                if t < best_t:
                    best_v, best_t = v, t
            cur[knob] = best_v
    final = measure(cur)
    reset_shapes()
    return {"dim": dim, "shapes": cur, "us_per_iteration": 1e6 * final,
            "default_us_per_iteration": 1e6 * base, "design": d.name,
            "cells": int(d.num_movable), "iterations": iters,
            "runs": sum(1 for t in timings.values() if t != float("inf")) * repeats}


def c_table(entries) -> str:
    lines = ["constexpr DimShapes kTunedShapes[] = {",
             "    {0, {kTileChunk, kSortPeriod, 0, 8}},"]
    for e in entries:
        s = e["shapes"]
        lines.append(f"    {{{e['dim']}, {{{s['scatter_chunk']}, {s['sort_period']}, "
                     f"{s['col_pairs']}, {s['graph_slots']}}}}},")
    lines.append("};")
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--dims", type=int, nargs="+", default=[128, 256, 512, 1024])
    ap.add_argument("--iters", type=int, default=60)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--write", action="store_true", help="store the table in shapes.json")
    args = ap.parse_args(argv)
    from . import hip
    if hip.device_count() < 1:
        print("the autotuner times kernels: it needs a GPU", file=sys.stderr)
        return 2
    reset_shapes()
    t0 = time.time()
    entries = []
    for dim in args.dims:
        e = tune_dim(dim, args.iters, args.repeats, log=lambda m: print(m, flush=True))
        print(f"{dim}: {e['shapes']} {e['us_per_iteration']:.1f} us/iteration "
              f"(defaults {e['default_us_per_iteration']:.1f})", flush=True)
        entries.append(e)
    table = {"tool": "python -m superplace_amd.tune " + " ".join(
                 ["--dims"] + [str(x) for x in args.dims] +
                 ["--iters", str(args.iters), "--repeats", str(args.repeats)]),
             "seconds": round(time.time() - t0, 1), "entries": entries}
    print(json.dumps(table, indent=1))
    print(c_table(entries))
    if args.write:
        with open(SHAPES_JSON, "w") as f:
            json.dump(table, f, indent=1)
            f.write("\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
