"""Summaries of rocprofv3 CSV output for the profiles/ folder.

    python -m superplace_amd.profile_summary trace <dir> [--title T] [--command C]
        per-kernel calls / average / min / max / share from <dir>/run_kernel_trace.csv,
        and the Poisson solve (the FFT passes of one evaluation) summed;
    python -m superplace_amd.profile_summary pmc <dir> [--title T] [--command C]
        per-kernel medians of every counter in <dir>/run_counter_collection.csv;
    python -m superplace_amd.profile_summary traffic <fetch_dir> <write_dir>
            <calib_fetch_dir> <calib_write_dir> --json OUT [--title T] [--command C]
        HBM bytes per dispatch from FETCH_SIZE and WRITE_SIZE (KiB), each scaled
        by what the same counter reads for a plain 256 MiB copy (the
        calibration runs), per kernel and per kernel family; OUT is the JSON
        table bench.py reads (profiles/hbm_traffic.json).

Markdown goes to stdout.  Dispatches that return at once (the device loop's
launches after its stage ended) are left out of medians by taking the
median over all dispatches of a kernel.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import os
import re
import statistics
import sys

CALIB_BYTES = 256 * 1024 * 1024   # the calibration copy: 256 MiB read, 256 MiB written

# kernel -> family of the device loop (the bench's kernel families)
FAMILIES = {
    "k_scatter": "scatter",
    "k_finalize": "finalize",
    "k_dct_rows_rho": "poisson", "k_dct_cols": "poisson", "k_idct_rows": "poisson",
    "k_dct_rows": "poisson_host", "k_gemm": "poisson_dense",
    "k_wa_pins": "wa_nets",
    "k_dct_cols_wa": "k_dct_cols_wa",
    "k_node_grad": "node_grad",
    "k_update": "update",
    "k_tile_count": "sort", "k_tile_place": "sort",
    "k_reduce": "reduce",
}
POISSON_PASSES = ("k_dct_rows_rho", "k_dct_rows", "k_dct_cols", "k_dct_cols_wa", "k_idct_rows")


def short(name: str) -> str:
    m = re.search(r"(k_\w+)(<[^>]*>)?", name)
    return m.group(0) if m else name.split("(")[0][:48]


def base(name: str) -> str:
    return short(name).split("<")[0]


def read(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def header(title, command):
    out = []
    if title:
        out.append(f"# {title}\n")
    if command:
        out.append("Command (from the repository root, on one MI355X):\n")
        out.append("```\n" + command + "\n```\n")
    return out


def trace(args):
    rows = read(os.path.join(args.dir, "run_kernel_trace.csv"))
    per = collections.defaultdict(list)
    seq = []
    for r in rows:
        t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        k = short(r["Kernel_Name"])
        per[k].append((t1 - t0) / 1e3)
        seq.append((t0, t1, k))
    total = sum(sum(v) for v in per.values())
    out = header(args.title, args.command)
    out.append("| kernel | calls | avg us | median us | min us | max us | share |")
    out.append("|---|---|---|---|---|---|---|")
    for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"| `{k}` | {len(v)} | {statistics.mean(v):.2f} | {statistics.median(v):.2f} | "
                   f"{min(v):.2f} | {max(v):.2f} | {100 * sum(v) / total:.1f}% |")
    # Poisson solves: a first pass followed by the other passes of the same solve
    seq.sort()
    solves, cur = [], None
    for t0, t1, k in seq:
        b = k.split("<")[0]
        if b in ("k_dct_rows_rho", "k_dct_rows"):
            if cur:
                solves.append(cur)
            cur = [(t0, t1, k)]
        elif b in POISSON_PASSES and cur is not None:
            cur.append((t0, t1, k))
            if b == "k_idct_rows":
                solves.append(cur)
                cur = None
    full = [s for s in solves if len(s) == 3 and (s[0][1] - s[0][0]) > 3000]
    if full:
        sums = [sum((t1 - t0) / 1e3 for t0, t1, _ in s) for s in full]
        spans = [(s[-1][1] - s[0][0]) / 1e3 for s in full]
        out.append("")
        out.append(f"Poisson solve = the 3 FFT pass launches of one evaluation. Over {len(full)} "
                   f"solves (early exits after convergence excluded): summed kernel durations "
                   f"mean {statistics.mean(sums):.2f} us, median {statistics.median(sums):.2f} us; "
                   f"first start to last end (the span a HIP event pair around the solve "
                   f"times) mean {statistics.mean(spans):.2f} us, median "
                   f"{statistics.median(spans):.2f} us.")
    print("\n".join(out))


def per_kernel_counters(d):
    rows = read(os.path.join(d, "run_counter_collection.csv"))
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in rows:
        per[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return per


def pmc(args):
    per = per_kernel_counters(args.dir)
    names = sorted({c for v in per.values() for c in v})
    out = header(args.title, args.command)
    out.append("| kernel | dispatches | " + " | ".join(f"median {c}" for c in names) + " |")
    out.append("|---|---|" + "---|" * len(names))
    for k, v in sorted(per.items()):
        n = max(len(x) for x in v.values())
        out.append(f"| `{k}` | {n} | " +
                   " | ".join(f"{statistics.median(v[c]):.6g}" if c in v else "" for c in names) + " |")
    print("\n".join(out))


def calib_factor(d, counter):
    per = per_kernel_counters(d)
    vals = [x for k, v in per.items() if "copy" in k.lower() for x in v.get(counter, [])]
    if not vals:
        raise SystemExit(f"no copy kernel with {counter} in {d}")
    kib = max(vals)   # the 256 MiB copy is the largest copy dispatch
    return kib * 1024.0 / CALIB_BYTES


def traffic(args):
    ff = calib_factor(args.calib_fetch, "FETCH_SIZE")
    wf = calib_factor(args.calib_write, "WRITE_SIZE")
    fetch = per_kernel_counters(args.fetch)
    write = per_kernel_counters(args.write)
    kernels = {}
    for k in sorted(set(fetch) | set(write)):
        if not k.startswith("k_"):
            continue
        fb = statistics.median(fetch[k]["FETCH_SIZE"]) * 1024 / ff if "FETCH_SIZE" in fetch[k] else 0.0
        wb = statistics.median(write[k]["WRITE_SIZE"]) * 1024 / wf if "WRITE_SIZE" in write[k] else 0.0
        kernels[k] = {"fetch_bytes": fb, "write_bytes": wb}
    fams = collections.defaultdict(float)
    for k, v in kernels.items():
        f = FAMILIES.get(base(k))
        if f:
            # a Poisson solve is its passes' sum; the fused column pass (with
            # the WA blocks) is its own family
            fams[f] += v["fetch_bytes"] + v["write_bytes"]
    out = header(args.title, args.command)
    out.append(f"Calibration: a plain copy of 256 MiB reads FETCH_SIZE = {ff * CALIB_BYTES / 1024:.0f} KiB "
               f"and WRITE_SIZE = {wf * CALIB_BYTES / 1024:.0f} KiB per dispatch, so bytes = KiB x 1024 / "
               f"{ff:.3f} (fetch) and KiB x 1024 / {wf:.3f} (write).\n")
    out.append("Median over dispatches:\n")
    out.append("| kernel | fetch MB | write MB |")
    out.append("|---|---|---|")
    for k, v in kernels.items():
        out.append(f"| `{k}` | {v['fetch_bytes'] / 1e6:.2f} | {v['write_bytes'] / 1e6:.2f} |")
    out.append("\nPer family and launch (the Poisson family is one solve: its pass launches):\n")
    out.append("| family | MB per launch |")
    out.append("|---|---|")
    for f, b in sorted(fams.items(), key=lambda kv: -kv[1]):
        out.append(f"| {f} | {b / 1e6:.2f} |")
    print("\n".join(out))
    table = {"command": args.command, "calibration": {
                 "kernel": "plain copy of 256 MiB (read 256 MiB, write 256 MiB)",
                 "fetch_factor": ff, "write_factor": wf,
                 "note": "counter KiB x 1024 / factor = bytes"},
             "kernels": kernels,
             "families": {f: {"bytes_per_launch": b} for f, b in fams.items()}}
    with open(args.json, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    for name in ("trace", "pmc"):
        p = sub.add_parser(name)
        p.add_argument("dir")
        p.add_argument("--title", default="")
        p.add_argument("--command", default="")
    p = sub.add_parser("traffic")
    for a in ("fetch", "write", "calib_fetch", "calib_write"):
        p.add_argument(a)
    p.add_argument("--json", required=True)
    p.add_argument("--title", default="")
    p.add_argument("--command", default="")
    args = ap.parse_args(argv)
    {"trace": trace, "pmc": pmc, "traffic": traffic}[args.cmd](args)
    return 0


if __name__ == "__main__":
    sys.exit(main())
