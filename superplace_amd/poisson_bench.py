"""The Poisson backend benchmark of the reference's superplace-poisson-bench
tool (tools/poisson_bench/main.cpp) for the HIP backend: the same input
(rho[i] = sin(0.013 i) + cos(0.037 i) + (i mod 97)/97 over an n x n grid,
bin steps 1.25 x 0.75), 10 untimed solves, `rounds` timed ones through the
host-span PoissonBackend::solve (host copies included, as for the
reference's CUDA backend), and one JSON line with the tool's keys plus
"backend".

    python -m superplace_amd.poisson_bench [n=512] [workers=4] [rounds=100] [omp_threads=1]
                                           [transforms=auto|dense|bf16x3]

transforms (no counterpart in the tool): auto takes the FFT passes on
power-of-two grids and the dense MFMA products elsewhere; dense takes the
products on every grid; bf16x3 the products as split-BF16 products.
"""
from __future__ import annotations

import json
import sys
import time

import numpy as np

from . import hip


def density(n: int) -> np.ndarray:
    i = np.arange(n * n, dtype=np.int64)
    f = i.astype(np.float32)
    rho = (np.sin(f * np.float32(0.013)) + np.cos(f * np.float32(0.037))
           + (i % 97).astype(np.float32) / np.float32(97.0))
    return rho.astype(np.float32).reshape(n, n)


def run(n=512, workers=4, rounds=100, omp_threads=1, transforms="auto") -> dict:
    if n < 2 or rounds < 1 or omp_threads < 1:
        raise ValueError("invalid benchmark dimensions/iterations/threads")
    if transforms not in ("auto", "dense", "bf16x3"):
        raise ValueError("transforms must be auto, dense or bf16x3")
    rho = density(n)
    out = [np.zeros((n, n), np.float32) for _ in range(3)]
    t0 = time.perf_counter()
    be = hip.make_poisson_backend(hip.BackendConfig(n, n, 1.25, 0.75, fftw_threads=workers,
                                                    dense=transforms != "auto",
                                                    bf16x3=transforms == "bf16x3"))
    plan_ms = (time.perf_counter() - t0) * 1e3
    for _ in range(10):
        be.solve(rho, *out)
    t1 = time.perf_counter()
    for _ in range(rounds):
        be.solve(rho, *out)
    solve_ms = (time.perf_counter() - t1) * 1e3 / rounds
    res = dict(grid=n, fftw_threads=workers, fftw_plan_threads=be.cpu_threads(),
               omp_threads=omp_threads, iterations=rounds, plan_ms=plan_ms, solve_ms=solve_ms,
               checksum=float(np.sum(out[0], dtype=np.float64)), backend=be.name())
    be.close()
    return res


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    args = [int(a) for a in argv[:4]] + list(argv[4:5])
    try:
        print(json.dumps(run(*args), separators=(",", ":")))
    except Exception as e:   # the tool prints the message and exits 1
        print(e, file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
