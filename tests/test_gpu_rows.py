"""The device loop's decisions, row by row.

EPlace::run (src/gp/eplace.cpp:223-526) writes one metrics row per
iteration: the evaluations so far, HPWL, overflow, the schedules' lambda and
gamma and the Nesterov step (NesterovOptimizer::step, backtracking and BB
step, src/numeric/cpu/nesterov.cpp:51-240).  The engine takes these decisions
on the device (decideStep in the reduction kernel).  Here they are compared
with the oracle's restatement of the same loop, started from the same
positions, row by row:

* FP64 against the FP64 oracle: the evaluation counts exactly, every other
  column within TOL64_ROW relative, over the first ROWS rows (the two sides
  differ only in summation order and in the fixed-point rounding of the
  GPU's density charges: overflow, lambda and gamma come out equal, HPWL
  and the step within ~1e-13);
* FP32 against the FP32 oracle (the reference's arithmetic): the first row
  within TOL32_FIRST, then a band that widens as the FP32 rounding of the
  two implementations moves their trajectories apart;
* the three ways of driving the loop (host decisions, device decisions
  slot by slot, hipGraph-captured slots) give the same rows bit for bit.
"""
import numpy as np
import pytest

import oracle
from superplace_amd import design as D
from superplace_amd import hip

pytestmark = pytest.mark.gpu

ROWS = 40
TOL64_ROW = 1e-11   # measured: at most 1.3e-13 over 40 rows of both designs
TOL32_FIRST = 2e-6
KEYS = ("hpwl", "overflow", "lambda_", "gamma", "step_size")


def designs():
    return {"fixture": (D.with_fillers(D.fixture_eplace(3000, 12)), 0),
            "synth2k": (D.config_small(), 128)}


def orow(r):
    return {"evaluations": r.evaluations, "hpwl": r.hpwl, "overflow": r.overflow,
            "lambda_": r.lambda_, "gamma": r.gamma, "step_size": r.step_size}


def relerr(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def run_both(name, prec, max_iter):
    d, dim = designs()[name]
    ores, orows, _, _ = oracle.eplace_run(
        d, oracle.default_config(precision=prec, gp_bin_dim=dim, gp_max_iter=max_iter))
    res, rows = hip.GpEngine(d, hip.GpConfig(precision=prec, gp_bin_dim=dim,
                                             gp_max_iter=max_iter)).run()
    return res, rows, ores, [orow(r) for r in orows]


@pytest.mark.parametrize("name", ["fixture", "synth2k"])
def test_fp64_rows_match_oracle(name):
    res, rows, ores, orows = run_both(name, 64, ROWS)
    n = min(len(rows), len(orows))
    assert n >= 20, (len(rows), len(orows))
    worst = {k: 0.0 for k in KEYS}
    per_row = []
    for i in range(n):
        a, b = rows[i], orows[i]
        assert a["evaluations"] == b["evaluations"], (i, a, b)
        per_row.append(max(relerr(a[k], b[k]) for k in KEYS))
        for k in KEYS:
            worst[k] = max(worst[k], relerr(a[k], b[k]))
    print(name, n, "rows; worst relative difference", worst, flush=True)
    print(name, "per row", ["%.1e" % w for w in per_row], flush=True)
    for k in KEYS:
        assert worst[k] <= TOL64_ROW, (k, worst)
    if res["iterations"] == ores.iterations:
        assert res["hpwl"] == pytest.approx(ores.hpwl, rel=1e-9)


@pytest.mark.parametrize("name", ["fixture", "synth2k"])
def test_fp32_rows_track_oracle(name):
    res, rows, ores, orows = run_both(name, 32, ROWS)
    n = min(len(rows), len(orows))
    assert n >= 20
    a, b = rows[0], orows[0]
    assert a["evaluations"] == b["evaluations"]
    for k in KEYS:
        assert relerr(a[k], b[k]) <= TOL32_FIRST, (k, a[k], b[k])
    worst = []
    for i in range(n):
        a, b = rows[i], orows[i]
        worst.append(max(relerr(a[k], b[k]) for k in ("hpwl", "overflow", "lambda_")))
    print(name, "FP32 worst relative difference per row", ["%.1e" % w for w in worst],
          flush=True)
    for i, w in enumerate(worst):
        assert w <= 1e-5 * (1 + i), (i, w)


@pytest.mark.parametrize("prec", [32, 64])
def test_loop_modes_rows_bitwise(prec):
    """Host decisions (0), device decisions launched slot by slot (1) and the
    hipGraph-captured loop (2) give the same rows and positions bit for bit."""
    d, dim = designs()["fixture"]
    out = {}
    for mode in (0, 1, 2):
        eng = hip.GpEngine(d, hip.GpConfig(precision=prec, loop_mode=mode, graph_slots=5))
        res, rows = eng.run()
        out[mode] = (res, rows, eng.positions())
    res0, rows0, (x0, y0) = out[0]
    for mode in (1, 2):
        res, rows, (x, y) = out[mode]
        assert len(rows) == len(rows0)
        for i, (p, q) in enumerate(zip(rows, rows0)):
            for k in ("evaluations",) + KEYS:
                assert p[k] == q[k], (mode, i, k, p[k], q[k])
        assert res["hpwl"] == res0["hpwl"] and res["overflow"] == res0["overflow"]
        assert np.array_equal(x, x0) and np.array_equal(y, y0)
