"""Scripted MetricsSink call sequences shared by make_metrics_golden.py (run
through the reference's own MetricsSink) and tests/test_metrics.py (run
through sp_metrics_*).  An operation is a list:
  ["begin", stage] ["push", row] ["end", ms] ["duration", label, ms]
  ["num", key, value] ["str", key, value] ["write"]  (summary.json)
Rows use the field names of IterMetrics (src/util/metrics.h:12-28)."""
import numpy as np

FIELDS = ["iter", "hpwl", "overflow", "lambda", "gamma", "step_size", "grad_norm_wl",
          "grad_norm_den", "max_displacement", "elapsed_ms", "eval_lambda", "eval_gamma",
          "delta_hpwl", "lambda_mu", "evaluations"]


def _rows(rng, n, start_hpwl):
    rows = []
    hp = start_hpwl
    for i in range(n):
        dh = float(rng.normal(-0.01, 0.02) * hp)
        hp = hp + dh
        rows.append(dict(iter=i, hpwl=hp, overflow=float(max(0.0, 0.95 - 0.03 * i)),
                         **{"lambda": float(1e-4 * 1.05 ** i)}, gamma=float(80.0 / (1 + i)),
                         step_size=float(rng.uniform(0, 30)), grad_norm_wl=float(rng.uniform(0, 1e3)),
                         grad_norm_den=float(rng.uniform(0, 1e-3)), max_displacement=0.0,
                         elapsed_ms=float(0.11 * (i + 1)), eval_lambda=float(1e-4 * 1.05 ** i),
                         eval_gamma=float(80.0 / (1 + i)), delta_hpwl=dh,
                         lambda_mu=float(rng.choice([0.95, 1.0, 1.05, 1.0312345])),
                         evaluations=int(i + 1 + (i % 3 == 0))))
    return rows


def scenarios():
    rng = np.random.default_rng(7)
    flow = [["str", "benchmark", "adaptec1"], ["str", "aux_path", "/data/adaptec1/adaptec1.aux"],
            ["duration", "read", 812.5], ["begin", "init"]]
    for i in range(7):
        flow.append(["push", dict(iter=i, hpwl=4.1e7 * (1 + 0.3 / (i + 1)), grad_norm_wl=1e-6 * (i + 1),
                                  elapsed_ms=36.0 + i)])
    flow += [["end", 259.25], ["begin", "gp"], ["begin", "mGP"]]
    flow += [["push", r] for r in _rows(rng, 40, 1.0275e8)]
    flow += [["end", 4711.0625], ["num", "mGP_initial_hpwl", 1.0275e8],
             ["num", "mGP_initial_overflow", float(np.float32(0.9543))],
             ["num", "mGP_converged", 1], ["num", "mGP_rollback_iter", -1],
             ["num", "mgp_hpwl", 7.734567890123e7], ["str", "poisson_backend", "hip-fft"],
             ["num", "fftw_threads", 0], ["end", 5000.0],
             ["str", "note", 'quotes " back\\slash\nnewline\ttab \x01 ctl, café'],
             ["str", "run_status", "completed"], ["write"]]
    failure = [["begin", "gp"], ["begin", "mGP"]]
    failure += [["push", r] for r in _rows(rng, 5, 3.3e5)]
    failure += [["str", "run_status", "failed"],
                ["str", "failure_message", "non-finite gradient at iter 4"],
                ["str", "failure_kind", "non_finite"], ["str", "failure_stage", "mGP"],
                ["end", 0.0], ["write"]]
    nested = [["begin", "A"], ["push", dict(iter=0, hpwl=1.0)], ["begin", "B"],
              ["push", dict(iter=0, hpwl=2.0)], ["end", 1.5], ["push", dict(iter=1, hpwl=3.0)],
              ["begin", "B"], ["push", dict(iter=0, hpwl=4.0)], ["end", 2.5], ["end", 3.5],
              ["push", dict(iter=9, hpwl=5.0)], ["end", 4.5], ["write"]]
    return {"flow": flow, "failure": failure, "nested": nested}


def full_row(r):
    out = {f: 0 for f in FIELDS}
    out["lambda_mu"] = 1.0
    out.update(r)
    return out
