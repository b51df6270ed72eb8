"""metrics.csv / summary.json writers (sp_metrics_*, superplace_amd.metrics)
against the reference's own MetricsSink (src/util/metrics.cpp): the files it
wrote for the scripted call sequences of tests/golden/metrics_scenarios.py
are committed under tests/golden/metrics/ (make_metrics_golden.py); ours
must be byte-identical.  Host code only."""
import json
import os
import sys

import pytest

from superplace_amd import hip, metrics

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
from metrics_scenarios import full_row, scenarios  # noqa: E402


def run(ops, out_dir):
    m = metrics.MetricsSink(str(out_dir))
    for op in ops:
        kind = op[0]
        if kind == "begin":
            m.begin_stage(op[1])
        elif kind == "push":
            r = full_row(op[1])
            md = r.pop("max_displacement")
            m.push(r, max_displacement=[md])
        elif kind == "end":
            m.end_stage(op[1])
        elif kind == "duration":
            m.record_duration(op[1], op[2])
        elif kind in ("num", "str"):
            m.set_summary(op[1], op[2])
        elif kind == "write":
            m.write_summary(os.path.join(str(out_dir), "summary.json"))
    m.close()


@pytest.mark.parametrize("name", ["flow", "failure", "nested"])
def test_files_byte_identical(name, tmp_path):
    run(scenarios()[name], tmp_path)
    for fn in ("metrics.csv", "summary.json"):
        got = (tmp_path / fn).read_bytes()
        want = open(os.path.join(GOLD, "metrics", name, fn), "rb").read()
        assert got == want, fn


def test_summary_is_json_and_failure_contract(tmp_path):
    """The failure summary keeps the keys main.cpp:125-138 writes."""
    run(scenarios()["failure"], tmp_path)
    s = json.loads((tmp_path / "summary.json").read_text())
    assert s["run_status"] == "failed"
    for k in ("failure_message", "failure_kind", "failure_stage"):
        assert k in s
    assert s["durations_ms"]["mGP"] == 0


def test_non_finite_rejected(tmp_path):
    m = metrics.MetricsSink(str(tmp_path))
    m.begin_stage("mGP")
    with pytest.raises(hip.NonFinite):
        m.push(dict(iter=0, hpwl=float("nan")))
    with pytest.raises(hip.NonFinite):
        m.set_summary("mgp_hpwl", float("inf"))
    with pytest.raises(hip.InvalidArgument):
        m.write_summary(str(tmp_path / "missing_dir" / "summary.json"))


def test_gp_result_keys(tmp_path):
    """The stage keys EPlace::run writes (eplace.cpp:266-270, :509-513)."""
    m = metrics.MetricsSink(str(tmp_path))
    res = dict(iterations=500, converged=1, stagnated=0, rollback_iter=-1, evaluations=812,
               hpwl=7.7e7, overflow=0.0998, initial_hpwl=1.02e8, initial_overflow=0.95,
               initial_lambda=1.5e-4, seconds_loop=0.07)
    m.gp_result("mGP", res, hpwl_reference=3.5e5, restarts=0)
    m.write_summary(str(tmp_path / "s.json"))
    s = json.loads((tmp_path / "s.json").read_text())
    want = {"mGP_" + k for k in ("initial_hpwl", "initial_overflow", "initial_lambda",
                                 "hpwl_reference", "restarts", "evaluations", "converged",
                                 "stagnated", "rollback_iter")}
    assert set(s) == want
    assert s["mGP_rollback_iter"] == -1 and s["mGP_evaluations"] == 812
