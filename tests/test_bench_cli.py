"""bench.py's host-side parts, without a GPU: the one-GPU guard, the
automatic bin dimension it sizes the batch jobs with (against the oracle's
BinGrid, bin_grid.cpp:172-189), the committed traffic table it reads, and
the CPU-baseline record."""
import os
import subprocess
import sys

import pytest

import oracle
from superplace_amd import design as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


@pytest.mark.parametrize("seed0", [0, 7])
def test_auto_bin_dim_is_the_reference_rule(seed0):
    for d, dim in bench.job_designs(3, 3000, 20000, seed0, 4, 1024):
        assert dim == oracle.Grid(d, 1.0, 0).dim, d.name


def test_auto_bin_dim_clamps():
    d = D.config_small()
    auto = oracle.Grid(d, 1.0, 0).dim
    assert bench.auto_bin_dim(d, 4, 1024) == auto
    assert bench.auto_bin_dim(d, 2 * auto, 1024) == 2 * auto
    assert bench.auto_bin_dim(d, 4, auto // 2) == auto // 2


def test_traffic_table_has_the_dominant_family():
    for fam in ("poisson", "scatter", "wa_nets", "node_grad", "update"):
        v = bench.measured_traffic(fam)
        assert v is not None and v > 0, fam
    assert bench.measured_traffic("no_such_family") is None


def test_cpu_baseline_record():
    r = bench.cpu_baseline(D.config_small(), 3, min_seconds=0.0)
    assert r["value"] > 0 and r["unit"] == "iters/s" and r["kind"] == "port"
    assert r["cores"] >= 1
    assert r["sample"].startswith("1 runs of the first 3 mGP iterations")


def test_algorithmic_bytes_follow_the_measurement_plan():
    """Per-launch algorithmic work at the benchmark sizes (n = 370,176
    optimised nodes, 944k pins, 221k nets, 512^2 bins): scatter 8.4 MB,
    rho/tau 4.2, WA + HPWL 19.4 + node pass 19.7 (27.8 + gather 11.4
    together), Nesterov 23.7 MB; DCT 1.879 GFLOP; one evaluation at roofline
    about 21 us."""
    z = dict(n=370176, B=512 * 512, dim=512, P=944000, E=221000, Nmov=211000, Nfix=543)
    w = bench.algorithmic_work(z, False)
    mb = {k: v[0] / 1e6 for k, v in w.items()}
    assert abs(mb["scatter"] - 8.39) < 0.01 and abs(mb["finalize"] - 4.19) < 0.01
    assert abs(mb["update"] - 23.69) < 0.01
    assert abs(mb["wa_nets"] + mb["node_grad"] - (27.8 + 11.4)) < 0.2
    assert abs(w["poisson"][1] / 1e9 - 1.879) < 0.001
    assert abs(bench.algorithmic_work(z, True)["poisson"][0] - 24 * z["B"]) == 0
    assert 20.5 < bench.iteration_roofline_us(z, 1.0) < 22.0
