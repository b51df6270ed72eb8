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
