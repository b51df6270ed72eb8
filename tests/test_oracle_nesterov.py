"""Pins the oracle's Nesterov optimizer against the reference's own
NesterovOptimizer (golden vectors from oracle/_ref) and the behavioural
checks of test/test_nesterov.cpp:36-235."""
import glob
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ref_nesterov_*.npz"))))
def test_matches_reference_bitwise(path):
    g = np.load(path)
    steps, bb, bt, init, hi = g["params"]
    obj, pos, ev = oracle.nesterov_quadratic(g["coef"], g["target"], g["pos0"], int(steps),
                                             int(bb), int(bt), float(init), float(hi))
    assert np.array_equal(pos, g["pos"])
    assert ev == int(g["evals"])
    assert obj == float(g["obj"])


def test_backtrack_recovers_from_huge_step():
    c, t, p0 = np.full(32, 4.0), np.full(32, 3.0), np.zeros(32)
    ob, _, eb = oracle.nesterov_quadratic(c, t, p0, 3, 1, 1, 12.5)
    op, _, ep = oracle.nesterov_quadratic(c, t, p0, 3, 1, 0, 12.5)
    assert ob < 1e-6 and op > 1e3 and eb == ep + 1


def test_backtrack_noop_when_safe_and_convergence():
    c, t = np.full(64, 2.0), 10.0 + 0.5 * np.arange(64)
    _, pa, ea = oracle.nesterov_quadratic(c, t, np.zeros(64), 100, 1, 1)
    _, pb, eb = oracle.nesterov_quadratic(c, t, np.zeros(64), 100, 1, 0)
    assert np.array_equal(pa, pb) and ea == eb
    _, p, _ = oracle.nesterov_quadratic(c, t, np.zeros(64), 200, 0, 0)
    assert np.abs(p - t).max() < 1e-3
    _, p, _ = oracle.nesterov_quadratic(np.ones(16), np.full(16, 1000.0), np.zeros(16), 100,
                                        1, 0, 0.0, 5.0)
    assert p.max() <= 5.0 + 1e-4 and p.max() > 5.0 - 1e-2
