"""Pins the oracle's WA wirelength, HPWL and schedules against the
reference's own code (golden vectors from oracle/_ref), the 2-pin closed
form, central differences (test/test_wirelength.cpp:67-112) and hand values."""
import os

import numpy as np
import pytest

import oracle
from tests.helpers import small_netlist
from superplace_amd import design as D

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = {
    "three_pin": (3, 0, [[0, 1, 2]], 11, [50.0]),
    "fixed_pin": (3, 1, [[0, 1, 2, 3]], 22, [50.0]),
    "shared": (5, 0, [[0, 1, 2], [1, 3], [2, 3, 4], [0, 4]], 33, [50.0]),
    "gamma_sweep": (6, 2, [[0, 1, 2, 6], [2, 3, 4], [4, 5, 7], [0, 5]], 44, [500.0, 50.0, 5.0]),
    "forty_pin": (40, 0, [list(range(40))], 55, [20.0]),
    "tiny_gamma": (8, 0, [[0, 1, 2, 3], [4, 5, 6, 7]], 66, [0.01]),
}


@pytest.mark.parametrize("name", list(CASES))
def test_wa_matches_reference_golden(name):
    nm, nf, nets, seed, gammas = CASES[name]
    d = small_netlist(nm, nf, nets, seed)
    g = np.load(os.path.join(GOLD, f"ref_wa_{name}.npz"))
    for gam in gammas:
        wl32, gr32 = oracle.wa(d, gam, 1000, precision=32)
        assert wl32 == pytest.approx(float(g[f"wl_{gam}"]), rel=1e-6, abs=1e-4)
        assert np.allclose(gr32, g[f"grad_{gam}"], atol=1e-6 * max(1, np.abs(g[f"grad_{gam}"]).max()))
        wl64, gr64 = oracle.wa(d, gam, 1000, precision=64)
        assert np.isfinite(wl64) and np.all(np.isfinite(gr64))
        assert np.allclose(gr64, g[f"grad_{gam}"], atol=1e-4 * max(1, np.abs(gr64).max()))


def test_wa_hpwl_synth_golden():
    d = D.synth_design(300, 3, 320, die=200.0, util=0.6, tail="powerlaw", tail_max=150,
                       seed=5, with_fillers=False, locality=4.0)
    g = np.load(os.path.join(GOLD, "ref_wa_synth300.npz"))
    wl, gr = oracle.wa(d, 7.5, 100, precision=32)
    assert wl == pytest.approx(float(g["wl"]), rel=1e-6)
    assert np.allclose(gr, g["grad"], atol=1e-6 * np.abs(g["grad"]).max())
    assert oracle.hpwl(d, 32) == pytest.approx(float(g["hpwl"]), rel=1e-12)
    assert oracle.hpwl(d, 64) == pytest.approx(float(g["hpwl"]), rel=1e-6)


def test_two_pin_closed_form():
    # WA of a 2-pin net with separation delta: delta*tanh(delta/(2 gamma));
    # d/dx2 = tanh(t) + t sech^2(t), t = delta/(2 gamma).
    for delta, gamma in [(30.0, 50.0), (200.0, 10.0), (3.0, 100.0)]:
        rows = (np.array([0.0]), np.array([1000.0]), np.array([0.0]), np.array([1000.0]))
        d = D.make_design([100.0, 100.0 + delta], [500.0, 500.0], [2, 2], [2, 2], [0, 0], 2,
                          [0, 1], [0, 0], [0, 0], [0, 0], 1, rows)
        wl, g = oracle.wa(d, gamma, 100, precision=64)
        t = delta / (2 * gamma)
        assert wl == pytest.approx(delta * np.tanh(t), rel=1e-12)
        assert g[1] == pytest.approx(np.tanh(t) + t / np.cosh(t) ** 2, rel=1e-10)
        assert g[0] == pytest.approx(-g[1], rel=1e-10)


def test_central_difference_gradient():
    d = small_netlist(6, 2, [[0, 1, 2, 6], [2, 3, 4], [4, 5, 7], [0, 5]], 44)
    for gamma in (500.0, 50.0, 5.0, 20.0):
        _, g = oracle.wa(d, gamma, 1000)
        denom = max(np.abs(g).max(), 1e-9)
        h = 0.5
        for i in range(d.num_movable):
            for ax in (0, 1):
                arr = d.node_x if ax == 0 else d.node_y
                x0 = arr[i]
                arr[i] = x0 + h
                wp, _ = oracle.wa(d, gamma, 1000)
                arr[i] = x0 - h
                wm, _ = oracle.wa(d, gamma, 1000)
                arr[i] = x0
                assert abs((wp - wm) / (2 * h) - g[ax * d.num_movable + i]) / denom < 1e-3


def test_wa_monotone_to_hpwl_and_finite():
    d = small_netlist(10, 0, [[0, 1, 2, 3, 4], [5, 6, 7, 8, 9], [0, 9]], 77)
    prev = -1
    for gamma in (200.0, 50.0, 10.0, 1.0):
        wl, _ = oracle.wa(d, gamma, 1000)
        assert wl >= prev - 1e-9
        prev = wl
    assert prev <= oracle.hpwl(d) + 1e-9
    wl, g = oracle.wa(small_netlist(8, 0, [[0, 1, 2, 3], [4, 5, 6, 7]], 66), 0.01, 1000, 32)
    assert np.isfinite(wl) and np.all(np.isfinite(g))


def test_hpwl_hand_value():
    # test/test_hpwl.cpp:13-56: HPWL = 130
    rows = (np.array([0.0]), np.array([80.0]), np.array([0.0]), np.array([80.0]))
    d = D.make_design([10, 50, 30], [10, 20, 60], [10] * 3, [10] * 3, [0, 0, 0], 3,
                      [0, 1, 2, 0, 1], [0, 0, 0, 1, 1], [0, 0, 0, 3, -2], [0, 0, 0, 4, -1], 2, rows)
    assert oracle.hpwl(d, 64) == pytest.approx(130.0, abs=1e-9)
    assert oracle.hpwl(d, 32) == pytest.approx(130.0, abs=1e-9)


def test_large_nets_skipped_in_wa_not_hpwl():
    # a 120-pin net: ignored by WA (degree > 100), counted by HPWL
    n = 120
    rows = (np.array([0.0]), np.array([1000.0]), np.array([0.0]), np.array([1000.0]))
    xs = np.linspace(10, 990, n)
    d = D.make_design(xs, np.full(n, 500.0), [1] * n, [1] * n, [0] * n, n, list(range(n)),
                      [0] * n, [0] * n, [0] * n, 1, rows)
    wl, g = oracle.wa(d, 10.0, 100)
    assert wl == 0.0 and np.abs(g).max() == 0.0
    assert oracle.hpwl(d) == pytest.approx(980.0)


def test_filler_plan_matches_product_python():
    d = D.fixture_eplace(3000, 12)
    assert oracle.filler_plan(d)[:3] == D.filler_plan(d, 1.0)[:3]
    d2 = D.config_small(with_fillers=False)
    a, b = oracle.filler_plan(d2), D.filler_plan(d2, 1.0)
    assert a[0] == b[0] and a[1] == pytest.approx(b[1]) and a[3] == pytest.approx(b[3], rel=1e-9)
