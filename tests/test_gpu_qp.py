"""GPU quadratic initial placement (sp_qp_place, sp_qp_b2b_system) against
the oracle (oracle.qp_place / b2b_system, itself pinned against the
reference's bound pins and B2B pairs in tests/test_oracle_qp.py).

Both sides do the same operations in the same order: the systems are
assembled in the same row order, nothing is fused into multiply-adds, and
the oracle sums its dot products and HPWL in the fixed order of the GPU's
block reduction.  So the whole trajectory (every outer iteration's HPWL and
solver error, the final positions) is compared bit for bit."""
import dataclasses
import os

import numpy as np
import pytest

import oracle
from superplace_amd import design as D, hip, qp

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cases():
    return {"fixture": D.fixture_eplace(300, 2, seed=3),
            "synth": D.synth_design(1500, 5, 1600, 420.0, 0.7, seed=21, name="synth",
                                    with_fillers=False)}


@pytest.mark.parametrize("name", ["fixture", "synth"])
def test_b2b_system_bitwise(name):
    d = cases()[name]
    g = np.load(os.path.join(GOLD, f"ref_b2b_{name}.npz"), allow_pickle=False)
    w = g["net_weight"]
    for pl in ("centre", "scatter"):
        x, y = g[f"x_{pl}"], g[f"y_{pl}"]
        dd = dataclasses.replace(d, node_x=x, node_y=y)
        for axis in (0, 1):
            print(name, pl, axis, flush=True)
            got = qp.b2b_system(dd, axis, net_weight=w)
            ref = oracle.b2b_system(d, x, y, axis, net_weight=w)
            for a, b, what in zip(got, ref, ("diag", "rhs", "row_start", "col", "val")):
                assert a.shape == b.shape, what
                assert np.array_equal(a, b), (pl, axis, what, np.abs(a - b).max())


def anchored_cases():
    # every connected group of movable nodes reaches a fixed node (the
    # synthetic design of cases() holds a free two-cell group: singular system)
    return {"fixture": D.fixture_eplace(300, 2, seed=3),
            "synth": D.synth_design(1500, 8, 1700, 420.0, 0.7, seed=22, name="synth",
                                    with_fillers=False),
            "fixture2k": D.fixture_eplace(2000, 6, seed=5)}


@pytest.mark.parametrize("name", ["fixture", "synth", "fixture2k", "synth_free"])
def test_qp_place_matches_oracle(name):
    d = anchored_cases()[name] if name != "synth_free" else cases()["synth"]
    print(name, flush=True)
    x, y, rows, its = qp.qp_place(d)
    ox, oy, ohp, oer = oracle.qp_place(d)
    assert its == len(ohp)
    assert np.array_equal(np.array([r["hpwl"] for r in rows]), ohp)
    assert np.array_equal(np.array([r["err"] for r in rows]), oer)
    assert np.array_equal(x, ox) and np.array_equal(y, oy)
    assert np.array_equal(x[d.num_movable:], d.node_x[d.num_movable:d.num_nodes])


def test_qp_free_group_quality():
    """With a group of cells tied to nothing fixed (singular system), where
    the group ends up depends on rounding, as for the reference on another
    machine; the placement quality does not."""
    d = cases()["synth"]
    _, _, rows, _ = qp.qp_place(d)
    _, _, ohp, _ = oracle.qp_place(d)
    assert abs(rows[-1]["hpwl"] - ohp[-1]) < 0.01 * ohp[-1]


def test_qp_place_deterministic_and_weighted():
    d = cases()["fixture"]
    w = (1.0 + (np.arange(d.num_nets) % 4) * 0.25).astype(np.float32)
    a = qp.qp_place(d, net_weight=w)
    b = qp.qp_place(d, net_weight=w)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    ox, oy, ohp, _ = oracle.qp_place(d, net_weight=w)
    assert np.allclose([r["hpwl"] for r in a[2]], ohp, rtol=1e-9)


def test_qp_solver_cap_and_config_errors():
    d = cases()["fixture"]
    x1, y1, rows1, its1 = qp.qp_place(d, qp.QpConfig(max_iter=1, solver_max_iter=0))
    # no solver iterations: every movable node stays at the core centre
    cx = np.float32(0.5) * (np.float32(d.core[0]) + np.float32(d.core[2]))
    assert its1 == 1 and np.all(x1[:d.num_movable] == cx)
    with pytest.raises(hip.InvalidArgument):
        qp.qp_place(d, qp.QpConfig(max_iter=0))
    with pytest.raises(hip.InvalidArgument):
        qp.b2b_system(d, 2)
