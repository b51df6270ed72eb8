"""Abacus legalization (sp_abacus_legalize, superplace_amd.legalize) against
the reference's own legalizeCells (src/lg/abacus.cpp:172-284): the node
centres and AbacusResult it produced on the cases of
tests/golden/abacus_cases.py are in tests/golden/ref_abacus.npz
(make_abacus_golden.py).  The bar is bitwise: every centre, every count,
mean and max displacement.  Host code only."""
import os
import sys

import numpy as np
import pytest

from superplace_amd import hip, legalize

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
from abacus_cases import cases  # noqa: E402

KEYS = ["precondition_ok", "cells", "failed", "rows", "segments", "rounded_widths",
        "selfcheck_overlaps", "selfcheck_offsite", "disp_mean", "disp_max"]
CASES = cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_matches_reference(name):
    d, x, y, step, opt = CASES[name]
    g = np.load(os.path.join(GOLD, "ref_abacus.npz"), allow_pickle=False)
    lx, ly, res = legalize.legalize_cells(d, x, y, row_step=step,
                                          reference_trials=bool(opt & 1),
                                          area_weight=bool(opt & 2))
    want = g[f"{name}_res"]
    got = np.array([res[k] for k in KEYS], dtype=np.float64)
    assert np.array_equal(got, want), (got, want)
    assert np.array_equal(lx.view(np.uint32), g[f"{name}_x"].view(np.uint32))
    assert np.array_equal(ly.view(np.uint32), g[f"{name}_y"].view(np.uint32))


def test_legal_result_properties():
    """Whatever the reference says: cells on rows and sites, no overlaps
    inside a row, fixed nodes unmoved."""
    d, x, y, step, _ = CASES["synth20k"]
    lx, ly, res = legalize.legalize_cells(d, x, y, row_step=step)
    assert res["failed"] == 0 and res["selfcheck_overlaps"] == 0 and res["selfcheck_offsite"] == 0
    n = d.num_nodes
    fixed = (d.node_flags[:n] & 2) != 0
    assert np.array_equal(lx[fixed], x[fixed]) and np.array_equal(ly[fixed], y[fixed])
    cells = ~fixed
    bottoms = ly[cells] - 0.5 * d.node_h[:n][cells]
    assert np.isin(bottoms, d.row_ly).all()
    order = np.lexsort((lx[cells], bottoms))
    l = (lx[cells] - 0.5 * d.node_w[:n][cells])[order]
    r = (lx[cells] + 0.5 * d.node_w[:n][cells])[order]
    same_row = bottoms[order][1:] == bottoms[order][:-1]
    assert np.all(l[1:][same_row] >= r[:-1][same_row] - 1e-3)


def test_segments_and_errors():
    d, x, y, step, _ = CASES["mac_fixed"]
    row, lo, hi = legalize.segments(d, row_step=step)
    assert len(row) > d.num_rows and np.all(hi > lo)
    bad = np.array(step, np.float32).copy()
    bad[0] = 0.0
    with pytest.raises(hip.InvalidArgument):
        legalize.legalize_cells(d, x, y, row_step=bad)
    x2 = x.copy()
    x2[0] = np.nan
    with pytest.raises(hip.InvalidArgument):
        legalize.legalize_cells(d, x2, y, row_step=step)
