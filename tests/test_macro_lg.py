"""Macro legalization (sp_macro_legalize, superplace_amd.legalize) against
the reference's own legalizeMacros (src/lg/macro_sa.cpp:493-1272): what it
produced on the cases of tests/golden/mlg_cases.py is in
tests/golden/ref_mlg.npz and tests/golden/mlg/<case>.json
(make_mlg_golden.py).  The bar is bitwise: every centre, flag, pin offset,
orientation, MacroSaResult field (the wall-clock saMs aside) and summary
key.  The cases include the node layouts of the reference's own tests
(test/test_macro_sa.cpp); their assertions are restated below on this
library's output.  Host code only."""
import dataclasses
import json
import os
import sys
import tempfile

import numpy as np
import pytest

from superplace_amd import hip, legalize, metrics
from superplace_amd.design import F_FIXED, F_MACRO

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
from mlg_cases import cases  # noqa: E402

RES_KEYS = ["num_macros", "moves", "accepted", "repaired", "chained", "hot_macros",
            "overlap_before", "overlap_after_sa", "overlap_after", "sa_ms", "t0",
            "typical_worsen", "hpwl_before", "hpwl_after"]
CASES = cases()


def run(name, **over):
    d, cfg = CASES[name]
    cfg = dict(cfg, **over)
    sink = metrics.MetricsSink()
    out = legalize.legalize_macros(d, sink=sink, **cfg)
    with tempfile.TemporaryDirectory() as t:
        path = os.path.join(t, "summary.json")
        sink.write_summary(path)
        with open(path) as f:
            out["summary"] = json.load(f)
    return d, out


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("name", sorted(CASES))
def test_matches_reference(name):
    d, o = run(name)
    g = np.load(os.path.join(GOLD, "ref_mlg.npz"), allow_pickle=False)
    n = d.num_nodes
    got = np.array([o["result"][k] for k in RES_KEYS], np.float64)
    got[RES_KEYS.index("sa_ms")] = 0.0
    assert np.array_equal(got, g[f"{name}_res"]), dict(zip(RES_KEYS, zip(got, g[f"{name}_res"])))
    assert np.array_equal(bits(o["x"]), bits(g[f"{name}_x"]))
    assert np.array_equal(bits(o["y"]), bits(g[f"{name}_y"]))
    assert np.array_equal(o["flags"][:n], g[f"{name}_flags"])
    assert np.array_equal(bits(o["pin_offset_x"]), bits(g[f"{name}_offx"]))
    assert np.array_equal(o["orient"][:n], g[f"{name}_orient"])
    summary = dict(o["summary"])
    durations = summary.pop("durations_ms", {})
    with open(os.path.join(GOLD, "mlg", f"{name}.json")) as f:
        assert summary == json.load(f)
    assert ("mLG" in durations) == (o["result"]["num_macros"] > 0)


# ---------------------------------------------------------------- properties
def boxes(d, x, y, idx):
    hw, hh = 0.5 * d.node_w[idx], 0.5 * d.node_h[idx]
    return np.stack([x[idx] - hw, y[idx] - hh, x[idx] + hw, y[idx] + hh], 1)


def pair_overlap(a, b):
    dx = np.minimum(a[:, None, 2], b[None, :, 2]) - np.maximum(a[:, None, 0], b[None, :, 0])
    dy = np.minimum(a[:, None, 3], b[None, :, 3]) - np.maximum(a[:, None, 1], b[None, :, 1])
    return np.where((dx > 0) & (dy > 0), dx.astype(np.float64) * dy, 0.0)


def macro_overlap(d, x, y, with_blocks=True):
    """Independent pairwise overlap of the macros (and of macros with the
    fixed nodes of positive area, terminal_NI excluded)."""
    n = d.num_nodes
    fl = d.node_flags[:n]
    m = np.nonzero((fl & F_MACRO) != 0)[0]
    m = m[m < d.num_movable] if not with_blocks else m
    mov = np.nonzero(((fl & F_MACRO) != 0) & (np.arange(n) < d.num_movable))[0]
    a = boxes(d, x, y, mov)
    o = pair_overlap(a, a)
    total = float(np.triu(o, 1).sum())
    if with_blocks:
        blk = np.arange(d.num_movable, n)
        blk = blk[((fl[blk] & 8) == 0) & (d.node_w[blk] > 0) & (d.node_h[blk] > 0)]
        if len(blk):
            total += float(pair_overlap(a, boxes(d, x, y, blk)).sum())
    return total


def macro_indices(d):
    n = d.num_movable
    return np.nonzero(((d.node_flags[:n] & F_MACRO) != 0) & ((d.node_flags[:n] & F_FIXED) == 0))[0]


@pytest.mark.parametrize("name", ["spec_stacked16", "dense150", "dense400", "sparse3000"])
def test_result_is_legal(name):
    """Zero overlap (independent pairwise check), macros inside the core on
    the site/row grid, fixed afterwards, every other node where it was,
    flipped macros with negated pin x offsets and mirrored orientation."""
    d, o = run(name)
    x, y = o["x"], o["y"]
    n = d.num_nodes
    m = macro_indices(d)
    assert macro_overlap(d, x, y) == 0.0
    assert o["result"]["overlap_after"] == 0.0
    lx, ly, hx, hy = d.core
    b = boxes(d, x, y, m)
    assert np.all(b[:, 0] >= lx - 1e-3) and np.all(b[:, 2] <= hx + 1e-3)
    assert np.all(b[:, 1] >= ly - 1e-3) and np.all(b[:, 3] <= hy + 1e-3)
    rh = d.row_height_mode
    site = float(d.meta["row_step"][0])
    assert np.all(np.abs(np.remainder(b[:, 1] - ly + 0.5 * rh, rh) - 0.5 * rh) < 1e-2)
    assert np.all(np.abs(np.remainder(b[:, 0] - lx + 0.5 * site, site) - 0.5 * site) < 1e-2)
    assert np.all((o["flags"][m] & F_FIXED) != 0)
    others = np.setdiff1d(np.arange(n), m)
    assert np.array_equal(x[others], d.node_x[others]) and np.array_equal(y[others], d.node_y[others])
    assert np.array_equal(o["flags"][others], d.node_flags[others])
    flipped = np.zeros(n, bool)
    for node in m:
        pins = d.flat_node2pin[d.node2pin_start[node]:d.node2pin_start[node + 1]]
        if len(pins) and np.array_equal(o["pin_offset_x"][pins], -d.pin_offset_x[pins]) and \
                np.any(d.pin_offset_x[pins] != 0):
            flipped[node] = True
    o0 = d.meta["node_orient"][:n]
    mirror = {0: 4, 4: 0, 2: 6, 6: 2}
    want = np.array([mirror.get(int(v), int(v)) if f else int(v) for v, f in zip(o0, flipped)])
    assert np.array_equal(o["orient"][:n], want)


def test_removes_artificial_overlap():
    """testRemovesArtificialOverlap: 16 macros on one point, fixed corners."""
    d, o = run("spec_stacked16")
    assert o["result"]["num_macros"] == 16
    assert macro_overlap(d, d.node_x, d.node_y, with_blocks=False) > 0
    assert macro_overlap(d, o["x"], o["y"], with_blocks=False) == 0.0
    assert o["result"]["overlap_after"] == 0.0


def test_macros_become_fixed_and_second_pass_is_empty():
    """testMacrosBecomeFixed: all 8 fixed, still macros; a second pass
    finds no movable macro."""
    d, o = run("spec_fixed8")
    assert np.all((o["flags"] & F_FIXED) != 0) and np.all((o["flags"] & F_MACRO) != 0)
    d2 = dataclasses.replace(d, node_flags=o["flags"].copy(), node_x=o["x"].copy(),
                             node_y=o["y"].copy())
    r2 = legalize.legalize_macros(d2)["result"]
    assert r2["num_macros"] == 0 and r2["moves"] == 0


def test_wirelength_controlled():
    """testWirelengthControlled: a legal start stays legal, HPWL within 5%."""
    d, o = run("spec_grid9")
    r = o["result"]
    assert r["overlap_before"] == 0.0 and r["overlap_after"] == 0.0
    assert r["hpwl_after"] / r["hpwl_before"] < 1.05


def test_reported_hpwl_matches_full_recompute():
    """testIncrementalMatchesFullHpwl: hpwl_after equals a full HPWL of the
    output (mirrored pins included), computed here in float64 per net."""
    d, o = run("spec_static")
    w = d.meta.get("net_weight")
    px = o["x"][d.pin2node] + o["pin_offset_x"]
    py = o["y"][d.pin2node] + d.pin_offset_y
    total = 0.0
    for k in range(d.num_nets):
        p = d.flat_net2pin[d.net2pin_start[k]:d.net2pin_start[k + 1]]
        if len(p) >= 2:
            total += float(px[p].max() - px[p].min()) + float(py[p].max() - py[p].min())
    assert w is None
    assert abs(o["result"]["hpwl_after"] - total) <= 1e-6 * max(1.0, total)


def test_push_chain():
    """testPushChain: the column is pushed down 30 as a whole."""
    d, o = run("spec_chain")
    r = o["result"]
    assert r["overlap_before"] > 0 and r["overlap_after"] == 0.0 and r["chained"] >= 4
    disp = np.hypot(o["x"][:5] - d.node_x[:5], o["y"][:5] - d.node_y[:5])
    assert disp.max() <= 30.0 + 1e-3


def test_escape_blockage_array():
    """testEscapeBlockageArray: the nearest zero-overlap contact position is
    1010 away (1100 with grid candidates only)."""
    d, o = run("spec_array")
    assert o["result"]["overlap_before"] > 0 and o["result"]["overlap_after"] == 0.0
    disp = float(np.hypot(np.float64(o["x"][0]) - d.node_x[0], np.float64(o["y"][0]) - d.node_y[0]))
    assert abs(disp - 1010.0) <= 1e-3


def test_separation_tie_prefers_small_move():
    """testSeparationTiePrefersSmallDisplacement: where SA left overlap, the
    wide macro moves at most 10 (vertically), not across the core."""
    exercised = 0
    for seed in (1002, 1, 2, 3, 4):
        d, o = run(f"spec_tie{seed}")
        assert o["result"]["overlap_after"] == 0.0
        if o["result"]["overlap_after_sa"] > 0:
            exercised += 1
            assert np.hypot(o["x"][0] - 5000.0, o["y"][0] - 5000.0) <= 10.0 + 1e-3
    assert exercised > 0


def test_large_macro_small_shift():
    """testLargeMacroSmallShift: a 30-deep overlap is undone by a shift of
    at most 60."""
    d, o = run("spec_shift")
    assert o["result"]["overlap_after"] == 0.0
    disp = np.hypot(o["x"][:3] - d.node_x[:3], o["y"][:3] - d.node_y[:3])
    assert disp.max() <= 60.0


def test_determinism():
    """testDeterminism: the same seed gives the same bits."""
    _, a = run("spec_determinism")
    _, b = run("spec_determinism")
    assert np.array_equal(bits(a["x"]), bits(b["x"])) and np.array_equal(bits(a["y"]), bits(b["y"]))
    _, c = run("dense60", seed=5)
    _, e = run("dense60")
    assert not np.array_equal(bits(c["x"]), bits(e["x"]))


def test_errors():
    d, cfg = CASES["fixture6"]
    with pytest.raises(hip.InvalidArgument):
        legalize.legalize_macros(d, row_step=np.zeros(d.num_rows, np.float32))
