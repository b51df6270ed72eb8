"""The stage flow's host parts (superplace_amd.flow) and the legality of
what the legalizers deliver, judged by the reference's own standalone
checker (tools/check/main.cpp built into oracle/_ref as superplace-check):
Abacus after jitter, and mLG followed by Abacus on a design with movable
macros, pass it; the unlegalized placements fail it.  Host code only."""
import dataclasses
import json
import os

import numpy as np
import pytest

import oracle
from superplace_amd import bookshelf, flow, legalize
from superplace_amd.design import F_FIXED, F_MACRO

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bookshelf")


def need_checker():
    if not oracle.CHECKER_PATH.exists():
        pytest.skip("superplace-check is not built (oracle/_ref)")


def aux(name):
    return os.path.join(GOLD, name, f"{name}.aux")


def jitter(d, seed, sx, sy, movable_only=True):
    g = np.random.default_rng(seed)
    n = d.num_nodes
    x, y = d.node_x.copy(), d.node_y.copy()
    m = np.arange(n) < d.num_movable
    m &= (d.node_flags[:n] & F_FIXED) == 0
    x[:n][m] += g.normal(0, sx, m.sum()).astype(np.float32)
    y[:n][m] += g.normal(0, sy, m.sum()).astype(np.float32)
    lx, ly, hx, hy = d.core
    return dataclasses.replace(d, node_x=np.clip(x, lx, hx).astype(np.float32),
                               node_y=np.clip(y, ly, hy).astype(np.float32))


@pytest.mark.parametrize("name", ["synth", "thin1"])
def test_abacus_output_passes_checker(name, tmp_path):
    need_checker()
    d = jitter(bookshelf.read(aux(name)), 1, 4.0, 3.0)
    before = tmp_path / "before.pl"
    bookshelf.write_pl(str(before), d)
    ok, report = oracle.check_placement(aux(name), before)
    assert not ok, report
    out, res = flow.lg_stage(d, flow.FlowConfig())
    assert res["failed"] == 0
    pl = tmp_path / "lg.pl"
    bookshelf.write_pl(str(pl), out)
    ok, report = oracle.check_placement(aux(name), pl)
    assert ok, report


def test_macro_and_cell_legalization_pass_checker(tmp_path):
    """mLG (macros pushed onto each other) then Abacus; the .pl carries the
    frozen macros' /FIXED and the flipped macros' orientations."""
    need_checker()
    d = bookshelf.read(aux("mac"))
    m = np.nonzero(((d.node_flags & F_MACRO) != 0)[:d.num_movable])[0]
    x, y = d.node_x.copy(), d.node_y.copy()
    x[m] = np.float32(0.5 * (d.core[0] + d.core[2])) + np.arange(len(m)) * 23.0
    y[m] = np.float32(0.5 * (d.core[1] + d.core[3])) + np.arange(len(m)) * 13.0
    d = jitter(dataclasses.replace(d, node_x=x, node_y=y), 2, 10.0, 8.0)
    same, res = flow.lg_stage(d, flow.FlowConfig())   # macros not legalized yet: no-op
    assert res is None and same is d
    o = legalize.legalize_macros(d)
    assert o["result"]["overlap_after"] == 0.0
    n = d.num_nodes
    d2 = d.copy()
    d2.node_x[:n], d2.node_y[:n] = o["x"], o["y"]
    d2.node_flags = o["flags"].copy()
    d2.pin_offset_x = o["pin_offset_x"]
    d2.meta = dict(d.meta, node_orient=o["orient"])
    out, res = flow.lg_stage(d2, flow.FlowConfig())
    assert res["failed"] == 0
    pl = tmp_path / "mac.lg.pl"
    bookshelf.write_pl(str(pl), out)
    ok, report = oracle.check_placement(aux("mac"), pl)
    assert ok, report
    names = d.meta["node_names"]
    text = pl.read_text().splitlines()
    codes = ["N", "E", "S", "W", "FN", "FE", "FS", "FW"]
    for i in m:
        line = next(t for t in text if t.split("\t")[0] == names[i])
        assert line.endswith(" /FIXED")
        assert line.split(": ")[1].split()[0] == codes[o["orient"][i]]


def test_hpwl_matches_reference_definition():
    """flow.hpwl (summary figures) against the oracle's chunked HPWL."""
    d = bookshelf.read(aux("synth"))
    a = flow.hpwl(d)
    b = oracle.hpwl(d, precision=32)
    assert abs(a - b) <= 1e-9 * max(1.0, b)


def test_failure_summary_contract(tmp_path):
    """main.cpp:119-138: a failing run writes summary.json with run_status
    failed, failure_message, failure_kind (the invariant's, else
    "exception") and failure_stage (the invariant's, else "pipeline")."""
    cfg = flow.FlowConfig(stages=("bogus",))
    with pytest.raises(ValueError):
        flow.run(aux("fixture"), str(tmp_path), cfg)
    s = json.loads((tmp_path / "fixture" / "summary.json").read_text())
    assert s["run_status"] == "failed" and s["failure_kind"] == "exception"
    assert s["failure_stage"] == "pipeline" and "bogus" in s["failure_message"]
    assert s["benchmark"] == "fixture" and "read" in s["durations_ms"]
    # lg before gp on movable macros: a logged no-op, the run completes
    d, path = flow.run(aux("mac"), str(tmp_path), flow.FlowConfig(stages=("lg",)))
    s = json.loads(open(path).read())
    assert s["run_status"] == "completed" and "lg_cells" not in s
    assert (tmp_path / "mac" / "mac.lg.pl").exists()


def test_mixed_design_legalizers_pass_checker(tmp_path):
    """The mixed-size generator (63 movable macros, rows of its site pitch)
    written as BookShelf and read back: mLG from the clustered start, then
    Abacus, pass the checker."""
    need_checker()
    from superplace_amd import design as D
    g = D.config_mixed(with_fillers=False)
    path = bookshelf.write_bookshelf(g, str(tmp_path), "mixed", site=g.meta["site_pitch"])
    d = bookshelf.read(path)
    assert d.num_movable == 2063 and int(((d.node_flags & F_MACRO) != 0).sum()) >= 63
    o = legalize.legalize_macros(d)
    n = d.num_nodes
    d2 = d.copy()
    d2.node_x[:n], d2.node_y[:n] = o["x"], o["y"]
    d2.node_flags = o["flags"].copy()
    d2.pin_offset_x = o["pin_offset_x"]
    d2.meta = dict(d.meta, node_orient=o["orient"])
    out, res = flow.lg_stage(d2, flow.FlowConfig())
    assert res["failed"] == 0 and res["rounded_widths"] == 0
    pl = tmp_path / "mixed.lg.pl"
    bookshelf.write_pl(str(pl), out)
    ok, report = oracle.check_placement(path, pl)
    assert ok, report
