"""Filler plan and scatter (sp_filler_plan / sp_filler_scatter, through
design.filler_plan / design.add_fillers) against the reference's
initializeFillers and scatterFillers (src/gp/filler.cpp:33-139).

Golden fixtures (tests/golden/bookshelf/<case>_fillers.npz, made by
make_filler_golden.py) hold what the reference computed on the database its
own BookShelf reader built from the same files, for target densities 1.0,
0.8, 0.6 and seeds 1002, 7.  The bar is bitwise: counts, sizes and the three
FP64 areas equal, filler centres equal as float32 (long arrays through their
head, tail and SHA-256).  Host code only."""
import hashlib
import os

import numpy as np
import pytest

from superplace_amd import bookshelf, design as D, hip

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bookshelf")
CASES = {"tiny": "tiny/tiny.aux", "fixture": "fixture/fixture.aux", "thin1": "thin1/thin1.aux",
         "synth": "synth/synth.aux", "mac": "mac/mac.aux"}
TARGETS = (1.0, 0.8, 0.6)
SEEDS = (1002, 7)


def golden(case):
    with np.load(os.path.join(GOLD, f"{case}_fillers.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def same_positions(g, key, ours):
    ours = np.ascontiguousarray(ours, np.float32)
    ref = g[key]
    if key + "_sha" in g:
        assert ours.size == int(g[key + "_n"])
        half = ref.size // 2
        assert np.array_equal(ours[:half].view(np.uint32), ref[:half].view(np.uint32))
        assert np.array_equal(ours[-half:].view(np.uint32), ref[half:].view(np.uint32))
        assert hashlib.sha256(ours.tobytes()).hexdigest() == str(g[key + "_sha"])
    else:
        assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("case", sorted(CASES))
def test_plan_and_scatter_match_reference(case):
    g = golden(case)
    d = bookshelf.read(os.path.join(GOLD, CASES[case]))
    for t in TARGETS:
        for s in SEEDS:
            info = g[f"t{t}_s{s}_info"]
            count, w, h, white, scaled, area = D.filler_plan(d, t)
            assert count == int(info[0]), (t, s)
            assert np.float32(w) == np.float32(info[1]) and np.float32(h) == np.float32(info[2])
            assert (white, scaled, area) == (info[3], info[4], info[5]), (t, s)
            out = D.add_fillers(d, t, s)
            assert out.num_fillers == count
            n = out.num_nodes
            same_positions(g, f"t{t}_s{s}_x", out.node_x[n:])
            same_positions(g, f"t{t}_s{s}_y", out.node_y[n:])
            assert np.all(out.node_w[n:] == np.float32(w)) and np.all(out.node_h[n:] == np.float32(h))
            assert np.all(out.node_flags[n:] == D.F_FILLER)
            assert np.all(out.node2pin_start[n:] == d.num_pins)


@pytest.mark.parametrize("case", ["fixture", "mac"])
def test_rescatter_with_seed_offset(case):
    """scatterFillers(db, cfg, offset) draws with Config::random_seed + offset."""
    g = golden(case)
    d = bookshelf.read(os.path.join(GOLD, CASES[case]))
    out = D.add_fillers(d, 1.0, 1002 + 3)
    same_positions(g, "rescatter_x", out.node_x[out.num_nodes:])
    same_positions(g, "rescatter_y", out.node_y[out.num_nodes:])


def test_oracle_plan_agrees():
    """The oracle's restatement of the plan gives the same numbers."""
    import oracle
    for case in ("fixture", "mac", "synth"):
        d = bookshelf.read(os.path.join(GOLD, CASES[case]))
        for t in TARGETS:
            assert oracle.filler_plan(d, t) == D.filler_plan(d, t)


def test_fillers_inside_core_and_invariants():
    """Whatever the seed: fillers lie inside the core shrunk by half their
    size, and count x single area rounds the filler area."""
    d = bookshelf.read(os.path.join(GOLD, CASES["mac"]))
    lx, ly, hx, hy = d.core
    for s in (1, 2, 3):
        out = D.add_fillers(d, 0.9, s)
        n = out.num_nodes
        w, h = out.node_w[n], out.node_h[n]
        assert np.all(out.node_x[n:] >= np.float32(lx) + np.float32(0.5) * w)
        assert np.all(out.node_x[n:] <= np.float32(hx) - np.float32(0.5) * w)
        assert np.all(out.node_y[n:] >= np.float32(ly) + np.float32(0.5) * h)
        assert np.all(out.node_y[n:] <= np.float32(hy) - np.float32(0.5) * h)
    count, w, h, white, scaled, area = D.filler_plan(d, 0.9)
    assert abs(count - area / (float(np.float32(w)) * h)) <= 0.5


def test_rejects_bad_arguments():
    d = bookshelf.read(os.path.join(GOLD, CASES["tiny"]))
    with pytest.raises(hip.InvalidArgument):
        D.filler_plan(d, 0.0)
    with pytest.raises(hip.InvalidArgument):
        D.filler_plan(d, float("nan"))
