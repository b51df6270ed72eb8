"""Launch shapes (sp_hip_shapes) change timings only: a stage run under every
candidate value of each knob, and a batch whose members run different
shapes, give the results and positions of the default shapes bit for bit."""
import numpy as np
import pytest

from superplace_amd import design as D
from superplace_amd import hip, tune

pytestmark = pytest.mark.gpu

KEYS = ("iterations", "evaluations", "converged", "stagnated", "hpwl", "overflow")


def run(d, dim, shapes):
    tune.set_shapes(dim, shapes)
    try:
        eng = hip.GpEngine(d, hip.GpConfig(gp_bin_dim=dim))
        taken = eng.info["shapes"].as_dict()
        res, _ = eng.run(max_rows=1)
        x, y = eng.positions()
        eng.close()
    finally:
        tune.reset_shapes()
    return taken, res, x, y


@pytest.mark.parametrize("knob", sorted(tune.KNOBS))
def test_every_knob_value_gives_the_same_bits(knob):
    d = D.with_fillers(D.fixture_eplace(3000, 12))
    _, ref, rx, ry = run(d, 128, tune.DEFAULT)
    assert ref["converged"]
    for v in tune.KNOBS[knob]:
        shapes = dict(tune.DEFAULT, **{knob: v})
        taken, res, x, y = run(d, 128, shapes)
        print(knob, v, taken, res["iterations"], flush=True)
        assert taken == shapes
        for k in KEYS:
            assert res[k] == ref[k], (knob, v, k)
        assert np.array_equal(x, rx) and np.array_equal(y, ry), (knob, v)


def test_batch_members_with_different_shapes():
    designs = [D.with_fillers(D.fixture_eplace(600, 2)), D.with_fillers(D.fixture_eplace(3000, 12))]
    cfgs = [hip.GpConfig(gp_bin_dim=64), hip.GpConfig(gp_bin_dim=128)]
    alone = []
    for d, c in zip(designs, cfgs):
        eng = hip.GpEngine(d, c)
        res, _ = eng.run(max_rows=1)
        alone.append((res, *eng.positions()))
        eng.close()
    tune.set_shapes(64, dict(tune.DEFAULT, scatter_chunk=256, sort_period=1))
    tune.set_shapes(128, dict(tune.DEFAULT, scatter_chunk=2048, sort_period=4, graph_slots=4))
    try:
        res, xs, ys = hip.run_batch(designs, cfgs)
    finally:
        tune.reset_shapes()
    for (r, x, y), b, bx, by in zip(alone, res, xs, ys):
        for k in KEYS:
            assert b[k] == r[k], k
        assert np.array_equal(bx, x) and np.array_equal(by, y)
