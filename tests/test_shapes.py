"""The launch-shape table of the library (sp_hip_shapes_*), on the CPU: the
compiled-in table mirrors the autotuner's last output
(superplace_amd/shapes.json), overrides take precedence by size then fall
back to the size-0 entry, reset restores the compiled-in table, and invalid
shapes are refused."""
import json
import os

import pytest

from superplace_amd import hip, tune


def test_compiled_table_mirrors_shapes_json():
    tune.reset_shapes()
    if not os.path.exists(tune.SHAPES_JSON):
        pytest.skip("no tuned table yet")
    with open(tune.SHAPES_JSON) as f:
        table = json.load(f)
    for e in table["entries"]:
        assert tune.get_shapes(e["dim"]) == e["shapes"], e["dim"]


def test_override_precedence_and_reset():
    tune.reset_shapes()
    base = tune.get_shapes(4096)   # a size no table holds: the size-0 entry
    assert base == tune.DEFAULT
    own = dict(tune.DEFAULT, scatter_chunk=1024, graph_slots=16)
    tune.set_shapes(4096, own)
    assert tune.get_shapes(4096) == own
    every = dict(tune.DEFAULT, sort_period=4)
    tune.set_shapes(0, every)
    assert tune.get_shapes(4096) == own and tune.get_shapes(2048) == every
    tune.reset_shapes()
    assert tune.get_shapes(4096) == base


@pytest.mark.parametrize("bad", [dict(scatter_chunk=300), dict(sort_period=0),
                                 dict(col_pairs=3), dict(graph_slots=65)])
def test_invalid_shapes_refused(bad):
    with pytest.raises(hip.InvalidArgument):
        tune.set_shapes(128, dict(tune.DEFAULT, **bad))
