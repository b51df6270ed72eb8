"""The engine's own node order.  A design whose small nets join far-apart
node indices (here: the 2,000-cell synthetic design with its cells
shuffled) is renumbered inside the engine by breadth-first search over the
nets; positions, node indices and results still cross the C ABI in the
caller's order.  Shuffled and original must therefore agree: the same
evaluation at the same positions (the FP64 reductions sum in another order,
nothing else changes), and the same stage run."""
import numpy as np
import pytest

from superplace_amd import design as D
from superplace_amd import hip

pytestmark = pytest.mark.gpu


def shuffled(base, seed=5):
    nm = base.num_movable
    perm = np.concatenate([np.random.default_rng(seed).permutation(nm),
                           np.arange(nm, base.num_nodes)])          # new -> old
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    d = D.make_design(base.node_x[perm], base.node_y[perm], base.node_w[perm],
                      base.node_h[perm], base.node_flags[perm], nm, inv[base.pin2node],
                      base.pin2net, base.pin_offset_x, base.pin_offset_y, base.num_nets,
                      (base.row_ly, base.row_height, base.row_lx, base.row_hx), name="shuffled")
    return D.with_fillers(d), perm


@pytest.fixture(scope="module")
def pair():
    base = D.config_small(with_fillers=False)
    return D.with_fillers(base), *shuffled(base)


@pytest.mark.parametrize("prec", [32, 64])
def test_evaluation_is_independent_of_node_order(pair, prec):
    d0, d1, perm = pair
    cfg = hip.GpConfig(precision=prec, gp_bin_dim=128)
    e0, e1 = hip.GpEngine(d0, cfg), hip.GpEngine(d1, cfg)
    s0, o0 = e0.eval(lam=0.3, gamma=18.0, debug=True)
    s1, o1 = e1.eval(lam=0.3, gamma=18.0, debug=True)
    # integer density grids: equal whatever the order
    assert np.array_equal(o0["node_q"], o1["node_q"]) and np.array_equal(o0["filler_q"], o1["filler_q"])
    tol = 1e-12 if prec == 64 else 1e-6
    for k in ("tau", "wl", "hpwl"):
        assert s1[k] == pytest.approx(s0[k], rel=tol), k
    # the per-node WA gradient, mapped through each engine's opt index to
    # caller node indices of the original design
    nm = d0.num_movable
    i0, i1 = e0.opt_idx, e1.opt_idx
    g0 = dict(zip(i0.tolist(), o0["grad_wl"][:len(i0)]))
    orig1 = np.where(i1 < nm, perm[np.minimum(i1, nm - 1)], i1)   # shuffled index -> original
    g1 = dict(zip(orig1.tolist(), o1["grad_wl"][:len(i1)]))
    keys = [k for k in g0 if k < nm]
    a = np.array([g0[k] for k in keys])
    b = np.array([g1[k] for k in keys])
    assert np.max(np.abs(a - b)) <= (1e-12 if prec == 64 else 1e-6) * np.max(np.abs(a))


def test_stage_run_is_independent_of_node_order(pair):
    d0, d1, perm = pair
    cfg = hip.GpConfig(precision=64, gp_bin_dim=128, gp_max_iter=40)
    e0, e1 = hip.GpEngine(d0, cfg), hip.GpEngine(d1, cfg)
    r0, _ = e0.run()
    r1, _ = e1.run()
    assert r1["iterations"] == r0["iterations"]
    assert r1["hpwl"] == pytest.approx(r0["hpwl"], rel=1e-6)
    x0, y0 = e0.positions()
    x1, y1 = e1.positions()
    nm = d0.num_movable
    # caller order: shuffled node j is original node perm[j]
    dx = np.abs(x1[:nm] - x0[perm[:nm]]).max()
    dy = np.abs(y1[:nm] - y0[perm[:nm]]).max()
    assert max(dx, dy) <= 1e-3, (dx, dy)
