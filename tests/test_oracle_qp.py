"""Oracle of the quadratic initial placement (QuadraticPlacer,
src/init/quadratic_placer.cpp; Bound2Bound model, src/init/b2b_model.h)
pinned against the reference's own computeNetBoundPins and forEachB2BPair
(golden pairs in tests/golden/ref_b2b_*.npz, made by make_b2b_golden.py),
the B2B gradient identity, and a direct sparse solve."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle
from superplace_amd import design as D

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cases():
    return {"fixture": D.fixture_eplace(300, 2, seed=3),
            "synth": D.synth_design(1500, 5, 1600, 420.0, 0.7, seed=21, name="synth",
                                    with_fillers=False)}


def system_matrix(diag, rs, col, val):
    n = len(diag)
    rows = np.repeat(np.arange(n), np.diff(rs))
    return (sp.csr_matrix((val, (rows, col)), shape=(n, n)) + sp.diags(diag)).tocsr()


@pytest.mark.parametrize("name", ["fixture", "synth"])
def test_bound_pins_and_system_match_reference(name):
    d = cases()[name]
    g = np.load(os.path.join(GOLD, f"ref_b2b_{name}.npz"), allow_pickle=False)
    nm = d.num_movable
    wts = g["net_weight"]
    for pl in ("centre", "scatter"):
        x, y = g[f"x_{pl}"], g[f"y_{pl}"]
        bounds = np.stack(oracle.bound_pins(d, x, y))
        assert np.array_equal(bounds, g[f"bounds_{pl}"]), pl
        for axis in (0, 1):
            pos = (x, y)[axis]
            off = (d.pin_offset_x, d.pin_offset_y)[axis].astype(np.float32)
            pq, w = g[f"pairs_{pl}_{axis}"], g[f"w_{pl}_{axis}"].astype(np.float64)
            # assemble (quadratic_placer.cpp:75-98) from the reference's pairs
            A = np.zeros((nm, nm))
            b = np.zeros(nm)
            for (p, q), wt in zip(pq.T, w):
                i, j = d.pin2node[p], d.pin2node[q]
                pi = np.float64(np.float32(pos[i]) + off[p])
                pj = np.float64(np.float32(pos[j]) + off[q])
                if i < nm and j < nm:
                    A[i, i] += wt; A[j, j] += wt; A[i, j] -= wt; A[j, i] -= wt
                    b[i] += -wt * (np.float64(off[p]) - off[q])
                    b[j] += -wt * (np.float64(off[q]) - off[p])
                elif i < nm:
                    A[i, i] += wt
                    b[i] += wt * (pj - off[p])
                elif j < nm:
                    A[j, j] += wt
                    b[j] += wt * (pi - off[q])
            zero = np.diag(A) == 0
            A[zero, zero] = 1e-6
            b[zero] += 1e-6 * pos[:nm][zero]
            diag, rhs, rs, col, val = oracle.b2b_system(d, x, y, axis, net_weight=wts)
            M = system_matrix(diag, rs, col, val).toarray()
            scale = np.abs(A).max()
            assert np.abs(M - A).max() <= 1e-12 * scale, (pl, axis)
            # summation order differs (row order here, pair order there): rounding only
            assert np.abs(rhs - b).max() <= 1e-10 * max(np.abs(b).max(), 1.0), (pl, axis)
            assert (A == A.T).all()


def test_gradient_identity():
    """At distinct pin positions with a negligible distance floor, the B2B
    quadratic's gradient A x - b equals the HPWL subgradient: per net +w for
    the pin holding the maximum, -w for the minimum, 0 for the others
    (b2b_model.h:50-55: the 2P-3 pairs telescope to the net's span)."""
    d = D.fixture_eplace(200, 0, seed=9)
    rng = np.random.default_rng(4)
    n = d.num_nodes
    x = d.node_x[:n].astype(np.float32).copy()
    y = d.node_y[:n].astype(np.float32).copy()
    x[:d.num_movable] = rng.uniform(d.core[0], d.core[2], d.num_movable)
    y[:d.num_movable] = rng.uniform(d.core[1], d.core[3], d.num_movable)
    wts = rng.uniform(0.5, 2.0, d.num_nets).astype(np.float32)
    for axis, pos, off in ((0, x, d.pin_offset_x), (1, y, d.pin_offset_y)):
        diag, rhs, rs, col, val = oracle.b2b_system(d, x, y, axis, net_weight=wts,
                                                   min_distance=1e-9)
        A = system_matrix(diag, rs, col, val)
        grad = A @ pos[:d.num_movable].astype(np.float64) - rhs
        want = np.zeros(d.num_movable)
        bmin, bmax, tmin, tmax = oracle.bound_pins(d, x, y)
        if axis == 1:
            bmin, bmax = tmin, tmax
        for k in range(d.num_nets):
            pins = d.flat_net2pin[d.net2pin_start[k]:d.net2pin_start[k + 1]]
            nodes = d.pin2node[pins]
            if len(pins) < 2 or len(set(nodes)) < len(nodes):
                continue   # pins sharing a node pair with nothing (b2b_model.h:66)
            if d.pin2node[bmax[k]] < d.num_movable:
                want[d.pin2node[bmax[k]]] += wts[k]
            if d.pin2node[bmin[k]] < d.num_movable:
                want[d.pin2node[bmin[k]]] -= wts[k]
        shared = np.zeros(d.num_movable, bool)
        for k in range(d.num_nets):
            pins = d.flat_net2pin[d.net2pin_start[k]:d.net2pin_start[k + 1]]
            nodes = d.pin2node[pins]
            if len(set(nodes)) < len(nodes):
                shared[nodes[nodes < d.num_movable]] = True
        ok = ~shared
        assert ok.sum() > 100
        assert np.abs(grad[ok] - want[ok]).max() < 1e-6 * max(1.0, np.abs(want).max()), axis


def test_bicgstab_matches_direct_solve():
    # every movable component of this design reaches a fixed node, so the
    # system is nonsingular (the synthetic one holds a free two-cell component)
    d = cases()["fixture"]
    diag, rhs, rs, col, val = oracle.b2b_system(d, d.node_x, d.node_y, 0)
    A = system_matrix(diag, rs, col, val)
    xs = sla.spsolve(A.tocsc(), rhs)
    assert np.all(np.isfinite(xs))
    xb, err, its = oracle.bicgstab(diag, rhs, rs, col, val, np.zeros(len(diag)), tol=1e-13,
                                   max_iter=20000)
    assert err <= 1e-13 and its > 10
    assert np.abs(xb - xs).max() <= 1e-9 * np.abs(xs).max()
    # the iteration cap bounds the work, as qp_solver_max_iter does
    _, err5, its5 = oracle.bicgstab(diag, rhs, rs, col, val, np.zeros(len(diag)), tol=1e-13,
                                    max_iter=5)
    assert its5 == 5 and err5 > err


def test_qp_place_stop_rule_and_quality():
    d = cases()["synth"]
    cfg = oracle.qp_config()
    x, y, hp, er = oracle.qp_place(d, cfg)
    n = len(hp)
    assert 1 <= n <= cfg.max_iter
    # stops at the first iteration > 4 whose HPWL moved by < outer_tol, or at max_iter
    for i in range(5, n - 1):
        assert abs(hp[i] - hp[i - 1]) >= cfg.outer_tol * hp[i - 1]
    assert n == cfg.max_iter or abs(hp[-1] - hp[-2]) < cfg.outer_tol * hp[-2]
    # fixed nodes unmoved, movable nodes inside the core
    assert np.array_equal(x[d.num_movable:], d.node_x[d.num_movable:d.num_nodes])
    m = d.num_movable
    assert np.all(x[:m] - 0.5 * d.node_w[:m] >= d.core[0] - 1e-3)
    assert np.all(x[:m] + 0.5 * d.node_w[:m] <= d.core[2] + 1e-3)
    # far better than a random placement of the same cells
    rng = np.random.default_rng(1)
    xr, yr = d.node_x[:d.num_nodes].copy(), d.node_y[:d.num_nodes].copy()
    xr[:m] = rng.uniform(d.core[0], d.core[2], m)
    yr[:m] = rng.uniform(d.core[1], d.core[3], m)
    import dataclasses
    dr = dataclasses.replace(d, node_x=xr.astype(np.float32), node_y=yr.astype(np.float32))
    assert hp[-1] < 0.5 * oracle.hpwl(dr, 64)
