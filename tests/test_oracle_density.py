"""Pins the oracle's bin grid (bin_grid.cpp restated): the hand cases of
test/test_density.cpp:60-246, a brute-force FP64 overlap on a tiny grid, the
overfull tau floor of test/test_eplace.cpp:115-122 and the fixed-point mirror."""
import numpy as np
import pytest

import oracle
from superplace_amd import design as D
from tests.helpers import density_db

M, FX, FL = D.F_MACRO, D.F_FIXED, D.F_FILLER


@pytest.mark.parametrize("prec", [32, 64])
def test_scaling_rule(prec):
    dc = density_db(64.0, [(32, 32, 16, 16, 0)])
    dm = density_db(64.0, [(32, 32, 16, 16, M)])
    gc = oracle.Grid(dc, 0.5, 8, prec)
    gm = oracle.Grid(dm, 0.5, 8, prec)
    assert abs(gc.accumulate(dc.node_x, dc.node_y)[0].sum() - 256.0) < 1e-3
    assert abs(gm.accumulate(dm.node_x, dm.node_y)[0].sum() - 128.0) < 1e-3


def test_filler_separate_and_smoothing():
    d = density_db(64.0, [(32, 32, 16, 16, FL)])
    g = oracle.Grid(d, 0.5, 8)
    nd, fd = g.accumulate(d.node_x, d.node_y)
    assert abs(nd.sum()) < 1e-9 and abs(fd.sum() - 256.0) < 1e-9
    d = density_db(64.0, [(32, 32, 2, 2, 0)])
    g = oracle.Grid(d, 1.0, 8)
    nd, _ = g.accumulate(d.node_x, d.node_y)
    assert abs(nd.sum() - 4.0) < 1e-9
    assert (nd > 1e-6).sum() > 1


def test_overflow_cases():
    cells = [(4 + 8 * i, 4 + 8 * j, 8, 8, 0) for i in range(8) for j in range(8)]
    d = density_db(64.0, cells)
    g = oracle.Grid(d, 1.0, 8)
    assert abs(g.overflow(g.accumulate(d.node_x, d.node_y)[0])) < 1e-12
    d = density_db(64.0, [(4, 4, 8, 8, 0)] * 64)
    g = oracle.Grid(d, 1.0, 8)
    assert g.overflow(g.accumulate(d.node_x, d.node_y)[0]) > 0.5
    base = [(4, 4, 8, 8, 0)] * 32
    da = density_db(64.0, base)
    db = density_db(64.0, base + [(4, 4, 8, 8, FL)] * 32)
    ga, gb = oracle.Grid(da, 1.0, 8), oracle.Grid(db, 1.0, 8)
    ta = ga.overflow(ga.accumulate(da.node_x, da.node_y)[0])
    tb = gb.overflow(gb.accumulate(db.node_x, db.node_y)[0])
    assert ta == tb
    # closed form: 32 cells of 64 stacked in one 8x8 bin of 64 bins, target 1:
    # overflow area = 32*64 - 64, denominator 32*64
    assert abs(ta - (32 * 64 - 64) / (32 * 64)) < 1e-12


def test_charge_matches_density():
    cells = [(20 + 4 * i, 30, 6, 6, 0) for i in range(20)]
    cells += [(30 + 18 * i, 70, 16, 16, M) for i in range(5)]
    cells += [(25 + 7 * i, 100, 5, 5, FL) for i in range(12)]
    d = density_db(128.0, cells)
    g = oracle.Grid(d, 0.6, 16)
    nd, fd = g.accumulate(d.node_x, d.node_y)
    assert abs(g.total_charge(d.node_x, d.node_y) - (nd.sum() + fd.sum())) < 1e-9


def brute_force_density(d, dim, target):
    """Pure-python FP64 overlap of every smoothed footprint with every bin."""
    lx, ly, hx, hy = d.core
    sx, sy = (hx - lx) / dim, (hy - ly) / dim
    nd = np.zeros((dim, dim))
    fd = np.zeros((dim, dim))
    for i in range(d.total_nodes):
        f = int(d.node_flags[i])
        if f & (D.F_FIXED | D.F_NI):
            continue
        w, h, cx, cy = float(d.node_w[i]), float(d.node_h[i]), float(d.node_x[i]), float(d.node_y[i])
        ew, eh = max(w, sx), max(h, sy)
        scale = (w / sx if w < sx else 1.0) * (h / sy if h < sy else 1.0)
        wt = scale * (target if (f & D.F_MACRO) and not (f & D.F_FILLER) else 1.0)
        for bx in range(dim):
            for by in range(dim):
                ox = min(cx + ew / 2, lx + (bx + 1) * sx) - max(cx - ew / 2, lx + bx * sx)
                oy = min(cy + eh / 2, ly + (by + 1) * sy) - max(cy - eh / 2, ly + by * sy)
                if ox > 0 and oy > 0:
                    (fd if f & D.F_FILLER else nd)[bx, by] += ox * oy * wt
    return nd.ravel(), fd.ravel()


def test_brute_force_small_grid():
    rng = np.random.default_rng(3)
    cells = [(rng.uniform(5, 59), rng.uniform(5, 59), rng.uniform(1, 12), rng.uniform(1, 12),
              int(rng.choice([0, M, FL]))) for _ in range(60)]
    d = density_db(64.0, cells)
    g = oracle.Grid(d, 0.7, 8)
    nd, fd = g.accumulate(d.node_x, d.node_y)
    bn, bf = brute_force_density(d, 8, 0.7)
    assert np.allclose(nd, bn, atol=1e-9) and np.allclose(fd, bf, atol=1e-9)


def test_fixed_point_mirror():
    d = D.with_fillers(D.fixture_eplace(500, 4))
    g32 = oracle.Grid(d, 1.0, 64, precision=32)
    g64 = oracle.Grid(d, 1.0, 64, precision=64)
    nq, fq = g32.accumulate_fixed(d.node_x, d.node_y, 32)
    nd, fd = g64.accumulate(d.node_x, d.node_y)
    assert np.allclose(nq / 2.0 ** 32, nd, rtol=1e-5, atol=1e-4)
    assert np.allclose(fq / 2.0 ** 32, fd, rtol=1e-5, atol=1e-4)


def test_overfull_tau_floor():
    """Movable area 1.4x the core: tau >= 0.4/1.4 for any placement."""
    cells = [(8 + 16 * (i % 8), 11.2 + (128 - 22.4) * (i // 8) / 7, 16, 22.4, 0)
             for i in range(64)]
    d = density_db(128.0, cells)
    g = oracle.Grid(d, 1.0, 8)
    tau = g.overflow(g.accumulate(d.node_x, d.node_y)[0])
    assert tau >= 0.4 / 1.4 - 1e-6   # 22.4 is not exact in float32


def test_bin_dim_power_of_two_and_static_terms():
    cells = [(10 + (i % 20) * 12, 10 + (i // 20) * 12, 8, 8, 0) for i in range(500)]
    d = density_db(256.0, cells)
    g = oracle.Grid(d, 1.0, 0)
    assert g.dim >= 4 and (g.dim & (g.dim - 1)) == 0
    # rows cover the core: no base density; no fixed nodes: no terminal density
    assert np.abs(g.base).max() == 0 and np.abs(g.terminal).max() == 0
    # a fixed 8x8 block at a bin corner adds target * its area in terminal density
    d = density_db(64.0, [(32, 32, 4, 4, 0), (4, 4, 8, 8, FX)])
    g = oracle.Grid(d, 0.5, 8)
    assert abs(g.terminal.sum() - 0.5 * 64) < 1e-9
