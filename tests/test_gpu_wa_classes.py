"""WA pass of the device engine on every net size class it has, against the
FP64 oracle (wa_wirelength.cpp:119-225, hpwl.cpp:13-39).

The engine splits nets by degree: degree 2 and 3-4 take one lane each,
5-64 a group of lanes, larger nets one wavefront (their pins batched in
registers up to 128 pins, or 512 for the extrema of nets above
ignore_net_degree, else strided loops).  The design here is the 2,000-cell
synthetic design plus nets of chosen degrees around every one of those
boundaries, run with ignore_net_degree values that put the boundaries on
both sides of the WA/HPWL-only split."""
import numpy as np
import pytest

import oracle
from superplace_amd import design as D
from superplace_amd import hip

pytestmark = pytest.mark.gpu

EXTRA_DEGREES = [1, 2, 3, 4, 5, 8, 9, 16, 17, 32, 33, 64, 65, 100, 101, 127, 128, 129, 300,
                 511, 512, 513, 700, 1100]


def design_with_wide_nets(seed=11):
    base = D.config_small(with_fillers=False)
    rng = np.random.default_rng(seed)
    nm, nn = base.num_movable, base.num_nodes
    p2n, p2e = [base.pin2node], [base.pin2net]
    offx, offy = [base.pin_offset_x], [base.pin_offset_y]
    net = base.num_nets
    for deg in EXTRA_DEGREES:
        # mostly movable cells, one pin on a fixed macro for the larger nets
        nodes = rng.integers(0, nm, size=deg).astype(np.int32)
        if deg > 8:
            nodes[0] = rng.integers(nm, nn)
        p2n.append(nodes)
        p2e.append(np.full(deg, net, np.int32))
        offx.append(((rng.random(deg) - 0.5) * base.node_w[nodes]).astype(np.float32))
        offy.append(((rng.random(deg) - 0.5) * base.node_h[nodes]).astype(np.float32))
        net += 1
    d = D.make_design(base.node_x, base.node_y, base.node_w, base.node_h, base.node_flags,
                      nm, np.concatenate(p2n), np.concatenate(p2e), np.concatenate(offx),
                      np.concatenate(offy), net,
                      (base.row_ly, base.row_height, base.row_lx, base.row_hx),
                      name="wide-nets")
    return D.with_fillers(d)


@pytest.fixture(scope="module")
def wide():
    return design_with_wide_nets()


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("ignore", [2, 3, 100, 130, 2000])
def test_wa_every_size_class_matches_oracle(wide, ignore, prec):
    d = wide
    deg = np.diff(d.net2pin_start)
    assert deg.max() >= 1100 and (deg == 513).any()
    eng = hip.GpEngine(d, hip.GpConfig(precision=prec, gp_bin_dim=128, ignore_net_degree=ignore))
    gam = 18.0
    sc, out = eng.eval(lam=0.3, gamma=gam, debug=True)
    wl, gwl = oracle.wa(d, gam, ignore, precision=64)
    idx = eng.opt_idx
    nm = d.num_movable
    has = idx < nm
    wref = np.concatenate([np.where(has, gwl[np.minimum(idx, nm - 1)], 0),
                           np.where(has, gwl[nm + np.minimum(idx, nm - 1)], 0)])
    tol_grad, tol_wl, tol_hpwl = (2e-4, 2e-5, 1e-6) if prec == 32 else (1e-9, 1e-11, 1e-12)
    print(ignore, prec, rel(out["grad_wl"], wref), sc["wl"], wl, flush=True)
    assert rel(out["grad_wl"], wref) < tol_grad
    assert sc["wl"] == pytest.approx(wl, rel=tol_wl)
    assert sc["hpwl"] == pytest.approx(oracle.hpwl(d, 64), rel=tol_hpwl)
