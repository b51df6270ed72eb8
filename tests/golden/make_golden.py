"""Writes the small golden fixtures under tests/golden/.

* poisson_*.npz  : phi / xi_x / xi_y for small grids computed with scipy.fft
  (dct/dst types 2 and 3, norm=None, whose unnormalised definitions equal
  FFTW's REDFT10 / REDFT01 / RODFT01 used by src/numeric/cpu/cpu_poisson.cpp).
  This is an independent code path from the oracle's direct summation.
* ref_wa_*.npz, ref_hpwl.npz, ref_nesterov_*.npz : outputs of the reference's
  own WaWirelength::compute, computeHPWL and NesterovOptimizer, compiled from
  /root/reference by oracle/Makefile into oracle/_ref (run `make -C oracle ref`
  first).  Inputs are small seeded netlists / quadratics.

Run from the repo root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
from scipy.fft import dct, dst

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.dirname(os.path.abspath(__file__))


def scipy_poisson(rho, sx, sy):
    nx, ny = rho.shape
    F = dct(dct(rho, type=2, axis=1), type=2, axis=0)
    u = np.pi * np.arange(nx) / (nx * sx)
    v = np.pi * np.arange(ny) / (ny * sy)
    den = u[:, None] ** 2 + v[None, :] ** 2
    P = np.where(den > 0, F / (4.0 * nx * ny) / np.where(den > 0, den, 1), 0.0)
    phi = dct(dct(P, type=3, axis=1), type=3, axis=0)
    I = np.zeros_like(P)
    I[:-1, :] = P[1:, :] * u[1:, None]
    ex = dst(dct(I, type=3, axis=1), type=3, axis=0)
    J = np.zeros_like(P)
    J[:, :-1] = P[:, 1:] * v[None, 1:]
    ey = dct(dst(J, type=3, axis=1), type=3, axis=0)
    return phi, ex, ey


def poisson_fixtures():
    rng = np.random.default_rng(1002)
    for (nx, ny, sx, sy) in [(8, 8, 2.5, 2.5), (16, 12, 1.0, 3.0), (32, 32, 22.65625, 22.65625),
                             (12, 20, 4.0, 1.5)]:
        rho = rng.uniform(0.0, 1.0, (nx, ny))
        phi, ex, ey = scipy_poisson(rho, sx, sy)
        np.savez(os.path.join(OUT, f"poisson_{nx}x{ny}.npz"), rho=rho, phi=phi, ex=ex, ey=ey,
                 step=np.array([sx, sy]))


def small_netlist(n_mov, n_fix, nets, seed):
    """A netlist of the shape used by test/test_wirelength.cpp:19-57."""
    from superplace_amd.design import Rng, make_design, F_FIXED
    rng = Rng(seed)
    total = n_mov + n_fix
    x = rng.uniform(total, 100.0, 900.0)
    y = rng.uniform(total, 100.0, 900.0)
    w = np.concatenate([np.full(n_mov, 6.0), np.full(n_fix, 20.0)])
    h = np.concatenate([np.full(n_mov, 12.0), np.full(n_fix, 20.0)])
    flags = np.concatenate([np.zeros(n_mov, np.uint8), np.full(n_fix, F_FIXED, np.uint8)])
    p2n, p2e = [], []
    for k, net in enumerate(nets):
        for nd in net:
            p2n.append(nd)
            p2e.append(k)
    offx = rng.uniform(len(p2n), -4.0, 4.0)
    offy = rng.uniform(len(p2n), -4.0, 4.0)
    rows = (np.array([0.0]), np.array([1000.0]), np.array([0.0]), np.array([1000.0]))
    return make_design(x, y, w, h, flags, n_mov, p2n, p2e, offx, offy, len(nets), rows)


WA_CASES = {
    "three_pin": (3, 0, [[0, 1, 2]], 11, [50.0]),
    "fixed_pin": (3, 1, [[0, 1, 2, 3]], 22, [50.0]),
    "shared": (5, 0, [[0, 1, 2], [1, 3], [2, 3, 4], [0, 4]], 33, [50.0]),
    "gamma_sweep": (6, 2, [[0, 1, 2, 6], [2, 3, 4], [4, 5, 7], [0, 5]], 44, [500.0, 50.0, 5.0]),
    "forty_pin": (40, 0, [list(range(40))], 55, [20.0]),
    "tiny_gamma": (8, 0, [[0, 1, 2, 3], [4, 5, 6, 7]], 66, [0.01]),
}


def reference_fixtures():
    import oracle
    if oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref/libsp_ref.so missing: run make -C oracle ref")
    for name, (nm, nf, nets, seed, gammas) in WA_CASES.items():
        d = small_netlist(nm, nf, nets, seed)
        out = {}
        for g in gammas:
            wl, grad = oracle.ref_wa(d, g, ignore_net_degree=1000)
            out[f"wl_{g}"] = np.array(wl)
            out[f"grad_{g}"] = grad
        np.savez(os.path.join(OUT, f"ref_wa_{name}.npz"), **out)
    # a generator-made netlist with large nets, default ignore degree 100
    from superplace_amd import design as D
    d = D.synth_design(300, 3, 320, die=200.0, util=0.6, tail="powerlaw", tail_max=150,
                       seed=5, with_fillers=False, locality=4.0)
    wl, grad = oracle.ref_wa(d, 7.5, ignore_net_degree=100)
    hp = oracle.ref_hpwl(d)
    np.savez(os.path.join(OUT, "ref_wa_synth300.npz"), wl=np.array(wl), grad=grad,
             hpwl=np.array(hp))
    # Nesterov on the quadratics of test/test_nesterov.cpp
    cases = {
        "isotropic": (np.full(64, 2.0), 10.0 + 0.5 * np.arange(64), np.zeros(64), 200, 0, 0, 0.0,
                      np.inf),
        "illcond_bb": (0.01 * 10.0 ** (3.0 * np.arange(50) / 49), -5.0 + 0.3 * np.arange(50),
                       np.full(50, 100.0), 400, 1, 0, 0.0, np.inf),
        "illcond_bt": (0.01 * 10.0 ** (3.0 * np.arange(50) / 49), -5.0 + 0.3 * np.arange(50),
                       np.full(50, 100.0), 400, 1, 1, 0.0, np.inf),
        "huge_step_bt": (np.full(32, 4.0), np.full(32, 3.0), np.zeros(32), 3, 1, 1, 12.5, np.inf),
        "huge_step": (np.full(32, 4.0), np.full(32, 3.0), np.zeros(32), 3, 1, 0, 12.5, np.inf),
        "constraint": (np.ones(16), np.full(16, 1000.0), np.zeros(16), 100, 1, 0, 0.0, 5.0),
    }
    for name, (c, t, p0, steps, bb, bt, init, hi) in cases.items():
        obj, pos, ev = oracle.nesterov_quadratic(c, t, p0, steps, bb, bt, init, hi,
                                                 use_reference=True)
        np.savez(os.path.join(OUT, f"ref_nesterov_{name}.npz"), coef=np.float32(c),
                 target=np.float32(t), pos0=np.float32(p0),
                 params=np.array([steps, bb, bt, init, hi]), obj=np.array(obj), pos=pos,
                 evals=np.array(ev))


if __name__ == "__main__":
    poisson_fixtures()
    reference_fixtures()
    print("golden fixtures written to", OUT)
