"""Pins the oracle's Poisson solve (cpu_poisson.cpp:149-207 restated).

Checks: the 1-D transforms equal scipy's unnormalised DCT/DST (FFTW's
REDFT10/REDFT01/RODFT01), the 2-D solve equals scipy-made golden fixtures,
closed forms for a single cosine mode (phi and xi), and the six analytic
properties of test/test_poisson.cpp:81-262.
"""
import os

import numpy as np
import pytest
from scipy.fft import dct, dst

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("n", [2, 3, 8, 17, 64])
def test_r2r_matches_scipy(n):
    x = np.random.default_rng(n).standard_normal(n)
    assert np.allclose(oracle.r2r_1d(0, x), dct(x, type=2), atol=1e-12)
    assert np.allclose(oracle.r2r_1d(1, x), dct(x, type=3), atol=1e-12)
    assert np.allclose(oracle.r2r_1d(2, x), dst(x, type=3), atol=1e-12)


@pytest.mark.parametrize("name", ["8x8", "16x12", "32x32", "12x20"])
def test_poisson_golden(name):
    g = np.load(os.path.join(GOLD, f"poisson_{name}.npz"))
    phi, ex, ey = oracle.poisson(g["rho"], *g["step"])
    for got, ref in ((phi, g["phi"]), (ex, g["ex"]), (ey, g["ey"])):
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def _grid(n, f):
    xs = np.arange(n)[:, None] + 0.5
    ys = np.arange(n)[None, :] + 0.5
    return f(xs, ys)


def test_single_mode_closed_form():
    # rho = cos(u X) cos(v Y) => phi = rho / (u^2+v^2),
    # xi_x = u sin(u X) cos(v Y) / (u^2+v^2), xi_y likewise (X = (x+1/2) step).
    n, step, j0, k0 = 64, 2.5, 3, 5
    u, v = np.pi * j0 / (n * step), np.pi * k0 / (n * step)
    rho = _grid(n, lambda X, Y: np.cos(np.pi * j0 * X / n) * np.cos(np.pi * k0 * Y / n))
    phi, ex, ey = oracle.poisson(rho, step, step)
    den = u * u + v * v
    assert np.allclose(phi, rho / den, atol=1e-9 * np.abs(rho / den).max())
    exx = _grid(n, lambda X, Y: u * np.sin(np.pi * j0 * X / n) * np.cos(np.pi * k0 * Y / n)) / den
    eyy = _grid(n, lambda X, Y: v * np.cos(np.pi * j0 * X / n) * np.sin(np.pi * k0 * Y / n)) / den
    assert np.allclose(ex, exx, atol=1e-9 * np.abs(exx).max())
    assert np.allclose(ey, eyy, atol=1e-9 * np.abs(eyy).max())


def _gauss(n, cx, cy, s):
    return _grid(n, lambda X, Y: np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * s * s)))


def test_analytic_properties():
    n, step = 64, 2.5
    # 1: Gaussian peak -> field points away, Neumann boundary, DC removed
    rho = _gauss(n, n / 2, n / 2, n / 10)
    rho -= rho.mean()
    phi, ex, ey = oracle.poisson(rho, step, step)
    scale = np.abs(ex).max() + np.abs(ey).max()
    X = np.arange(n)[:, None] + 0.5 - n / 2
    Y = np.arange(n)[None, :] + 0.5 - n / 2
    r = np.sqrt(X ** 2 + Y ** 2)
    mask = (r >= 3) & (r <= 0.4 * n) & (np.sqrt(ex ** 2 + ey ** 2) >= 1e-3 * scale)
    assert mask.sum() > 100
    assert ((X * ex + Y * ey)[mask] > 0).mean() >= 0.99
    normal = max(np.abs(ex[0]).max(), np.abs(ex[-1]).max(), np.abs(ey[:, 0]).max(),
                 np.abs(ey[:, -1]).max())
    assert normal < 0.05 * scale
    # 2: uniform -> zero field
    phi, ex, ey = oracle.poisson(np.zeros((n, n)), step, step)
    assert np.abs(ex).max() < 1e-12 and np.abs(ey).max() < 1e-12
    # 3: mirror symmetry in x -> xi_x antisymmetric
    rho = _grid(n, lambda Xx, Yy: np.exp(-((Xx - n / 2) ** 2) / 50) * (1 + 0.3 * np.sin(0.2 * (Yy - 0.5))))
    rho -= rho.mean()
    _, ex, _ = oracle.poisson(rho, step, step)
    assert np.abs(ex + ex[::-1]).max() < 1e-9 * np.abs(ex).max()
    # 5: xi = -grad phi by central differences (catches the DST index shift)
    rho = _grid(n, lambda Xx, Yy: np.cos(np.pi * 2 * Xx / n) * np.cos(np.pi * 3 * Yy / n)
                + 0.5 * np.cos(np.pi * Xx / n))
    rho -= rho.mean()
    phi, ex, ey = oracle.poisson(rho, step, step)
    gx = (phi[2:, 1:-1] - phi[:-2, 1:-1]) / (2 * step)
    gy = (phi[1:-1, 2:] - phi[1:-1, :-2]) / (2 * step)
    assert np.abs(ex[1:-1, 1:-1] + gx).max() < 0.01 * np.abs(ex).max()
    assert np.abs(ey[1:-1, 1:-1] + gy).max() < 0.01 * np.abs(ey).max()
    # 6: determinism
    a = oracle.poisson(rho, step, step)
    b = oracle.poisson(rho, step, step)
    assert all(np.array_equal(p, q) for p, q in zip(a, b))


def test_planted_index_shift_is_caught():
    """A plausible mistake (no index shift before RODFT01) breaks the
    central-difference property, so that property pins the shift."""
    n, step = 32, 1.0
    rho = _grid(n, lambda X, Y: np.cos(np.pi * 2 * X / n) * np.cos(np.pi * 3 * Y / n))
    F = dct(dct(rho, type=2, axis=1), type=2, axis=0)
    u = np.pi * np.arange(n) / (n * step)
    den = u[:, None] ** 2 + u[None, :] ** 2
    P = np.where(den > 0, F / (4 * n * n) / np.where(den > 0, den, 1), 0)
    phi = dct(dct(P, type=3, axis=1), type=3, axis=0)
    wrong = dst(dct(P * u[:, None], type=3, axis=1), type=3, axis=0)   # unshifted
    gx = (phi[2:, 1:-1] - phi[:-2, 1:-1]) / (2 * step)
    assert np.abs(wrong[1:-1, 1:-1] + gx).max() > 0.1 * np.abs(wrong).max()
