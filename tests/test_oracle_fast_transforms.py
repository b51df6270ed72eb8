"""The oracle's FFT transforms (the CPU baseline's path) against its direct
sums, which the golden vectors and closed forms pin (test_oracle_poisson)."""
import numpy as np
import pytest

import oracle


@pytest.mark.parametrize("n", [2, 4, 16, 128, 1024])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_fast_1d_matches_direct(kind, n):
    x = np.random.default_rng(n + kind).standard_normal(n)
    direct = oracle.r2r_1d(kind, x)
    fast = oracle.r2r_1d_fast(kind, x)
    assert np.abs(fast - direct).max() <= 1e-12 * np.abs(direct).max()


@pytest.mark.parametrize("nx,ny", [(32, 32), (64, 128), (256, 256)])
def test_fast_poisson_matches_direct(nx, ny):
    rho = np.random.default_rng(nx * ny).random((nx, ny))
    oracle.set_fast_poisson(False)
    ref = oracle.poisson(rho, 3.0, 2.0)
    oracle.set_fast_poisson(True)
    try:
        fast = oracle.poisson(rho, 3.0, 2.0)
    finally:
        oracle.set_fast_poisson(False)
    for a, b in zip(fast, ref):
        assert np.abs(a - b).max() <= 1e-11 * np.abs(b).max()


def test_fast_path_falls_back_for_other_sizes():
    rho = np.random.default_rng(5).random((24, 40))
    ref = oracle.poisson(rho, 1.0, 1.0)
    oracle.set_fast_poisson(True)
    try:
        out = oracle.poisson(rho, 1.0, 1.0)
    finally:
        oracle.set_fast_poisson(False)
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(a, b)
