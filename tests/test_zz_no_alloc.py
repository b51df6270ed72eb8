"""sp_hip_poisson_solve_noalloc: the FP32 Poisson solve over the library's
static scratch array, allocating nothing, against sp_hip_poisson_solve_dev
(an FP32 handle) on the same grids, two calls in a row (the second on a
smaller grid, so it runs over what the first left in the shared array), and
its size checks.  Both paths run the same kernels with the same tables, so
the results are compared bit for bit."""
import ctypes as C

import numpy as np
import pytest

import oracle
from superplace_amd import hip
from tests.hiprt import DevArray

pytestmark = pytest.mark.gpu


def last_error():
    return hip.lib().sp_hip_last_error().decode()


def noalloc(nx, ny, sx, sy, d_in, d_out):
    L = hip.lib()
    f = L.sp_hip_poisson_solve_noalloc
    f.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]
    return f(nx, ny, sx, sy, C.c_void_p(d_in), C.c_void_p(d_out), None)


def density(nx, ny, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.0, 2.0, (nx, ny)).astype(np.float32)
    return rho - rho.mean(dtype=np.float64).astype(np.float32)


def with_handle(rho, sx, sy):
    nx, ny = rho.shape
    be = hip.HipPoissonBackend(hip.BackendConfig(nx, ny, sx, sy))
    d_in = DevArray(rho)
    d_out = DevArray.zeros(3 * nx * ny, np.float32)
    be.solve_device(d_in.ptr, d_out.ptr)
    out = d_out.get()
    be.close()
    return out


@pytest.mark.parametrize("first,second", [((512, 512), (256, 128)), ((1024, 1024), (64, 64))])
def test_noalloc_matches_handle_twice(first, second):
    for k, (nx, ny) in enumerate((first, second)):
        print(f"no-alloc solve {nx}x{ny} (call {k + 1})", flush=True)
        sx, sy = 22.65625 * (k + 1), 17.5
        rho = density(nx, ny, 10 + k)
        d_in = DevArray(rho)
        d_out = DevArray.zeros(3 * nx * ny, np.float32)
        assert noalloc(nx, ny, sx, sy, d_in.ptr, d_out.ptr) == 0, last_error()
        got = d_out.get()
        ref = with_handle(rho, sx, sy)
        assert np.array_equal(got, ref), np.abs(got - ref).max()
        if nx * ny <= 256 * 128:
            # and the field itself against the oracle's direct sums
            _, ex, ey = oracle.poisson(rho.astype(np.float64), sx, sy)
            n = nx * ny
            for g, r in ((got[n:2 * n], ex.ravel()), (got[2 * n:], ey.ravel())):
                assert np.abs(g - r).max() / np.abs(r).max() < 1e-5


@pytest.mark.parametrize("nx,ny", [(2048, 64), (64, 2048), (96, 64), (32, 32), (512, 100)])
def test_noalloc_rejects_sizes_outside_the_scratch(nx, ny):
    d_in = DevArray.zeros(8, np.float32)
    d_out = DevArray.zeros(8, np.float32)
    assert noalloc(nx, ny, 1.0, 1.0, d_in.ptr, d_out.ptr) == hip.SP_HIP_EINVAL
    assert "powers of two" in last_error()
