"""Error study of split-BF16 transform products for the Poisson solve (the
DCT-as-GEMM form of the dense path), against the field criterion of the
reference's CPU/GPU parity test: max |field error| / max |field| < 1e-5
(test/test_poisson_parity.cpp:22).

Every matrix product of the dense solve (forward REDFT10 along y and x, the
spectral divide, the REDFT01/RODFT01 products of the fields) is formed from
bfloat16 pieces of its FP32 operands, x = hi + mid + lo with each piece the
round-to-nearest-even bfloat16 of what the previous ones leave, and summed
in FP32 as the matrix cores accumulate:

- two pieces, products hi*hi + hi*lo + lo*hi (three bf16 MFMAs per FP32
  product): the dropped lo*lo term and the residues beyond lo are about
  2^-16 of each product, which lands at 1-2e-5 of the field: it FAILS the
  criterion on a clustered density at 256^2 (and at 512^2: 2.1e-5);
- three pieces, products hi*hi + hi*mid + mid*hi + hi*lo + mid*mid + lo*hi
  (six bf16 MFMAs): about 2e-6, the FP32 products' own level.

So the FP32-accurate bf16 path costs six bf16 products per FP32 product:
2.5 / 6 = 0.42 PFLOP/s of matrix peak against 0.157 for FP32 MFMA.  The
oracle's FP64 solve is the truth."""
import numpy as np
import pytest

import oracle


def bf16(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32).view(np.float32)


def pieces(x, n):
    x = np.asarray(x, np.float32)
    out, rest = [], x
    for _ in range(n):
        p = bf16(rest)
        out.append(p)
        rest = (rest - p).astype(np.float32)
    return out


def product(a, b, n):
    """FP32-accumulated sum of the bf16 piece products with i + j < n
    (i, j = piece indices from 0 = hi)."""
    pa, pb = pieces(a, n), pieces(b, n)
    acc = np.zeros((a.shape[0], b.shape[1]), np.float32)
    for i in range(n):
        for j in range(n - i):
            acc += pa[i] @ pb[j]
    return acc


def dense_solve(rho, sx, sy, n):
    nx, ny = rho.shape
    pi = np.pi
    y, k, x, j = np.arange(ny), np.arange(ny), np.arange(nx), np.arange(nx)
    cy = 2 * np.cos(pi * np.outer(y + 0.5, k) / ny)                       # [y][k]
    cx = 2 * np.cos(pi * np.outer(j, x + 0.5) / nx)                       # [j][x]
    u, v = pi * j / (nx * sx), pi * k / (ny * sy)
    ay = np.where(k == 0, 1.0, 2.0)[:, None] * np.cos(pi * np.outer(k, y + 0.5) / ny)   # [k][y]
    ax = np.where(j == 0, 1.0, 2.0)[None, :] * np.cos(pi * np.outer(x + 0.5, j) / nx)   # [x][j]
    sxu = np.where(j == 0, 0.0, 2.0 * u)[None, :] * np.sin(pi * np.outer(x + 0.5, j) / nx)
    bsv = np.where(k == 0, 0.0, 2.0 * v)[:, None] * np.sin(pi * np.outer(k, y + 0.5) / ny)
    mats = [m.astype(np.float32) for m in (cy, cx, ay, ax, sxu, bsv)]
    cy, cx, ay, ax, sxu, bsv = mats
    f = product(cx, product(rho, cy, n), n).astype(np.float64)
    den = u[:, None] ** 2 + v[None, :] ** 2
    den[0, 0] = 1.0
    p = (f / den / (4.0 * nx * ny)).astype(np.float32)
    p[0, 0] = 0.0
    t1 = product(p, ay, n)
    ex = product(sxu, t1, n)
    ey = product(ax, product(p, bsv, n), n)
    return ex.astype(np.float64), ey.astype(np.float64)


def clustered(n, seed=5):
    xs, ys = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    r2 = (xs - n / 2) ** 2 + (ys - n / 2) ** 2
    rho = 3 * np.exp(-r2 / (0.02 * n * n)) + 0.1 * np.random.default_rng(seed).uniform(0, 1, (n, n))
    rho = rho.astype(np.float32)
    return rho - np.float32(rho.mean(dtype=np.float64))


def field_error(n, pieces_n):
    rho = clustered(n)
    _, ex0, ey0 = oracle.poisson(rho, 22.65, 22.65, 64)
    ex, ey = dense_solve(rho, 22.65, 22.65, pieces_n)
    return max(np.abs(ex - ex0).max() / np.abs(ex0).max(), np.abs(ey - ey0).max() / np.abs(ey0).max())


def test_two_piece_split_misses_the_field_criterion():
    err = field_error(256, 2)
    print("two pieces:", err)
    assert err > 1e-5


def test_three_piece_split_meets_it():
    err = field_error(256, 3)
    print("three pieces:", err)
    assert err < 5e-6
