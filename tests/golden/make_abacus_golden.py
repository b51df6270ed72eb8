"""Writes tests/golden/ref_abacus.npz: for every case of abacus_cases.py, the
node centres and the AbacusResult the reference's own legalizeCells
(src/lg/abacus.cpp, built into oracle/_ref by `make -C oracle ref`) gives.

Run from the repo root: python tests/golden/make_abacus_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)
from abacus_cases import cases, row_sites  # noqa: E402


def main():
    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libsp_ref.so"))
    L.ref_abacus.restype = C.c_int
    L.ref_abacus.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 5 + [C.c_int] + [C.c_void_p] * 5 + \
        [C.c_int, C.c_void_p]
    p = lambda a: a.ctypes.data
    out = {}
    for name, (d, x, y, step, opt) in cases().items():
        n = d.num_nodes
        x = np.ascontiguousarray(x, np.float32).copy()
        y = np.ascontiguousarray(y, np.float32).copy()
        arrs = [np.ascontiguousarray(a) for a in (
            d.node_w[:n].astype(np.float32), d.node_h[:n].astype(np.float32),
            d.node_flags[:n].astype(np.uint8), d.row_ly.astype(np.float32),
            d.row_height.astype(np.float32), d.row_lx.astype(np.float32),
            np.asarray(step, np.float32), row_sites(d, step))]
        res = np.zeros(10)
        rc = L.ref_abacus(d.num_movable, n, p(x), p(y), p(arrs[0]), p(arrs[1]), p(arrs[2]),
                          d.num_rows, p(arrs[3]), p(arrs[4]), p(arrs[5]), p(arrs[6]), p(arrs[7]),
                          opt, p(res))
        assert rc == 0, name
        out[f"{name}_x"], out[f"{name}_y"], out[f"{name}_res"] = x, y, res
        print(name, res[:8].astype(int), res[8:])
    np.savez_compressed(os.path.join(HERE, "ref_abacus.npz"), **out)


if __name__ == "__main__":
    main()
