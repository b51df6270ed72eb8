"""Quadratic (Kraftwerk2 Bound2Bound) initial placement on the GPU through the
library's sp_qp_place / sp_qp_b2b_system, mirroring QuadraticPlacer::place
(src/init/quadratic_placer.cpp:149-228) and its Config fields
(src/util/config.h:49-74)."""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import hip


@dataclasses.dataclass
class QpConfig:
    max_iter: int = 20            # qp_max_iter
    tol: float = 1e-6             # qp_tol
    solver_max_iter: int = 100    # qp_solver_max_iter
    outer_tol: float = 1e-4       # qp_outer_tol
    min_distance: float = 1.0     # qp_min_distance
    ignore_net_degree: int = 100

    def struct(self) -> hip.SpQpConfig:
        return hip.SpQpConfig(self.max_iter, self.tol, self.solver_max_iter, self.outer_tol,
                              self.min_distance, self.ignore_net_degree)


def _weights(net_weight):
    if net_weight is None:
        return None, None
    w = np.ascontiguousarray(net_weight, np.float32)
    return w, w.ctypes.data


def qp_place(design, cfg: QpConfig | None = None, net_weight=None, max_rows: int = 64):
    """Returns (x, y, rows, iterations): node positions (num_nodes), and per
    outer iteration a dict with iter, hpwl, err (max of the two solver
    errors) and elapsed_ms."""
    cfg = cfg or QpConfig()
    da = hip._DesignArrays(design)
    n = design.num_nodes
    x = np.zeros(max(n, 1), np.float32)
    y = np.zeros(max(n, 1), np.float32)
    rows = (hip.SpHipIterMetrics * max(max_rows, 1))()
    its = C.c_int32(0)
    _, wp = _weights(net_weight)
    cs = cfg.struct()
    hip.check(hip.lib().sp_qp_place(C.byref(da.s), wp, C.byref(cs), x.ctypes.data,
                                    y.ctypes.data, rows, max_rows, C.byref(its)))
    out = [dict(iter=rows[i].iter, hpwl=rows[i].hpwl, err=rows[i].grad_norm_wl,
                elapsed_ms=rows[i].elapsed_ms) for i in range(min(its.value, max_rows))]
    return x[:n], y[:n], out, its.value


def b2b_system(design, axis: int, net_weight=None, min_distance: float = 1.0,
               ignore_net_degree: int = 100):
    """(diag, rhs, row_start, col, val) of one axis' B2B system at the
    design's positions, as the GPU assembles it."""
    da = hip._DesignArrays(design)
    _, wp = _weights(net_weight)
    nnz = C.c_int64(0)
    L = hip.lib()
    hip.check(L.sp_qp_b2b_system(C.byref(da.s), wp, int(axis), float(min_distance),
                                 int(ignore_net_degree), None, None, None, None, None, 0,
                                 C.byref(nnz)))
    m = design.num_movable
    diag, rhs = np.zeros(max(m, 1)), np.zeros(max(m, 1))
    rs = np.zeros(m + 1, np.int64)
    col = np.zeros(max(nnz.value, 1), np.int32)
    val = np.zeros(max(nnz.value, 1))
    hip.check(L.sp_qp_b2b_system(C.byref(da.s), wp, int(axis), float(min_distance),
                                 int(ignore_net_degree), diag.ctypes.data, rhs.ctypes.data,
                                 rs.ctypes.data, col.ctypes.data, val.ctypes.data, nnz.value,
                                 C.byref(nnz)))
    k = nnz.value
    return diag[:m], rhs[:m], rs, col[:k], val[:k]
