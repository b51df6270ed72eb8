"""Abacus standard-cell legalization through the library (sp_abacus_*),
mirroring legalizeCells / buildAbacusSegments (src/lg/abacus.h:34-35)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import hip

FULL_TRIALS, AREA_WEIGHT = 1, 2   # AbacusOptions::referenceTrials, ::areaWeight


def _steps(design, row_step):
    if row_step is None:
        row_step = design.meta.get("row_step")
    if row_step is None:
        row_step = np.ones(design.num_rows, np.float32)
    return np.ascontiguousarray(row_step, np.float32)


def legalize_cells(design, x=None, y=None, row_step=None, reference_trials=False,
                   area_weight=False):
    """Returns (x, y, result dict): centres of all num_nodes nodes after
    legalization (inputs default to the design's), AbacusResult fields."""
    n = design.num_nodes
    x = np.ascontiguousarray(design.node_x[:n] if x is None else x, np.float32).copy()
    y = np.ascontiguousarray(design.node_y[:n] if y is None else y, np.float32).copy()
    steps = _steps(design, row_step)
    da = hip._DesignArrays(design)
    res = hip.SpAbacusResult()
    flags = (FULL_TRIALS if reference_trials else 0) | (AREA_WEIGHT if area_weight else 0)
    hip.check(hip.lib().sp_abacus_legalize(C.byref(da.s), steps.ctypes.data, x.ctypes.data,
                                           y.ctypes.data, flags, C.byref(res)))
    return x, y, {k: getattr(res, k) for k, _ in hip.SpAbacusResult._fields_}


def segments(design, row_step=None):
    """(row, lx, hx) arrays of the free row segments."""
    steps = _steps(design, row_step)
    da = hip._DesignArrays(design)
    cnt = C.c_int32(0)
    L = hip.lib()
    hip.check(L.sp_abacus_segments(C.byref(da.s), steps.ctypes.data, None, None, None, 0,
                                   C.byref(cnt)))
    m = max(cnt.value, 1)
    row, lx, hx = np.zeros(m, np.int32), np.zeros(m, np.float32), np.zeros(m, np.float32)
    hip.check(L.sp_abacus_segments(C.byref(da.s), steps.ctypes.data, row.ctypes.data,
                                   lx.ctypes.data, hx.ctypes.data, m, C.byref(cnt)))
    k = cnt.value
    return row[:k], lx[:k], hx[:k]
