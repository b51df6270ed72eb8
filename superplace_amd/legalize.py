"""Legalization through the library: Abacus standard-cell legalization
(sp_abacus_*, mirroring legalizeCells / buildAbacusSegments,
src/lg/abacus.h:34-35) and simulated-annealing macro legalization
(sp_macro_legalize, mirroring legalizeMacros, src/lg/macro_sa.h:40)."""
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


def legalize_macros(design, x=None, y=None, seed=1002, moves=0, init_accept=0.9,
                    temp_ratio=1e-4, disp_weight=1.0, sink=None, row_step=None):
    """Macro legalization through the library (sp_macro_legalize), mirroring
    legalizeMacros(db, cfg, sink) (src/lg/macro_sa.h:40).

    Positions default to the design's; the net weights, node orientations
    and names come from design.meta when the design was read from BookShelf
    (all 1, N and the node index otherwise).  sink is a metrics.MetricsSink
    or None.  Returns a dict: x, y (centres of all num_nodes nodes), flags,
    pin_offset_x, orient (after the pass: macros fixed, flipped macros'
    pins mirrored) and result (the MacroSaResult fields)."""
    n = design.num_nodes
    x = np.ascontiguousarray(design.node_x[:n] if x is None else x, np.float32).copy()
    y = np.ascontiguousarray(design.node_y[:n] if y is None else y, np.float32).copy()
    flags = np.ascontiguousarray(design.node_flags, np.uint8).copy()
    offx = np.ascontiguousarray(design.pin_offset_x, np.float32).copy()
    orient = np.ascontiguousarray(design.meta.get("node_orient", np.zeros(design.total_nodes)),
                                  np.uint8).copy()
    wts = design.meta.get("net_weight")
    wts = None if wts is None else np.ascontiguousarray(wts, np.float32)
    steps = _steps(design, row_step)
    site = float(steps[0]) if design.num_rows else 1.0
    names = design.meta.get("node_names")
    name_arr = None
    if names is not None:
        enc = [s.encode() for s in names[:n]]
        name_arr = (C.c_char_p * len(enc))(*enc)
    cfg = hip.SpMlgConfig(random_seed=int(seed) & 0xFFFFFFFFFFFFFFFF, moves=int(moves),
                          init_accept=init_accept, temp_ratio=temp_ratio, disp_weight=disp_weight)
    res = hip.SpMlgResult()
    da = hip._DesignArrays(design)
    hip.check(hip.lib().sp_macro_legalize(
        C.byref(da.s), None if wts is None else wts.ctypes.data, site, design.row_height_mode,
        C.byref(cfg), x.ctypes.data, y.ctypes.data, flags.ctypes.data, offx.ctypes.data,
        orient.ctypes.data, None if sink is None else sink._h,
        None if name_arr is None else C.cast(name_arr, C.c_void_p), C.byref(res)))
    return dict(x=x, y=y, flags=flags, pin_offset_x=offx, orient=orient,
                result={k: getattr(res, k) for k, _ in hip.SpMlgResult._fields_})
