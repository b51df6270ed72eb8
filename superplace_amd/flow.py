"""The placement flow around the library's entry points, for a caller outside
the reference tree: the stages superplace-cli chains (src/main.cpp:87-118)
with the same order, hand-offs, `.pl` snapshots and summary keys.

- init: quadratic initial placement (InitStage, src/stages/init_stage.cpp:15-53,
  with the quadratic placer) on the GPU (sp_qp_place);
- gp: GlobalPlaceStage::run (src/stages/gp_stage.cpp:56-244): fillers, mGP,
  and when movable macros exist mLG (host, bit-exact), filler re-scatter,
  FILLERONLY and cGP, each stage on the GPU engine (sp_hip_gp_*);
- lg: LegalizeStage::run (src/stages/lg_stage.cpp:11-41): Abacus (host,
  bit-exact).

After each stage <out>/<name>.<stage>.pl is written; summary.json gets
run_status "completed", or "failed" with failure_message / failure_kind /
failure_stage (src/main.cpp:119-138)."""
from __future__ import annotations

import dataclasses
import os
import time

import numpy as np

from . import bookshelf, hip, legalize, metrics, qp
from .design import F_FILLER, F_FIXED, F_MACRO, F_NI, add_fillers, filler_plan


@dataclasses.dataclass
class FlowConfig:
    """The Config fields the stages read (src/util/config.h)."""
    gp: hip.GpConfig = dataclasses.field(default_factory=hip.GpConfig)
    random_seed: int = 1002
    stages: tuple = ("gp", "lg")
    lg_cell_weight: str = "unit"        # "area": AbacusOptions::areaWeight
    macro_sa_moves: int = 0
    macro_sa_init_accept: float = 0.9
    macro_sa_temp_ratio: float = 1e-4
    macro_sa_disp_weight: float = 1.0
    qp: qp.QpConfig = dataclasses.field(default_factory=qp.QpConfig)


class StageError(RuntimeError):
    """An invariant of a stage failed (InvariantError, src/util/invariant.h)."""

    def __init__(self, kind, stage, message):
        super().__init__(message)
        self.kind, self.stage = kind, stage


def hpwl(d, x=None, y=None) -> float:
    """computeHPWL (src/db/hpwl.cpp:13-39): per net with >= 2 pins, the
    float extents of its pins summed in double."""
    x = d.node_x if x is None else x
    y = d.node_y if y is None else y
    px = (x[d.pin2node] + d.pin_offset_x).astype(np.float32)
    py = (y[d.pin2node] + d.pin_offset_y).astype(np.float32)
    starts = d.net2pin_start
    deg = np.diff(starts)
    keep = deg >= 2
    if not keep.any():
        return 0.0
    order = d.flat_net2pin
    sx, sy = px[order], py[order]
    idx = starts[:-1][keep]
    ext = (np.maximum.reduceat(sx, idx) - np.minimum.reduceat(sx, idx)).astype(np.float64) + \
          (np.maximum.reduceat(sy, idx) - np.minimum.reduceat(sy, idx)).astype(np.float64)
    return float(ext.sum())


def _without_fillers(d):
    if d.num_fillers == 0:
        return d
    n = d.num_nodes
    out = d.copy()
    for f in ("node_x", "node_y", "node_w", "node_h", "node_flags"):
        setattr(out, f, getattr(d, f)[:n].copy())
    out.node2pin_start = d.node2pin_start[:n + 1].copy()
    out.num_fillers = 0
    return out


def _with_positions(d, x, y):
    out = d.copy()
    out.node_x = np.asarray(x, np.float32).copy()
    out.node_y = np.asarray(y, np.float32).copy()
    return out


def _run_stage(d, cfg: FlowConfig, stage: int, name: str, sink):
    """EPlace(db, cfg, sink).run(stage): the stage's metrics rows and keys."""
    gcfg = dataclasses.replace(cfg.gp, stage=stage)
    eng = hip.GpEngine(d, gcfg)
    try:
        res, rows = eng.run()
        x, y = eng.positions()
    finally:
        eng.close()
    if sink is not None:
        sink.begin_stage(name)
        if rows:
            sink.push(rows)
        sink.end_stage(1e3 * res["seconds_loop"])
        ref = (res["initial_hpwl"] * gcfg.delta_hpwl_relative
               if gcfg.lambda_relative and stage != 1 and res["initial_hpwl"] > 0
               else gcfg.delta_hpwl_ref)
        sink.gp_result(name, res, ref)
    return _with_positions(d, x, y), res


def _overflow(d, cfg: FlowConfig) -> float:
    """tau of the current placement (a fresh BinGrid, as gp_stage.cpp:139-146)."""
    eng = hip.GpEngine(d, dataclasses.replace(cfg.gp, stage=0))
    try:
        sc = eng.eval(x=d.node_x, y=d.node_y, lam=1.0, gamma=1.0)
    finally:
        eng.close()
    return float(sc["tau"])


def init_stage(d, cfg: FlowConfig, sink=None):
    """InitStage with the quadratic placer: B2B QP on the GPU."""
    t0 = time.time()
    before = hpwl(d)
    x, y, _, _ = qp.qp_place(d, cfg.qp, net_weight=d.meta.get("net_weight"))
    n = d.num_nodes
    nx, ny = d.node_x.copy(), d.node_y.copy()
    nx[:n], ny[:n] = x, y
    d = _with_positions(d, nx, ny)
    out_of_core = _out_of_core(d, d.num_movable, 1e-3)
    if sink is not None:
        sink.set_summary("init_method", "quadratic")
        sink.set_summary("init_hpwl", hpwl(d))
        sink.set_summary("init_hpwl_before", before)
        sink.set_summary("init_time_ms", 1e3 * (time.time() - t0))
        sink.set_summary("init_out_of_core", out_of_core)
    return d


def check_geometry(d, stage: str, bounds: bool) -> None:
    """checkGeometry (src/util/invariant.h:20-33): finite geometry; with
    bounds, movable nodes (frozen macros included) and fillers inside the
    core to 0.01, terminal_NI exempt."""
    if not (np.all(np.isfinite(d.node_x)) and np.all(np.isfinite(d.node_y)) and
            np.all(np.isfinite(d.node_w)) and np.all(np.isfinite(d.node_h))):
        raise StageError("non_finite", stage, "non-finite node geometry")
    if not bounds:
        return
    i = np.arange(d.total_nodes)
    sel = ((i < d.num_movable) | ((d.node_flags & F_FILLER) != 0)) & ((d.node_flags & F_NI) == 0)
    i = i[sel]
    lx, ly, hx, hy = (np.float32(v) for v in d.core)
    hw, hh = np.float32(0.5) * d.node_w[i], np.float32(0.5) * d.node_h[i]
    eps = np.float32(0.01)
    ok = (d.node_x[i] - hw >= lx - eps) & (d.node_y[i] - hh >= ly - eps) & \
         (d.node_x[i] + hw <= hx + eps) & (d.node_y[i] + hh <= hy + eps)
    if not ok.all():
        raise StageError("out_of_core", stage, "node outside core")


def _out_of_core(d, count, tol, skip_fixed=False):
    lx, ly, hx, hy = d.core
    i = np.arange(count)
    if skip_fixed:
        i = i[(d.node_flags[i] & (F_FIXED | F_NI)) == 0]
    hw, hh = 0.5 * d.node_w[i], 0.5 * d.node_h[i]
    bad = (d.node_x[i] - hw < lx - tol) | (d.node_x[i] + hw > hx + tol) | \
          (d.node_y[i] - hh < ly - tol) | (d.node_y[i] + hh > hy + tol)
    return int(bad.sum())


def gp_stage(d, cfg: FlowConfig, sink=None, out_prefix=None):
    """GlobalPlaceStage::run: returns the design with its fillers and the
    delivered positions (macros fixed when mLG ran)."""
    t0 = time.time()
    d = _without_fillers(d)
    check_geometry(d, "gp/input", False)
    before = hpwl(d)
    n = d.num_nodes
    mov = np.arange(d.num_movable)
    movable_macros = int(np.sum(((d.node_flags[mov] & F_MACRO) != 0) &
                                ((d.node_flags[mov] & F_FIXED) == 0)))
    count = filler_plan(d, cfg.gp.target_density)[0]
    d = add_fillers(d, cfg.gp.target_density, cfg.random_seed)

    d, mgp = _run_stage(d, cfg, 0, "mGP", sink)
    if out_prefix and "bookshelf" in d.meta:
        bookshelf.write_pl(out_prefix + ".mgp.pl", d)
    final = mgp
    if sink is not None:
        sink.set_summary("mgp_hpwl", mgp["hpwl"])
        sink.set_summary("mgp_overflow", mgp["overflow"])
        sink.set_summary("mgp_iterations", mgp["iterations"])
    if movable_macros:
        is_macro = (d.node_flags[mov] & F_MACRO) != 0
        pre_x, pre_y = d.node_x[mov][is_macro].copy(), d.node_y[mov][is_macro].copy()
        out = legalize.legalize_macros(
            d, seed=cfg.random_seed, moves=cfg.macro_sa_moves,
            init_accept=cfg.macro_sa_init_accept, temp_ratio=cfg.macro_sa_temp_ratio,
            disp_weight=cfg.macro_sa_disp_weight, sink=sink)
        r = out["result"]
        if not (np.isfinite(r["overlap_after"]) and r["overlap_after"] <= 0):
            raise StageError("macro_overlap", "mLG", "residual macro overlap after legalization")
        d = d.copy()
        d.node_x[:n], d.node_y[:n] = out["x"], out["y"]
        d.node_flags = d.node_flags.copy()
        d.node_flags[:n] = out["flags"][:n]
        d.pin_offset_x = out["pin_offset_x"]
        d.meta = dict(d.meta, node_orient=out["orient"])
        check_geometry(d, "mLG", True)
        mx, my = d.node_x[mov][is_macro], d.node_y[mov][is_macro]
        disp = np.hypot(mx.astype(np.float64) - pre_x, my.astype(np.float64) - pre_y)
        if sink is not None:
            sink.set_summary("mlg_overflow_after", _overflow(d, cfg))
            sink.set_summary("mlg_macro_disp_mean", float(disp.mean()) if len(disp) else 0.0)
            sink.set_summary("mlg_macro_disp_max", float(disp.max()) if len(disp) else 0.0)
            sink.set_summary("mlg_overlap_before", r["overlap_before"])
            sink.set_summary("mlg_overlap_after", r["overlap_after"])
            sink.set_summary("mlg_hpwl", r["hpwl_after"])
            sink.set_summary("mlg_repaired", r["repaired"])
        fixed_x, fixed_y = mx.copy(), my.copy()
        # scatterFillers(db, cfg, 1): fresh filler positions around the fixed macros
        fresh = add_fillers(_without_fillers(d), cfg.gp.target_density, cfg.random_seed + 1)
        d = _with_positions(d, fresh.node_x, fresh.node_y)
        d, fo = _run_stage(d, cfg, 1, "FILLERONLY", sink)
        d, cgp = _run_stage(d, cfg, 2, "cGP", sink)
        moved = int(np.sum((d.node_x[mov][is_macro] != fixed_x) |
                           (d.node_y[mov][is_macro] != fixed_y)))
        if sink is not None:
            sink.set_summary("filleronly_iterations", fo["iterations"])
            sink.set_summary("cgp_hpwl", cgp["hpwl"])
            sink.set_summary("cgp_overflow", cgp["overflow"])
            sink.set_summary("cgp_iterations", cgp["iterations"])
            sink.set_summary("cgp_macros_moved", moved)
        if moved:
            raise StageError("macro_moved", "cGP", "a frozen macro changed position")
        final = cgp
    check_geometry(d, "gp/delivered", True)
    if sink is not None:
        sink.set_summary("gp_hpwl_before", before)
        sink.set_summary("gp_hpwl", final["hpwl"])
        sink.set_summary("gp_overflow", final["overflow"])
        sink.set_summary("gp_num_fillers", count)
        sink.set_summary("gp_movable_macros", movable_macros)
        sink.set_summary("gp_out_of_core", _out_of_core(d, d.total_nodes, 1e-2, skip_fixed=True))
        sink.set_summary("gp_time_ms", 1e3 * (time.time() - t0))
    return d


def lg_stage(d, cfg: FlowConfig, sink=None):
    """LegalizeStage::run: Abacus over the rows.  With movable macros not yet
    legalized it changes nothing and returns (d, None), as the reference
    only logs an error there."""
    mov = np.arange(d.num_movable)
    if np.any(((d.node_flags[mov] & F_MACRO) != 0) & ((d.node_flags[mov] & F_FIXED) == 0)):
        return d, None
    t0 = time.time()
    before = hpwl(d)
    x, y, res = legalize.legalize_cells(d, d.node_x[:d.num_nodes], d.node_y[:d.num_nodes],
                                        area_weight=cfg.lg_cell_weight == "area")
    nx, ny = d.node_x.copy(), d.node_y.copy()
    nx[:d.num_nodes], ny[:d.num_nodes] = x, y
    d = _with_positions(d, nx, ny)
    if sink is not None:
        sink.set_summary("lg_hpwl_before", before)
        sink.set_summary("lg_hpwl", hpwl(d))
        for k in ("cells", "failed", "disp_mean", "disp_max", "rows", "segments",
                  "selfcheck_overlaps"):
            sink.set_summary("lg_" + k, res[k])
        sink.set_summary("lg_time_ms", 1e3 * (time.time() - t0))
    return d, res


STAGES = {"init": init_stage, "gp": gp_stage, "lg": lg_stage}


def run(aux_path: str, out_dir: str, cfg: FlowConfig | None = None, name: str | None = None):
    """superplace-cli --stage <cfg.stages> on a BookShelf design: returns the
    final design and the path of summary.json."""
    cfg = cfg or FlowConfig()
    name = name or os.path.splitext(os.path.basename(aux_path))[0]
    out = os.path.join(out_dir, name)
    os.makedirs(out, exist_ok=True)
    summary = os.path.join(out, "summary.json")
    sink = metrics.MetricsSink(out)
    sink.set_summary("benchmark", name)
    sink.set_summary("aux_path", aux_path)
    try:
        t = time.time()
        d = bookshelf.read(aux_path)
        sink.record_duration("read", 1e3 * (time.time() - t))
        for stage in cfg.stages:
            if stage not in STAGES:
                raise ValueError(f"unknown stage '{stage}'")
            t = time.time()
            sink.begin_stage(stage)
            if stage == "gp":
                d = gp_stage(d, cfg, sink, out_prefix=os.path.join(out, name))
            elif stage == "lg":
                d, _ = lg_stage(d, cfg, sink)
            else:
                d = init_stage(d, cfg, sink)
            sink.end_stage(1e3 * (time.time() - t))
            bookshelf.write_pl(os.path.join(out, f"{name}.{stage}.pl"), d)
        sink.set_summary("run_status", "completed")
        sink.write_summary(summary)
        return d, summary
    except Exception as e:
        sink.set_summary("run_status", "failed")
        sink.set_summary("failure_message", str(e))
        sink.set_summary("failure_kind", getattr(e, "kind", "exception"))
        sink.set_summary("failure_stage", getattr(e, "stage", "pipeline"))
        try:
            sink.end_stage(0)
            sink.write_summary(summary)
        except Exception:
            pass
        raise
    finally:
        sink.close()
