"""Inputs of the Abacus legalization cases, shared by make_abacus_golden.py
(run through the reference's own legalizeCells) and tests/test_abacus.py
(run through sp_abacus_legalize).  Each case: a design, node centres,
per-row site steps (rows end at lx + step * whole sites) and options."""
import dataclasses
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _sited(d, step):
    """The design with rows ending on whole sites of width step."""
    step = np.full(d.num_rows, step, np.float32) if np.isscalar(step) else step.astype(np.float32)
    sites = np.floor((d.row_hx - d.row_lx) / step + 1e-6).astype(np.int32)
    hx = (d.row_lx.astype(np.float32) + step * sites.astype(np.float32)).astype(np.float32)
    core = (float(d.row_lx.min()), float(d.row_ly.min()), float(hx.max()),
            float((d.row_ly + d.row_height).max()))
    return dataclasses.replace(d, row_hx=hx, core=core), step, sites


def _jitter(d, seed, sx, sy):
    rng = np.random.default_rng(seed)
    n = d.num_nodes
    x = d.node_x[:n].astype(np.float32).copy()
    y = d.node_y[:n].astype(np.float32).copy()
    m = np.arange(n) < d.num_movable
    m &= (d.node_flags[:n] & 2) == 0
    x[m] += rng.normal(0, sx, m.sum()).astype(np.float32)
    y[m] += rng.normal(0, sy, m.sum()).astype(np.float32)
    lx, ly, hx, hy = d.core
    x = np.clip(x, lx, hx).astype(np.float32)
    y = np.clip(y, ly, hy).astype(np.float32)
    return x, y


def cases():
    from superplace_amd import bookshelf, design as D
    out = {}
    for name in ("fixture", "synth", "thin1"):
        d = bookshelf.read(os.path.join(HERE, "bookshelf", name, f"{name}.aux"))
        step = d.meta["row_step"]
        x, y = _jitter(d, 1, 4.0, 3.0)
        out[f"{name}_jitter"] = (d, x, y, step, 0)
        out[f"{name}_pl"] = (d, d.node_x[:d.num_nodes].copy(), d.node_y[:d.num_nodes].copy(), step, 0)
    # movable macros: Abacus refuses (precondition), as the reference does
    d = bookshelf.read(os.path.join(HERE, "bookshelf", "mac", "mac.aux"))
    out["mac_refused"] = (d, d.node_x[:d.num_nodes].copy(), d.node_y[:d.num_nodes].copy(),
                          d.meta["row_step"], 0)
    # the same with its macros fixed: they become obstacles
    f = d.node_flags.copy()
    f[(f & 1) != 0] |= 2
    dm = dataclasses.replace(d, node_flags=f)
    x, y = _jitter(dm, 2, 30.0, 20.0)
    out["mac_fixed"] = (dm, x, y, dm.meta["row_step"], 0)
    out["mac_fixed_area"] = (dm, x, y, dm.meta["row_step"], 2)
    out["mac_fixed_full"] = (dm, x, y, dm.meta["row_step"], 1)
    # a crowded synthetic design from a clustered start: 20k cells, 12 fixed macros
    s = D.synth_design(20000, 12, 21000, 1500.0, 0.75, init_sigma_frac=0.12, seed=8,
                       name="lg20k", with_fillers=False)
    sd, st, _ = _sited(s, np.float32(s.meta["site_pitch"]))
    out["synth20k"] = (sd, sd.node_x[:sd.num_nodes].copy(), sd.node_y[:sd.num_nodes].copy(), st, 0)
    return out


def row_sites(d, step):
    return np.round((d.row_hx - d.row_lx) / step).astype(np.int32)
