"""Inputs of the macro legalization cases, shared by make_mlg_golden.py
(run through the reference's own legalizeMacros) and tests/test_macro_lg.py
(run through sp_macro_legalize).  Each case: a design whose rows end on
whole sites, per-row site steps, and the macro_sa_* settings
(random_seed, moves, init_accept, temp_ratio, disp_weight).  Node names are
"n<i>" on both sides."""
import dataclasses
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT = dict(seed=1002, moves=0, init_accept=0.9, temp_ratio=1e-4, disp_weight=1.0)


def _named(d, step, orient=None, weights=None):
    n = d.num_nodes
    meta = dict(d.meta)
    meta["node_names"] = [f"n{i}" for i in range(n)]
    meta["row_step"] = np.asarray(step, np.float32)
    meta["node_orient"] = (np.zeros(d.total_nodes, np.uint8) if orient is None
                           else np.asarray(orient, np.uint8))
    if weights is not None:
        meta["net_weight"] = np.asarray(weights, np.float32)
    return dataclasses.replace(d, meta=meta)


def _sited(d, step):
    """The design with rows ending on whole sites of width step."""
    step = np.full(d.num_rows, step, np.float32)
    sites = np.floor((d.row_hx - d.row_lx) / step + 1e-6).astype(np.int32)
    hx = (d.row_lx.astype(np.float32) + step * sites.astype(np.float32)).astype(np.float32)
    core = (float(d.row_lx.min()), float(d.row_ly.min()), float(hx.max()),
            float((d.row_ly + d.row_height).max()))
    return dataclasses.replace(d, row_hx=hx, core=core), step


def dense(num_macros, num_cells, core, seed, site=2.0, rh=12.0, num_blocks=6):
    """Macros of 5..20 rows and 30..120 sites dropped uniformly (most of them
    overlap), standard cells, fixed blockages, zero-area NI terminals,
    2..5-pin nets over cells and macros with pins anywhere on their nodes,
    net weights in [0.5, 2] and mixed orientations."""
    from superplace_amd import design as D
    g = np.random.default_rng(seed)
    mw = g.integers(30, 121, num_macros) * site
    mh = g.integers(5, 21, num_macros) * rh
    cw = g.integers(2, 9, num_cells) * site
    bw = g.integers(20, 60, num_blocks) * site
    bh = g.integers(4, 12, num_blocks) * rh
    w = np.concatenate([cw, mw, bw, np.zeros(4)]).astype(np.float32)
    h = np.concatenate([np.full(num_cells, rh), mh, bh, np.zeros(4)]).astype(np.float32)
    nmov = num_cells + num_macros
    x = g.uniform(0.1 * core, 0.9 * core, len(w)).astype(np.float32)
    y = g.uniform(0.1 * core, 0.9 * core, len(w)).astype(np.float32)
    x[-4:] = [0.0, core, 0.0, core]
    y[-4:] = [0.0, 0.0, core, core]
    flags = np.concatenate([np.zeros(num_cells), np.full(num_macros, D.F_MACRO),
                            np.full(num_blocks, D.F_MACRO | D.F_FIXED),
                            np.full(4, D.F_FIXED | D.F_NI)]).astype(np.uint8)
    n = len(w)
    pin2node, pin2net, ox, oy = [], [], [], []
    nets = 0
    for _ in range(int(1.2 * nmov)):
        deg = int(g.integers(2, 6))
        nodes = g.choice(n, deg, replace=False)
        if g.uniform() < 0.5:   # most nets touch a macro
            nodes[0] = num_cells + int(g.integers(0, num_macros))
            nodes = np.unique(nodes)
        for nd in nodes:
            pin2node.append(int(nd))
            pin2net.append(nets)
            ox.append(float(g.uniform(-0.5, 0.5) * w[nd]))
            oy.append(float(g.uniform(-0.5, 0.5) * h[nd]))
        nets += 1
    nrows = int(core / rh)
    rows = (np.arange(nrows) * rh, np.full(nrows, rh), np.zeros(nrows), np.full(nrows, core))
    d = D.make_design(x, y, w, h, flags, nmov, pin2node, pin2net, ox, oy, nets, rows,
                      name=f"dense-{num_macros}")
    d, step = _sited(d, site)
    orient = g.choice(np.array([0, 2, 4, 6, 1], np.uint8), n)
    return _named(d, step, orient=orient, weights=g.uniform(0.5, 2.0, nets))


def spec_design(size, rh, specs, nets):
    """The fixture builder of the reference's macro SA tests
    (test/test_macro_sa.cpp:26-63): core [0,size]^2 of rows rh with unit
    sites, nodes (cx, cy, w, h, flags) movable first then fixed, each in
    spec order, nets of node lists with zero pin offsets, unit weights."""
    from superplace_amd import design as D
    order = [i for i, sp in enumerate(specs) if not sp[4] & D.F_FIXED] + \
            [i for i, sp in enumerate(specs) if sp[4] & D.F_FIXED]
    remap = {old: new for new, old in enumerate(order)}
    nmov = sum(1 for sp in specs if not sp[4] & D.F_FIXED)
    cols = [np.array([specs[i][k] for i in order], np.float32) for k in range(4)]
    flags = np.array([specs[i][4] for i in order], np.uint8)
    pin2node = [remap[i] for net in nets for i in net]
    pin2net = [k for k, net in enumerate(nets) for _ in net]
    zeros = [0.0] * len(pin2node)
    nrows = int(size / rh)
    rows = (np.arange(nrows) * rh, np.full(nrows, rh), np.zeros(nrows), np.full(nrows, size))
    d = D.make_design(*cols, flags, nmov, pin2node, pin2net, zeros, zeros, len(nets), rows,
                      name="spec")
    d, step = _sited(d, 1.0)
    return _named(d, step)


def reference_fixtures():
    """The node layouts of the reference's own macro SA tests
    (test/test_macro_sa.cpp:86-357), with its settings (seed 1002, 30000
    moves; 1 move where the test exercises only the repairs)."""
    from superplace_amd.design import F_MACRO as M, F_FIXED as F
    cfg = dict(DEFAULT, moves=30000)
    repair = dict(DEFAULT, moves=1)
    out = {}
    chain16 = [[i, i + 1] for i in range(15)]
    specs = [(500.0, 500.0, 90.0, 90.0, M)] * 16 + \
        [(950.0 if i % 2 else 50.0, 950.0 if i // 2 else 50.0, 80.0, 80.0, F) for i in range(4)]
    out["spec_stacked16"] = (spec_design(1000.0, 10.0, specs, chain16), cfg)
    specs = [(100.0 + i * 5.0, 100.0 + i * 5.0, 120.0, 120.0, M) for i in range(12)]
    out["spec_aligned12"] = (spec_design(1000.0, 10.0, specs, [[i, i + 1] for i in range(11)]),
                             cfg)
    specs = [(300.0, 300.0, 100.0, 100.0, M)] * 8
    out["spec_fixed8"] = (spec_design(1000.0, 10.0, specs, [[0, 1], [2, 3], [4, 5], [6, 7]]), cfg)
    specs = [(150.0 + (i % 3) * 300.0, 150.0 + (i // 3) * 300.0, 100.0, 100.0, M)
             for i in range(9)]
    out["spec_grid9"] = (spec_design(1000.0, 10.0, specs, [[i, i + 1] for i in range(8)]), cfg)
    specs = [(400.0 + i * 8.0, 400.0, 80.0, 80.0, M) for i in range(10)] + \
        [(50.0 + i * 45.0, 900.0, 4.0, 10.0, 0) for i in range(20)]
    nets = [[i, 10 + i, 10 + (i + 5) % 20] for i in range(10)]
    out["spec_static"] = (spec_design(1000.0, 10.0, specs, nets), cfg)
    ly = [800.0, 630.0, 430.0, 230.0, 30.0]
    specs = [(200.0, y + 100.0, 400.0, 200.0, M) for y in ly] + [(700.0, 500.0, 600.0, 1000.0, F)]
    out["spec_chain"] = (spec_design(1000.0, 10.0, specs, [[i, i + 1] for i in range(4)]), repair)
    specs = [(1500.0, 1500.0, 200.0, 80.0, M)] + \
        [(600.0 + 200.0 * i, 600.0 + 200.0 * j, 140.0, 140.0, F) for i in range(10)
         for j in range(10)]
    out["spec_array"] = (spec_design(3000.0, 10.0, specs, [[0, 1]]), repair)
    for seed in (1002, 1, 2, 3, 4):
        specs = [(5000.0, 5000.0, 2000.0, 1000.0, M), (5000.0, 5990.0, 2000.0, 1000.0, F)]
        out[f"spec_tie{seed}"] = (spec_design(10000.0, 10.0, specs, []), dict(repair, seed=seed))
    ly = [7000.0, 6030.0, 5030.0]
    specs = [(1000.0, y + 500.0, 2000.0, 1000.0, M) for y in ly] + \
        [(1000.0, 2500.0, 2000.0, 5000.0, F), (2250.0, 4000.0, 500.0, 8000.0, F)]
    out["spec_shift"] = (spec_design(8000.0, 10.0, specs, [[0, 1], [1, 2]]), repair)
    specs = [(500.0, 500.0, 100.0, 100.0, M)] * 14
    out["spec_determinism"] = (spec_design(1000.0, 10.0, specs, [[i, i + 1] for i in range(13)]),
                               cfg)
    return out


def cases():
    from superplace_amd import bookshelf, design as D
    out = reference_fixtures()
    # the reference's mGP fixture shape: macros clustered on the centre
    f = D.fixture_eplace(200, 6, seed=3)
    f, st = _sited(f, 1.0)
    out["fixture6"] = (_named(f, st), DEFAULT)
    f = D.fixture_eplace(400, 24, seed=5)
    f, st = _sited(f, 1.0)
    out["fixture24"] = (_named(f, st), DEFAULT)
    out["fixture24_fast"] = (_named(f, st), dict(DEFAULT, seed=7, moves=3000, init_accept=0.5,
                                                   temp_ratio=1e-2, disp_weight=0.0))
    # BookShelf: four movable macros pushed onto each other
    b = bookshelf.read(os.path.join(HERE, "bookshelf", "mac", "mac.aux"))
    m = np.nonzero(((b.node_flags & D.F_MACRO) != 0)[:b.num_movable])[0]
    x = b.node_x.copy()
    y = b.node_y.copy()
    x[m] = np.float32(b.core[0] + 0.5 * (b.core[2] - b.core[0])) + np.arange(len(m)) * 23.0
    y[m] = np.float32(b.core[1] + 0.5 * (b.core[3] - b.core[1])) + np.arange(len(m)) * 13.0
    bm = dataclasses.replace(b, node_x=x, node_y=y)
    out["mac_stacked"] = (_named(bm, b.meta["row_step"], orient=b.meta["node_orient"],
                                 weights=b.meta["net_weight"]), DEFAULT)
    # no movable macros: nothing to do
    f = D.fixture_eplace(100, 0, seed=3)
    f, st = _sited(f, 1.0)
    out["no_macros"] = (_named(f, st), DEFAULT)
    # dense random floorplans
    out["dense60"] = (dense(60, 600, 1500.0, seed=11), DEFAULT)
    out["dense150"] = (dense(150, 1500, 2400.0, seed=12), DEFAULT)
    out["dense150_s2"] = (dense(150, 1500, 2400.0, seed=12), dict(DEFAULT, seed=99))
    out["dense400"] = (dense(400, 2000, 3600.0, seed=13, site=1.0), DEFAULT)
    out["sparse3000"] = (dense(3000, 3000, 20400.0, seed=14), DEFAULT)
    # above 8000 macros the index self-check is skipped
    out["sparse9000"] = (dense(9000, 1000, 60000.0, seed=15), dict(DEFAULT, moves=200000))
    return out


def row_sites(d, step):
    return np.round((d.row_hx - d.row_lx) / step).astype(np.int32)
