"""Placement database (pure SoA + CSR) and seeded synthetic designs.

The layout follows the reference's PlaceDB (src/db/place_db.h:27-71): node
centre coordinates, node order [movable | fixed | filler], pins as node centre
plus offset, and two CSRs (net -> pins, node -> pins) built by the two-pass
rule of PlaceDB::finalizeCSR (src/db/place_db.cpp:67-95).

Synthetic netlists use a portable counter-based RNG (splitmix64) so that the
same seed gives the same design on every machine.  Fillers are planned and
scattered by the library exactly as the reference does (its Rng,
src/util/rng.h:15-18: std::mt19937 with libstdc++'s
uniform_real_distribution<float>).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

F_MACRO, F_FIXED, F_FILLER, F_NI = 1, 2, 4, 8   # src/db/place_db.h:17-22

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


class Rng:
    """splitmix64 stream: value i of the stream is mix(seed + (i+1)*golden)."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.counter = 0

    def _u64(self, n: int) -> np.ndarray:
        idx = np.arange(self.counter + 1, self.counter + n + 1, dtype=np.uint64)
        self.counter += n
        with np.errstate(over="ignore"):
            z = self.seed + idx * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int, lo=0.0, hi=1.0) -> np.ndarray:
        """float64 in [lo, hi): 53 high bits of each draw."""
        u = (self._u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def integers(self, n: int, lo: int, hi: int) -> np.ndarray:
        """int64 in [lo, hi] (inclusive), by floor of a uniform."""
        u = self.uniform(n)
        return np.minimum(lo + np.floor(u * (hi - lo + 1)).astype(np.int64), hi)


@dataclasses.dataclass
class Design:
    num_movable: int
    num_nodes: int
    num_fillers: int
    node_x: np.ndarray          # float32[total], centre
    node_y: np.ndarray
    node_w: np.ndarray
    node_h: np.ndarray
    node_flags: np.ndarray      # uint8[total]
    pin_offset_x: np.ndarray    # float32[P], relative to node centre
    pin_offset_y: np.ndarray
    pin2node: np.ndarray        # int32[P]
    pin2net: np.ndarray         # int32[P]
    net2pin_start: np.ndarray   # int32[E+1]
    flat_net2pin: np.ndarray    # int32[P]
    node2pin_start: np.ndarray  # int32[total+1]
    flat_node2pin: np.ndarray   # int32[P]
    row_ly: np.ndarray          # float32[R]
    row_height: np.ndarray
    row_lx: np.ndarray
    row_hx: np.ndarray
    core: tuple                 # (lx, ly, hx, hy), the rows' bounding box
    name: str = "design"
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def total_nodes(self) -> int:
        return self.num_nodes + self.num_fillers

    @property
    def num_pins(self) -> int:
        return int(self.pin2node.shape[0])

    @property
    def num_nets(self) -> int:
        return int(self.net2pin_start.shape[0] - 1)

    @property
    def num_rows(self) -> int:
        return int(self.row_ly.shape[0])

    @property
    def row_height_mode(self) -> float:
        """Most frequent row height (PlaceDB::computeRegions, place_db.cpp:97-110)."""
        if self.num_rows == 0:
            return 0.0
        vals, counts = np.unique(self.row_height, return_counts=True)
        return float(vals[np.argmax(counts)])

    @property
    def total_row_area(self) -> float:
        return float(np.sum((self.row_hx.astype(np.float64) - self.row_lx)
                            * self.row_height.astype(np.float64)))

    def copy(self) -> "Design":
        return dataclasses.replace(
            self, **{f.name: (getattr(self, f.name).copy()
                              if isinstance(getattr(self, f.name), np.ndarray)
                              else getattr(self, f.name))
                     for f in dataclasses.fields(self)})


def finalize_csr(pin2node: np.ndarray, pin2net: np.ndarray, num_nets: int, num_nodes: int):
    """Both CSRs, pins listed in increasing pin index inside each bucket
    (PlaceDB::finalizeCSR, src/db/place_db.cpp:67-95)."""
    pin2node = np.asarray(pin2node, dtype=np.int32)
    pin2net = np.asarray(pin2net, dtype=np.int32)
    net2pin_start = np.zeros(num_nets + 1, dtype=np.int32)
    np.cumsum(np.bincount(pin2net, minlength=num_nets), out=net2pin_start[1:])
    flat_net2pin = np.argsort(pin2net, kind="stable").astype(np.int32)
    node2pin_start = np.zeros(num_nodes + 1, dtype=np.int32)
    np.cumsum(np.bincount(pin2node, minlength=num_nodes), out=node2pin_start[1:])
    flat_node2pin = np.argsort(pin2node, kind="stable").astype(np.int32)
    return net2pin_start, flat_net2pin, node2pin_start, flat_node2pin


def make_design(node_x, node_y, node_w, node_h, node_flags, num_movable, pin2node, pin2net,
                pin_offset_x, pin_offset_y, num_nets, rows, name="design", meta=None) -> Design:
    """Assemble a Design (no fillers yet).  rows = (ly, height, lx, hx) arrays."""
    n = len(node_x)
    n2p, fn2p, nd2p, fnd2p = finalize_csr(pin2node, pin2net, num_nets, n)
    row_ly, row_h, row_lx, row_hx = (np.asarray(r, dtype=np.float32) for r in rows)
    core = (float(row_lx.min()), float(row_ly.min()), float(row_hx.max()),
            float((row_ly + row_h).max()))
    return Design(num_movable=int(num_movable), num_nodes=n, num_fillers=0,
                  node_x=np.asarray(node_x, np.float32), node_y=np.asarray(node_y, np.float32),
                  node_w=np.asarray(node_w, np.float32), node_h=np.asarray(node_h, np.float32),
                  node_flags=np.asarray(node_flags, np.uint8),
                  pin_offset_x=np.asarray(pin_offset_x, np.float32),
                  pin_offset_y=np.asarray(pin_offset_y, np.float32),
                  pin2node=np.asarray(pin2node, np.int32), pin2net=np.asarray(pin2net, np.int32),
                  net2pin_start=n2p, flat_net2pin=fn2p, node2pin_start=nd2p, flat_node2pin=fnd2p,
                  row_ly=row_ly, row_height=row_h, row_lx=row_lx, row_hx=row_hx, core=core,
                  name=name, meta=dict(meta or {}))


# ----------------------------------------------------------------------------- fillers
def filler_plan(d: Design, target_density: float):
    """Filler count and size (initializeFillers, src/gp/filler.cpp:47-121),
    computed by the library (sp_filler_plan) on the nodes below num_nodes.

    Returns (count, width, height, whitespace, node_area_scaled, filler_area)."""
    from . import hip
    import ctypes as C
    info = hip.SpFillerInfo()
    da = hip._DesignArrays(dataclasses.replace(d, num_fillers=0))
    hip.check(hip.lib().sp_filler_plan(C.byref(da.s), float(target_density), C.byref(info)))
    return (int(info.count), float(info.width), float(info.height), float(info.whitespace_area),
            float(info.node_area_scaled), float(info.total_filler_area))


def add_fillers(d: Design, target_density: float, seed: int) -> Design:
    """Append fillers (filler.cpp:122-132) with centres drawn in the core by
    the reference's generator (scatterFillers, filler.cpp:33-45: std::mt19937
    seeded with `seed`), through the library (sp_filler_scatter)."""
    from . import hip
    import ctypes as C
    count, w, h, *_ = filler_plan(d, target_density)
    fx = np.zeros(max(count, 1), np.float32)
    fy = np.zeros(max(count, 1), np.float32)
    da = hip._DesignArrays(dataclasses.replace(d, num_fillers=0))
    hip.check(hip.lib().sp_filler_scatter(C.byref(da.s), count, w, h, int(seed) & (2**64 - 1),
                                          hip._ptr(fx), hip._ptr(fy)))
    fx, fy = fx[:count], fy[:count]
    out = d.copy()
    if count == 0:
        return out
    out.node_x = np.concatenate([d.node_x[:d.num_nodes], fx.astype(np.float32)])
    out.node_y = np.concatenate([d.node_y[:d.num_nodes], fy.astype(np.float32)])
    out.node_w = np.concatenate([d.node_w[:d.num_nodes], np.full(count, w, np.float32)])
    out.node_h = np.concatenate([d.node_h[:d.num_nodes], np.full(count, h, np.float32)])
    out.node_flags = np.concatenate([d.node_flags[:d.num_nodes],
                                     np.full(count, F_FILLER, np.uint8)])
    out.num_fillers = count
    out.node2pin_start = np.concatenate([d.node2pin_start[:d.num_nodes + 1],
                                         np.full(count, d.num_pins, np.int32)]).astype(np.int32)
    out.meta = dict(d.meta, fillers=count, filler_w=w, filler_h=h)
    return out


def mt_uniform(seed: int, count: int, lo: float, hi: float) -> np.ndarray:
    """`count` successive Rng::uniform(lo, hi) draws of the reference's
    generator (src/util/rng.h:15-18: std::mt19937 seeded with `seed`,
    uniform_real_distribution<float>), through the library (sp_rng_uniform)."""
    from . import hip
    out = np.zeros(max(count, 1), np.float32)
    hip.check(hip.lib().sp_rng_uniform(int(seed) & (2**64 - 1), int(count), float(lo), float(hi),
                                       hip._ptr(out)))
    return out[:count]


# ----------------------------------------------------------------------------- generators
def _degree_tail_powerlaw(mean_tail: float, dmin: int, dmax: int) -> float:
    """Exponent a of p(d) ~ d^-a on [dmin, dmax] whose mean is mean_tail (bisection)."""
    d = np.arange(dmin, dmax + 1, dtype=np.float64)

    def mean(a):
        p = d ** -a
        return float((d * p).sum() / p.sum())

    lo, hi = 0.0, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mean(mid) > mean_tail:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def _sample_degrees(rng: Rng, n: int, tail_probs: np.ndarray, dmin_tail: int) -> np.ndarray:
    """P(2)=0.55, P(3)=0.2, remaining 0.25 over the tail distribution."""
    probs = np.concatenate([[0.55, 0.20], 0.25 * tail_probs / tail_probs.sum()])
    degs = np.concatenate([[2, 3], np.arange(dmin_tail, dmin_tail + len(tail_probs))])
    cdf = np.cumsum(probs)
    cdf[-1] = 1.0
    u = rng.uniform(n)
    return degs[np.searchsorted(cdf, u, side="right")].astype(np.int64)


def _fit_pin_count(degrees: np.ndarray, target: int, dmin_tail: int, dmax: int) -> np.ndarray:
    """The sampled degrees moved to exactly `target` pins: one pin more (or
    less) per tail net (degree >= dmin_tail), in net order, round after
    round, within [dmin_tail, dmax]."""
    deg = degrees.copy()
    diff = int(target) - int(deg.sum())
    tail = np.flatnonzero(deg >= dmin_tail)
    while diff != 0:
        step = 1 if diff > 0 else -1
        cand = tail[deg[tail] < dmax] if step > 0 else tail[deg[tail] > dmin_tail]
        if len(cand) == 0:
            raise ValueError("pin count out of reach of the degree range")
        k = min(abs(diff), len(cand))
        deg[cand[:k]] += step
        diff -= step * k
    return deg


def _netlist(rng: Rng, degrees: np.ndarray, num_movable: int, num_fixed: int,
             fixed_pin_frac: float, locality: float):
    """Pins of each net: a seed cell plus cells at two-sided geometric index
    offsets around it (netlist locality), a fraction of pins on fixed macros."""
    num_nets = len(degrees)
    pin2net = np.repeat(np.arange(num_nets, dtype=np.int32), degrees)
    P = len(pin2net)
    seed_cell = rng.integers(num_nets, 0, num_movable - 1)
    first = np.zeros(P, dtype=bool)
    starts = np.concatenate([[0], np.cumsum(degrees)[:-1]])
    first[starts] = True
    scale = locality * np.sqrt(degrees[pin2net].astype(np.float64))
    u = rng.uniform(P)
    sign = np.where(rng.uniform(P) < 0.5, -1, 1)
    off = np.floor(-scale * np.log1p(-u)).astype(np.int64) + 1
    cell = (seed_cell[pin2net] + sign * off) % num_movable
    cell = np.where(first, seed_cell[pin2net], cell)
    if num_fixed > 0:
        to_fixed = (rng.uniform(P) < fixed_pin_frac) & ~first
        fixed_pick = num_movable + rng.integers(P, 0, num_fixed - 1)
        cell = np.where(to_fixed, fixed_pick, cell)
    return cell.astype(np.int32), pin2net


def _place_fixed_macros(rng: Rng, count: int, die: float, side_lo: float, side_hi: float,
                        max_tries: int = 200):
    """Uniform random non-overlapping fixed macros inside [0, die]^2."""
    ws, hs, xs, ys = [], [], [], []
    for _ in range(count):
        w = float(rng.uniform(1, side_lo, side_hi)[0]) * die
        h = float(rng.uniform(1, side_lo, side_hi)[0]) * die
        for _t in range(max_tries):
            x = float(rng.uniform(1, 0.5 * w, die - 0.5 * w)[0])
            y = float(rng.uniform(1, 0.5 * h, die - 0.5 * h)[0])
            ok = True
            for (ox, oy, ow, oh) in zip(xs, ys, ws, hs):
                if abs(x - ox) < 0.5 * (w + ow) and abs(y - oy) < 0.5 * (h + oh):
                    ok = False
                    break
            if ok:
                ws.append(w); hs.append(h); xs.append(x); ys.append(y)
                break
    return (np.array(xs, np.float32), np.array(ys, np.float32),
            np.array(ws, np.float32), np.array(hs, np.float32))


def _cluster_init(rng: Rng, n: int, cx: float, cy: float, sigma: float):
    """Clustered start (sum of three uniforms ~ normal with the given sigma)."""
    a = sigma * 2.0   # sum of 3 U(-1,1) has variance 1 -> scale 1
    ux = (rng.uniform(n, -1, 1) + rng.uniform(n, -1, 1) + rng.uniform(n, -1, 1)) / math.sqrt(1.0)
    uy = (rng.uniform(n, -1, 1) + rng.uniform(n, -1, 1) + rng.uniform(n, -1, 1)) / math.sqrt(1.0)
    return cx + 0.5 * a * ux, cy + 0.5 * a * uy


def synth_design(num_cells: int, num_fixed: int, num_nets: int, die: float, util: float,
                 row_height: float = 12.0, tail: str = "geometric", tail_max: int = 32,
                 fixed_frac: float = 0.10, seed: int = 1002, init_sigma_frac: float = 0.05,
                 locality: float = 8.0, fixed_pin_frac: float = 0.01, name: str = "synth",
                 target_density: float = 1.0, with_fillers: bool = True,
                 num_macros: int = 0, macro_area_frac: float = 0.0,
                 num_pins: int | None = None) -> Design:
    """Seeded synthetic standard-cell design (the generator behind the
    benchmark workloads): num_cells movable cells of width U{1..8} sites and
    one row height, num_fixed fixed macros covering about fixed_frac of the
    die, num_nets nets with P(2)=0.55, P(3)=0.2 and a geometric (config-1
    style) or power-law tail.  The site pitch is chosen so that
    movable / (row area - fixed area) equals util (the reference's
    utilisation definition, src/db/db_stats.h:75-78).

    num_macros movable macros (after the cells in node order) together take
    macro_area_frac of the movable area: areas U(0.5, 1.5) times the mean,
    aspect ratios U(0.5, 2), heights rounded to whole rows and at least two
    rows (the reader's macro rule is taller than 1.5 rows), widths keeping
    the drawn area.

    num_pins: the sampled degrees are moved to exactly this many pins
    (_fit_pin_count) before the pins are placed."""
    rng = Rng(seed)
    nrows = int(die // row_height)
    rows = (np.arange(nrows, dtype=np.float32) * row_height,
            np.full(nrows, row_height, np.float32), np.zeros(nrows, np.float32),
            np.full(nrows, die, np.float32))
    core_h = nrows * row_height
    # fixed macros: side fraction chosen so that E[sum area] ~ fixed_frac * die^2
    if num_fixed > 0:
        mean_sq = fixed_frac / num_fixed
        s_hi = math.sqrt(3.0 * mean_sq / (0.25 + 0.5 + 1.0)) if mean_sq > 0 else 0.0
        fx, fy, fw, fh = _place_fixed_macros(rng, num_fixed, min(die, core_h), 0.5 * s_hi, s_hi)
    else:
        fx = fy = fw = fh = np.zeros(0, np.float32)
    n_fixed = len(fx)
    fixed_area = float(np.sum(fw.astype(np.float64) * fh))
    row_area = float(nrows) * row_height * die
    sites = rng.integers(num_cells, 1, 8).astype(np.float64)
    movable_area = util * (row_area - fixed_area)
    cell_area = movable_area * (1.0 - macro_area_frac) if num_macros > 0 else movable_area
    pitch = cell_area / (num_cells * row_height * sites.mean())
    cw = (sites * pitch).astype(np.float32)
    ch = np.full(num_cells, row_height, np.float32)
    if num_macros > 0:
        rel = rng.uniform(num_macros, 0.5, 1.5)
        area = rel / rel.sum() * movable_area * macro_area_frac
        aspect = rng.uniform(num_macros, 0.5, 2.0)
        mh = np.maximum(np.round(np.sqrt(area / aspect) / row_height), 2.0) * row_height
        mw = area / mh
        cw = np.concatenate([cw, mw.astype(np.float32)])
        ch = np.concatenate([ch, mh.astype(np.float32)])
    n_mov = num_cells + num_macros
    cx, cy = _cluster_init(rng, n_mov, 0.5 * die, 0.5 * core_h, init_sigma_frac * die)
    # degrees
    if tail == "geometric":
        q = 0.7
        tail_probs = q ** np.arange(0, tail_max - 4 + 1, dtype=np.float64)
        a = None
    else:
        a = _degree_tail_powerlaw((4.27 - 0.55 * 2 - 0.2 * 3) / 0.25, 4, tail_max)
        tail_probs = np.arange(4, tail_max + 1, dtype=np.float64) ** -a
    degrees = _sample_degrees(rng, num_nets, tail_probs, 4)
    if num_pins is not None:
        degrees = _fit_pin_count(degrees, num_pins, 4, tail_max)
    pin2node, pin2net = _netlist(rng, degrees, n_mov, n_fixed, fixed_pin_frac, locality)
    node_w = np.concatenate([cw, fw])
    node_h = np.concatenate([ch, fh])
    offx = (rng.uniform(len(pin2node), -0.5, 0.5) * node_w[pin2node]).astype(np.float32)
    offy = (rng.uniform(len(pin2node), -0.5, 0.5) * node_h[pin2node]).astype(np.float32)
    flags = np.concatenate([np.zeros(num_cells, np.uint8), np.full(num_macros, F_MACRO, np.uint8),
                            np.full(n_fixed, F_FIXED, np.uint8)])
    d = make_design(np.concatenate([cx, fx]), np.concatenate([cy, fy]), node_w, node_h, flags,
                    n_mov, pin2node, pin2net, offx, offy, num_nets, rows, name=name,
                    meta=dict(seed=seed, die=die, util=util, site_pitch=float(pitch),
                              fixed_macros=n_fixed, fixed_area_frac=fixed_area / (die * core_h),
                              tail=tail, tail_exponent=a, tail_max=tail_max,
                              pins=int(len(pin2node))))
    if with_fillers:
        d = add_fillers(d, target_density, seed + 1)
    return d


def config_small(seed: int = 1002, with_fillers: bool = True) -> Design:
    """2,000 cells (width U{1..8} sites, site 1, height 12), 8 fixed macros,
    2,200 nets with a geometric degree tail capped at 32, square die at 70%
    utilisation (128x128 bins in the workload)."""
    return synth_design(2000, 8, 2200, die=424.0, util=0.70, tail="geometric", tail_max=32,
                        fixed_frac=0.14, seed=seed, name="synth-2k", with_fillers=with_fillers,
                        locality=6.0)


def config_mixed(seed: int = 1002, with_fillers: bool = True) -> Design:
    """Mixed-size: the 2,000-cell design above plus 63 movable macros
    (aspect 0.5-2, together 56.9% of the movable area), 70% utilisation
    counting the macros (1024x1024 bins in the workload)."""
    return synth_design(2000, 8, 2200, die=424.0, util=0.70, tail="geometric", tail_max=32,
                        fixed_frac=0.14, seed=seed, name="synth-mixed", with_fillers=with_fillers,
                        locality=6.0, num_macros=63, macro_area_frac=0.569)


def config_adaptec1(seed: int = 1002, with_fillers: bool = True) -> Design:
    """adaptec1-shaped: 211k cells, 543 fixed macros, 221k nets with 944k
    pins (power-law degree tail up to 500), die 11,600 with rows of 12, 57%
    utilisation (512x512 bins in the workload)."""
    return synth_design(211_000, 543, 221_000, die=11600.0, util=0.57, tail="powerlaw",
                        tail_max=500, fixed_frac=0.10, seed=seed, name="synth-adaptec1",
                        with_fillers=with_fillers, locality=12.0, num_pins=944_000)


_FIX_CORE, _FIX_ROW = 1200.0, 12.0


def _fixture_netlist(w, h, n_mov: int, n_nodes: int):
    """3-pin nets by the fixtures' rule (test/test_eplace.cpp:80-94): for each
    movable i, nodes a=i, b=(7i+3) mod n_mov, c=(13i+11) mod n_nodes, skipped
    when two coincide; pins at (+w/4, -h/4) of their node."""
    i = np.arange(n_mov)
    a, b, c = i, (i * 7 + 3) % n_mov, (i * 13 + 11) % n_nodes
    keep = (a != b) & (b != c) & (a != c)
    nodes = np.stack([a[keep], b[keep], c[keep]], axis=1).reshape(-1)
    nets = int(keep.sum())
    pin2net = np.repeat(np.arange(nets), 3)
    offx = (np.float32(0.25) * w[nodes]).astype(np.float32)
    offy = (np.float32(-0.25) * h[nodes]).astype(np.float32)
    return nodes, pin2net, offx, offy, nets


def _fixture_rows():
    nrows = int(_FIX_CORE / _FIX_ROW)
    return (np.arange(nrows) * _FIX_ROW, np.full(nrows, _FIX_ROW), np.zeros(nrows),
            np.full(nrows, _FIX_CORE))


def fixture_eplace(num_cells: int, num_macros: int, seed: int = 7) -> Design:
    """The reference's mGP integration fixture (makeFixture,
    test/test_eplace.cpp:45-113), node for node: a 1200x1200 core of rows 12;
    cells 8x12, then macros 100x108, at the core centre plus
    Rng(seed).uniform(-20, 20) for x then y of each node in order; four
    60x60 fixed terminals at (40 | 1160, 40 | 1160); 3-pin nets by the
    fixture rule."""
    n_mov = num_cells + num_macros
    w = np.concatenate([np.full(num_cells, 8.0), np.full(num_macros, 100.0), np.full(4, 60.0)])
    h = np.concatenate([np.full(num_cells, _FIX_ROW), np.full(num_macros, 9 * _FIX_ROW),
                        np.full(4, 60.0)]).astype(np.float32)
    w = w.astype(np.float32)
    u = mt_uniform(seed, 2 * n_mov, -20.0, 20.0)
    centre = np.float32(0.5 * _FIX_CORE)
    x = np.concatenate([centre + u[0::2], [40.0, _FIX_CORE - 40.0, 40.0, _FIX_CORE - 40.0]])
    y = np.concatenate([centre + u[1::2], [40.0, 40.0, _FIX_CORE - 40.0, _FIX_CORE - 40.0]])
    flags = np.concatenate([np.zeros(num_cells, np.uint8), np.full(num_macros, F_MACRO, np.uint8),
                            np.full(4, F_FIXED, np.uint8)])
    nodes, pin2net, offx, offy, nets = _fixture_netlist(w, h, n_mov, n_mov + 4)
    return make_design(x, y, w, h, flags, n_mov, nodes, pin2net, offx, offy, nets,
                       _fixture_rows(), name=f"fixture-{num_cells}-{num_macros}")


def fixture_overfull() -> Design:
    """The reference's overfull fixture (makeOverfullFixture,
    test/test_eplace.cpp:123-180): fixture_eplace(3000, 0) with 40 movable
    macros of 208x208 appended to the movable nodes (centre plus
    Rng(11).uniform(-20, 20), x then y), so that the movable area is about
    1.4 times the core's; nets by the fixture rule over the new node order."""
    base = fixture_eplace(3000, 0)
    n = base.num_movable
    u = mt_uniform(11, 80, -20.0, 20.0)
    centre = np.float32(0.5 * _FIX_CORE)
    cat = np.concatenate
    x = cat([base.node_x[:n], centre + u[0::2], base.node_x[n:]])
    y = cat([base.node_y[:n], centre + u[1::2], base.node_y[n:]])
    w = cat([base.node_w[:n], np.full(40, 208.0, np.float32), base.node_w[n:]])
    h = cat([base.node_h[:n], np.full(40, 208.0, np.float32), base.node_h[n:]])
    flags = cat([base.node_flags[:n], np.full(40, F_MACRO, np.uint8), base.node_flags[n:]])
    nm = n + 40
    nodes, pin2net, offx, offy, nets = _fixture_netlist(w, h, nm, nm + 4)
    return make_design(x, y, w, h, flags, nm, nodes, pin2net, offx, offy, nets, _fixture_rows(),
                       name="fixture-overfull")


def with_fillers(d: Design, target_density: float = 1.0, seed: int = 1002) -> Design:
    return add_fillers(d, target_density, seed)


def auto_bin_dim(d: Design, lo: int = 4, hi: int = 1024) -> int:
    """The reference's automatic bin dimension (bin_grid.cpp:172-189: the
    largest power of two whose bins hold one average movable node at target
    density 1), clamped to [lo, hi]."""
    mov = np.arange(d.num_movable)
    keep = (d.node_flags[mov] & (F_FIXED | F_NI)) == 0
    area = float(np.sum(d.node_w[mov][keep].astype(np.float64) * d.node_h[mov][keep]))
    ideal = (d.core[2] - d.core[0]) * (d.core[3] - d.core[1]) / max(area / max(d.num_movable, 1), 1e-9)
    dim = 4
    while dim * 2 <= 1024 and (dim * 2) ** 2 <= ideal:
        dim *= 2
    return min(max(dim, lo), hi)


def job_cells(count: int, lo: int, hi: int, seed0: int, extremes: bool = False):
    """Cell counts log-uniform in [lo, hi], drawn from the splitmix64 stream of
    seed 1002 + seed0; with `extremes`, the smallest draw becomes lo and the
    largest hi."""
    u = Rng(1002 + seed0).uniform(count)
    cells = [int(round(lo * (hi / lo) ** float(x))) for x in u]
    if extremes and count >= 2:
        cells[int(np.argmin(u))] = lo
        cells[int(np.argmax(u))] = hi
    return cells


def job_designs(count: int, lo: int, hi: int, seed0: int, dim_lo: int, dim_hi: int,
                nets_per_cell: float = 1.05, extremes: bool = False):
    """The many-instance workloads: `count` seeded synthetic designs with cell
    counts from job_cells, the config-1 degree distribution, `nets_per_cell`
    nets per cell, 57% utilisation at adaptec1's cell density; each with its
    bin dimension by the automatic rule clamped to [dim_lo, dim_hi].  Returns
    [(design, bins per side)]."""
    out = []
    for k, cells in enumerate(job_cells(count, lo, hi, seed0, extremes)):
        d = synth_design(cells, max(8, cells // 400), int(cells * nets_per_cell),
                         25.25 * np.sqrt(cells), 0.57, seed=seed0 + k + 1, name=f"job{k}")
        out.append((d, auto_bin_dim(d, dim_lo, dim_hi)))
    return out


# Nets per cell of the serving stream: with the config-1 degree distribution
# (3.28 pins per net on average) a 1M-cell job has about 2.1M pins.
STREAM_NETS_PER_CELL = 0.64


def serving_stream(count: int, rate: float, seed0: int = 100, lo: int = 10000,
                   hi: int = 1000000, extremes: bool = False):
    """The serving workload: `count` jobs (job_designs over [lo, hi] cells,
    STREAM_NETS_PER_CELL nets per cell, bins 256-1024 by the automatic rule)
    arriving as a Poisson process of `rate` jobs per second (exponential gaps
    from the splitmix64 stream of seed 1002), the first at time 0.  Returns
    ([(design, bins per side)], arrival times)."""
    jobs = job_designs(count, lo, hi, seed0, 256, 1024, STREAM_NETS_PER_CELL, extremes)
    gaps = -np.log(1.0 - Rng(1002).uniform(count)) / rate
    return jobs, np.cumsum(gaps) - gaps[0]
