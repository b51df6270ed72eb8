"""Writes the Bound2Bound golden fixtures (tests/golden/ref_b2b_<case>.npz).

Expected outputs come from the reference's own computeNetBoundPins
(src/db/hpwl.cpp:70-102) and forEachB2BPair (src/init/b2b_model.h:56-82),
compiled into oracle/_ref by `make -C oracle ref`, on seeded designs at two
placements each (all movable nodes at the core centre, as QuadraticPlacer
starts, and a seeded scatter):
  bounds_<pl>             minX, maxX, minY, maxY pins per net (4 x nets)
  pairs_<pl>_<axis>       (p, q) per B2B pair in enumeration order
  w_<pl>_<axis>           their weights (float32)
with min distance 1.0 and nets of degree <= 100 (the defaults), and the
inputs (x_<pl>, y_<pl>, net weights).

Run from the repo root: python tests/golden/make_b2b_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.dirname(os.path.abspath(__file__))


def cases():
    from superplace_amd import design as D
    return {"fixture": D.fixture_eplace(300, 2, seed=3),
            "synth": D.synth_design(1500, 5, 1600, 420.0, 0.7, seed=21, name="synth",
                                    with_fillers=False)}


def placements(d, seed):
    n = d.num_nodes
    cx = np.float32(0.5) * (np.float32(d.core[0]) + np.float32(d.core[2]))
    cy = np.float32(0.5) * (np.float32(d.core[1]) + np.float32(d.core[3]))
    x0, y0 = d.node_x[:n].astype(np.float32).copy(), d.node_y[:n].astype(np.float32).copy()
    x0[:d.num_movable], y0[:d.num_movable] = cx, cy
    rng = np.random.default_rng(seed)
    x1, y1 = d.node_x[:n].astype(np.float32).copy(), d.node_y[:n].astype(np.float32).copy()
    x1[:d.num_movable] = rng.uniform(d.core[0], d.core[2], d.num_movable).astype(np.float32)
    y1[:d.num_movable] = rng.uniform(d.core[1], d.core[3], d.num_movable).astype(np.float32)
    return {"centre": (x0, y0), "scatter": (x1, y1)}


def main():
    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libsp_ref.so"))
    L.ref_b2b_pairs.restype = C.c_int64
    L.ref_b2b_pairs.argtypes = [C.c_int] * 4 + [C.c_void_p] * 10 + [C.c_int, C.c_float, C.c_int] + \
        [C.c_void_p] * 4 + [C.c_int64]
    p = lambda a: a.ctypes.data
    for name, d in cases().items():
        n = d.num_nodes
        wts = (1.0 + (np.arange(d.num_nets) % 3) * 0.5).astype(np.float32)
        arrs = dict(net_weight=wts)
        for pl, (x, y) in placements(d, 5).items():
            arrs[f"x_{pl}"], arrs[f"y_{pl}"] = x, y
            for axis in (0, 1):
                cap = 1 << 22
                bounds = np.zeros(4 * d.num_nets, np.int32)
                pp, qq = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
                ww = np.zeros(cap, np.float32)
                a = [np.ascontiguousarray(v) for v in (
                    x, y, d.node_w[:n].astype(np.float32), d.node_h[:n].astype(np.float32),
                    d.node_flags[:n].astype(np.uint8), d.pin2node.astype(np.int32),
                    d.pin2net.astype(np.int32), d.pin_offset_x.astype(np.float32),
                    d.pin_offset_y.astype(np.float32), wts)]
                cnt = L.ref_b2b_pairs(d.num_movable, n, d.num_pins, d.num_nets, *[p(v) for v in a],
                                      axis, 1.0, 100, p(bounds), p(pp), p(qq), p(ww), cap)
                assert cnt <= cap
                arrs[f"bounds_{pl}"] = bounds.reshape(4, d.num_nets)
                arrs[f"pairs_{pl}_{axis}"] = np.stack([pp[:cnt], qq[:cnt]])
                arrs[f"w_{pl}_{axis}"] = ww[:cnt]
        np.savez_compressed(os.path.join(OUT, f"ref_b2b_{name}.npz"), **arrs)
        print(name, {k: v.shape for k, v in arrs.items() if k.startswith("pairs")})


if __name__ == "__main__":
    main()
