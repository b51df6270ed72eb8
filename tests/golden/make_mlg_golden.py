"""Writes tests/golden/ref_mlg.npz and tests/golden/mlg/<case>.json: for
every case of mlg_cases.py, what the reference's own legalizeMacros
(src/lg/macro_sa.cpp, built into oracle/_ref by `make -C oracle ref`) gives:
node centres, node flags, pin x offsets, orientations, the MacroSaResult
fields (saMs, a wall-clock time, is stored as 0) and the summary keys it
wrote to its MetricsSink (durations_ms dropped: wall-clock).

Run from the repo root: python tests/golden/make_mlg_golden.py
"""
import ctypes as C
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)
from mlg_cases import cases, row_sites  # noqa: E402

RES_KEYS = ["num_macros", "moves", "accepted", "repaired", "chained", "hot_macros",
            "overlap_before", "overlap_after_sa", "overlap_after", "sa_ms", "t0",
            "typical_worsen", "hpwl_before", "hpwl_after"]


def ref_lib():
    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libsp_ref.so"))
    V = C.c_void_p
    L.ref_macro_sa.restype = C.c_int
    L.ref_macro_sa.argtypes = [C.c_int] * 4 + [V] * 11 + [C.c_int] + [V] * 5 + \
        [C.c_uint64, C.c_int, V, V, V]
    L.ref_metrics_create.restype = V
    L.ref_metrics_create.argtypes = [C.c_char_p]
    L.ref_metrics_write.argtypes = [V, C.c_char_p]
    L.ref_metrics_free.argtypes = [V]
    return L


def run_ref(L, d, cfg):
    """(outputs dict, summary dict) of the reference on one case."""
    n = d.num_nodes
    p = lambda a: a.ctypes.data
    x = np.ascontiguousarray(d.node_x[:n], np.float32).copy()
    y = np.ascontiguousarray(d.node_y[:n], np.float32).copy()
    flags = np.ascontiguousarray(d.node_flags[:n], np.uint8).copy()
    offx = np.ascontiguousarray(d.pin_offset_x, np.float32).copy()
    orient = np.ascontiguousarray(d.meta["node_orient"][:n], np.uint8).copy()
    step = np.asarray(d.meta["row_step"], np.float32)
    wts = d.meta.get("net_weight")
    wts = np.ones(d.num_nets, np.float32) if wts is None else np.asarray(wts, np.float32)
    arrs = [np.ascontiguousarray(a) for a in (
        d.node_w[:n].astype(np.float32), d.node_h[:n].astype(np.float32),
        d.pin2node.astype(np.int32), d.pin2net.astype(np.int32),
        d.pin_offset_y.astype(np.float32), wts, d.row_ly.astype(np.float32),
        d.row_height.astype(np.float32), d.row_lx.astype(np.float32), step,
        row_sites(d, step))]
    cfgf = np.array([cfg["init_accept"], cfg["temp_ratio"], cfg["disp_weight"]], np.float32)
    res = np.zeros(14)
    sink = L.ref_metrics_create(b"")
    rc = L.ref_macro_sa(d.num_movable, n, d.num_pins, d.num_nets, p(x), p(y), p(arrs[0]),
                        p(arrs[1]), p(flags), p(arrs[2]), p(arrs[3]), p(offx), p(arrs[4]),
                        p(arrs[5]), p(orient), d.num_rows, p(arrs[6]), p(arrs[7]), p(arrs[8]),
                        p(arrs[9]), p(arrs[10]), cfg["seed"], cfg["moves"], p(cfgf), sink,
                        p(res))
    assert rc == 0
    with tempfile.TemporaryDirectory() as t:
        path = os.path.join(t, "summary.json")
        L.ref_metrics_write(sink, path.encode())
        with open(path) as f:
            summary = json.load(f)
    L.ref_metrics_free(sink)
    summary.pop("durations_ms", None)
    res[RES_KEYS.index("sa_ms")] = 0.0
    return dict(x=x, y=y, flags=flags, offx=offx, orient=orient, res=res), summary


def main():
    L = ref_lib()
    out = {}
    os.makedirs(os.path.join(HERE, "mlg"), exist_ok=True)
    for name, (d, cfg) in cases().items():
        r, summary = run_ref(L, d, cfg)
        for k, v in r.items():
            out[f"{name}_{k}"] = v
        with open(os.path.join(HERE, "mlg", f"{name}.json"), "w") as f:
            json.dump(summary, f, indent=1, sort_keys=True)
        print(name, dict(zip(RES_KEYS, r["res"])), flush=True)
    np.savez_compressed(os.path.join(HERE, "ref_mlg.npz"), **out)


if __name__ == "__main__":
    main()
