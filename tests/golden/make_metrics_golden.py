"""Writes tests/golden/metrics/<scenario>/{metrics.csv,summary.json}: the
files the reference's own MetricsSink (src/util/metrics.cpp, built into
oracle/_ref by `make -C oracle ref`) writes for the call sequences of
metrics_scenarios.py.

Run from the repo root: python tests/golden/make_metrics_golden.py
"""
import ctypes as C
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
from metrics_scenarios import full_row, scenarios  # noqa: E402


def main():
    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libsp_ref.so"))
    L.ref_metrics_create.restype = C.c_void_p
    L.ref_metrics_create.argtypes = [C.c_char_p]
    for f in ("begin",):
        getattr(L, "ref_metrics_" + f).argtypes = [C.c_void_p, C.c_char_p]
    L.ref_metrics_push.argtypes = [C.c_void_p, C.c_int, C.c_double] + [C.c_float] * 7 + \
        [C.c_double, C.c_float, C.c_float, C.c_double, C.c_float, C.c_long]
    L.ref_metrics_end.argtypes = [C.c_void_p, C.c_double]
    L.ref_metrics_duration.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
    L.ref_metrics_number.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
    L.ref_metrics_string.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p]
    L.ref_metrics_write.argtypes = [C.c_void_p, C.c_char_p]
    L.ref_metrics_free.argtypes = [C.c_void_p]
    f32 = lambda v: float(np.float32(v))
    for name, ops in scenarios().items():
        tmp = tempfile.mkdtemp()
        h = L.ref_metrics_create(tmp.encode())
        for op in ops:
            kind = op[0]
            if kind == "begin":
                rc = L.ref_metrics_begin(h, op[1].encode())
            elif kind == "push":
                r = full_row(op[1])
                rc = L.ref_metrics_push(h, r["iter"], r["hpwl"], f32(r["overflow"]), f32(r["lambda"]),
                                        f32(r["gamma"]), f32(r["step_size"]), f32(r["grad_norm_wl"]),
                                        f32(r["grad_norm_den"]), f32(r["max_displacement"]),
                                        r["elapsed_ms"], f32(r["eval_lambda"]), f32(r["eval_gamma"]),
                                        r["delta_hpwl"], f32(r["lambda_mu"]), r["evaluations"])
            elif kind == "end":
                rc = L.ref_metrics_end(h, op[1])
            elif kind == "duration":
                rc = L.ref_metrics_duration(h, op[1].encode(), op[2])
            elif kind == "num":
                rc = L.ref_metrics_number(h, op[1].encode(), op[2])
            elif kind == "str":
                rc = L.ref_metrics_string(h, op[1].encode(), op[2].encode())
            elif kind == "write":
                rc = L.ref_metrics_write(h, os.path.join(tmp, "summary.json").encode())
            assert rc == 0, (name, op)
        L.ref_metrics_free(h)
        out = os.path.join(HERE, "metrics", name)
        os.makedirs(out, exist_ok=True)
        for fn in ("metrics.csv", "summary.json"):
            shutil.copy(os.path.join(tmp, fn), os.path.join(out, fn))
        shutil.rmtree(tmp)
        print(name, "ok")


if __name__ == "__main__":
    main()
