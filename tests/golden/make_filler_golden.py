"""Writes the filler golden fixtures (tests/golden/bookshelf/<case>_fillers.npz).

Inputs are the BookShelf designs under tests/golden/bookshelf/ (the
reference's own test data and seeded synthetic designs; `mac` is written
here from design.fixture_eplace(400, 4), a design with movable macros).
Expected outputs come from the reference's initializeFillers and
scatterFillers (src/gp/filler.cpp:33-139) run on the database its own
BookShelf reader built, from oracle/_ref (make -C oracle ref):
  t<target>_s<seed>_info  count, width, height, whitespace, nodeAreaScaled,
                          totalFillerArea
  t<target>_s<seed>_x/_y  filler centres
  rescatter_x/_y          scatterFillers(seed 1002, offset 3) after t1.0_s1002
Position arrays longer than KEEP entries are stored as their first and last
KEEP/2 entries plus the SHA-256 of the whole float32 array (key + "_sha").

Run from the repo root: python tests/golden/make_filler_golden.py
"""
import ctypes as C
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bookshelf")
CASES = {"tiny": "tiny/tiny.aux", "fixture": "fixture/fixture.aux", "thin1": "thin1/thin1.aux",
         "synth": "synth/synth.aux", "mac": "mac/mac.aux"}
TARGETS = (1.0, 0.8, 0.6)
SEEDS = (1002, 7)
KEEP = 4096


def store(arrs, key, v):
    v = np.ascontiguousarray(v, np.float32)
    if v.size > KEEP:
        arrs[key + "_sha"] = np.array(hashlib.sha256(v.tobytes()).hexdigest())
        arrs[key + "_n"] = np.int64(v.size)
        v = np.concatenate([v[:KEEP // 2], v[-KEEP // 2:]])
    arrs[key] = v


def main():
    from superplace_amd import bookshelf, design as D
    bookshelf.write_bookshelf(D.fixture_eplace(400, 4), os.path.join(OUT, "mac"), "mac")

    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libsp_ref.so"))
    L.ref_bookshelf_read.restype = C.c_void_p
    L.ref_bookshelf_read.argtypes = [C.c_char_p]
    L.ref_fillers.restype = C.c_int
    L.ref_fillers.argtypes = [C.c_void_p, C.c_float, C.c_uint64, C.c_void_p, C.c_void_p,
                              C.c_void_p, C.c_int]
    L.ref_fillers_rescatter.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p,
                                        C.c_void_p]
    L.ref_bookshelf_free.argtypes = [C.c_void_p]
    cap = 1 << 20
    for case, aux in CASES.items():
        h = L.ref_bookshelf_read(os.path.join(OUT, aux).encode())
        assert h, case
        arrs = {}
        for t in TARGETS:
            for s in SEEDS:
                info = np.zeros(6)
                x = np.zeros(cap, np.float32)
                y = np.zeros(cap, np.float32)
                n = L.ref_fillers(h, t, s, info.ctypes.data, x.ctypes.data, y.ctypes.data, cap)
                arrs[f"t{t}_s{s}_info"] = info
                store(arrs, f"t{t}_s{s}_x", x[:n])
                store(arrs, f"t{t}_s{s}_y", y[:n])
        n = L.ref_fillers(h, 1.0, 1002, np.zeros(6).ctypes.data, None, None, 0)
        x = np.zeros(max(n, 1), np.float32)
        y = np.zeros(max(n, 1), np.float32)
        L.ref_fillers_rescatter(h, 1002, 3, x.ctypes.data, y.ctypes.data)
        store(arrs, "rescatter_x", x[:n])
        store(arrs, "rescatter_y", y[:n])
        L.ref_bookshelf_free(h)
        np.savez_compressed(os.path.join(OUT, f"{case}_fillers.npz"), **arrs)
        print(case, {k: int(v[0]) for k, v in arrs.items() if k.endswith("info")})


if __name__ == "__main__":
    main()
