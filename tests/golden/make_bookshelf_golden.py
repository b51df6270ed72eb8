"""Writes the BookShelf golden fixtures under tests/golden/bookshelf/.

Inputs: the reference's own BookShelf test data (copied as data files into
tiny/, fixture/, thin1/) and a seeded synthetic design written as BookShelf
files (synth/).  Expected outputs come from the reference's own reader and
writer (src/io/bookshelf_reader.cpp, bookshelf_writer.cpp), compiled from
/root/reference into oracle/_ref by `make -C oracle ref`:
  <case>_ref.npz  every database array after BookshelfReader::read
  <case>_ref.pl   BookshelfWriter::writePl of deterministically moved centres

Run from the repo root: python tests/golden/make_bookshelf_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bookshelf")
CASES = {"tiny": "tiny/tiny.aux", "fixture": "fixture/fixture.aux", "thin1": "thin1/thin1.aux",
         "synth": "synth/synth.aux"}


def moved(n):
    """Deterministic centres for the writer check (values with fractions)."""
    i = np.arange(n, dtype=np.float64)
    return ((i * 37.25) % 91.5 + 0.125).astype(np.float32), ((i * 11.75) % 53.0 + 0.375).astype(np.float32)


def main():
    from superplace_amd import bookshelf, design as D
    d = D.synth_design(1000, 6, 1100, 300.0, 0.7, seed=11, name="synth", with_fillers=False)
    bookshelf.write_bookshelf(d, os.path.join(OUT, "synth"), "synth")

    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libsp_ref.so"))
    L.ref_bookshelf_read.restype = C.c_void_p
    L.ref_bookshelf_read.argtypes = [C.c_char_p]
    L.ref_bookshelf_counts.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.ref_bookshelf_get.argtypes = [C.c_void_p] + [C.c_void_p] * 16
    L.ref_bookshelf_node_name.restype = C.c_char_p
    L.ref_bookshelf_node_name.argtypes = [C.c_void_p, C.c_int]
    L.ref_bookshelf_write_pl.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p]
    L.ref_bookshelf_free.argtypes = [C.c_void_p]
    p = lambda a: a.ctypes.data
    for case, aux in CASES.items():
        h = L.ref_bookshelf_read(os.path.join(OUT, aux).encode())
        assert h, case
        cnt = np.zeros(5, np.int32)
        core = np.zeros(5, np.float32)
        L.ref_bookshelf_counts(h, p(cnt), p(core))
        mov, nodes, pins, nets, rows = (int(v) for v in cnt)
        a = dict(x=np.zeros(nodes, np.float32), y=np.zeros(nodes, np.float32),
                 w=np.zeros(nodes, np.float32), h=np.zeros(nodes, np.float32),
                 flags=np.zeros(nodes, np.uint8), orient=np.zeros(nodes, np.uint8),
                 offx=np.zeros(pins, np.float32), offy=np.zeros(pins, np.float32),
                 pin2node=np.zeros(pins, np.int32), pin2net=np.zeros(pins, np.int32),
                 net2pin_start=np.zeros(nets + 1, np.int32), flat_net2pin=np.zeros(pins, np.int32),
                 node2pin_start=np.zeros(nodes + 1, np.int32),
                 flat_node2pin=np.zeros(pins, np.int32), rows=np.zeros(6 * rows, np.float32),
                 net_weight=np.zeros(nets, np.float32))
        L.ref_bookshelf_get(h, *[p(v) for v in a.values()])
        names = "\n".join(L.ref_bookshelf_node_name(h, i).decode() for i in range(nodes))
        np.savez_compressed(os.path.join(OUT, f"{case}_ref.npz"), counts=cnt, core=core,
                            names=np.array(names), **a)
        mx, my = moved(nodes)
        assert L.ref_bookshelf_write_pl(h, os.path.join(OUT, f"{case}_ref.pl").encode(),
                                        p(mx), p(my)) == 0
        L.ref_bookshelf_free(h)
        print(case, "nodes", nodes, "movable", mov, "pins", pins, "nets", nets, "rows", rows)


if __name__ == "__main__":
    main()
