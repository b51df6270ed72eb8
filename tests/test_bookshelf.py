"""BookShelf reader and .pl writer (sp_bookshelf_*, superplace_amd.bookshelf)
against the reference's own reader and writer.

Golden fixtures (tests/golden/bookshelf/, made by make_bookshelf_golden.py)
hold every database array BookshelfReader::read produced
(src/io/bookshelf_reader.cpp) on the reference's own test designs
(test/check/tiny.*, test/gp_stagnation/fixture.*, test_data/thin1/thin1.*)
and a seeded 1000-cell design, and the file BookshelfWriter::writePl
(src/io/bookshelf_writer.cpp) wrote for fixed moved centres.  The bar is
bit-exact arrays and names and byte-identical .pl text.  Host code only."""
import os

import numpy as np
import pytest

from superplace_amd import bookshelf, hip

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bookshelf")
CASES = {"tiny": "tiny/tiny.aux", "fixture": "fixture/fixture.aux", "thin1": "thin1/thin1.aux",
         "synth": "synth/synth.aux"}
F_FIXED = 2   # node flag bit, src/db/place_db.h:17-22


def moved(n):
    i = np.arange(n, dtype=np.float64)
    return ((i * 37.25) % 91.5 + 0.125).astype(np.float32), ((i * 11.75) % 53.0 + 0.375).astype(np.float32)


def golden(case):
    with np.load(os.path.join(GOLD, f"{case}_ref.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("case", sorted(CASES))
def test_reader_matches_reference(case):
    g = golden(case)
    d = bookshelf.read(os.path.join(GOLD, CASES[case]))
    mov, nodes, pins, nets, rows = (int(v) for v in g["counts"])
    assert (d.num_movable, d.num_nodes, d.num_pins, d.num_nets, d.num_rows) == \
        (mov, nodes, pins, nets, rows)
    assert d.num_fillers == 0
    core = g["core"]
    assert np.array_equal(np.float32(d.core), core[:4])
    assert np.float32(d.meta["row_height"]) == core[4]
    pairs = [("x", d.node_x), ("y", d.node_y), ("w", d.node_w), ("h", d.node_h),
             ("flags", d.node_flags), ("orient", d.meta["node_orient"]),
             ("offx", d.pin_offset_x), ("offy", d.pin_offset_y), ("pin2node", d.pin2node),
             ("pin2net", d.pin2net), ("net2pin_start", d.net2pin_start),
             ("flat_net2pin", d.flat_net2pin), ("node2pin_start", d.node2pin_start),
             ("flat_node2pin", d.flat_node2pin), ("net_weight", d.meta["net_weight"])]
    for key, ours in pairs:
        ref = g[key]
        assert ours.shape == ref.shape, key
        assert np.array_equal(ours.view(np.uint8), ref.view(np.uint8)), key   # bitwise
    r = g["rows"].reshape(rows, 6)
    for col, ours in enumerate([d.row_ly, d.row_height, d.row_lx, d.row_hx,
                                d.meta["row_step"], d.meta["row_sites"].astype(np.float32)]):
        assert np.array_equal(ours, r[:, col]), col
    assert d.meta["node_names"] == str(g["names"]).split("\n")


@pytest.mark.parametrize("case", sorted(CASES))
def test_pl_writer_byte_identical(case, tmp_path):
    d = bookshelf.read(os.path.join(GOLD, CASES[case]))
    x, y = moved(d.num_nodes)
    out = tmp_path / "out.pl"
    bookshelf.write_pl(str(out), d, x, y)
    ref = open(os.path.join(GOLD, f"{case}_ref.pl"), "rb").read()
    assert out.read_bytes() == ref


def test_reference_fixture_facts():
    """What the reference's own tiny test states (test/check): 3 movable of 6
    nodes, 2 nets over 4 pins, a 20x20 core, fixed nodes after movable ones."""
    d = bookshelf.read(os.path.join(GOLD, CASES["tiny"]))
    assert (d.num_movable, d.num_nodes, d.num_nets, d.num_pins) == (3, 6, 2, 4)
    assert d.core == (0.0, 0.0, 20.0, 20.0)
    assert np.all(d.node_flags[:3] & F_FIXED == 0)
    assert np.all(d.node_flags[3:] & F_FIXED != 0)


def test_write_read_round_trip(tmp_path):
    """.pl written from moved centres and read back restores them to float
    precision of the 6-decimal text, and leaves everything else alone."""
    aux = os.path.join(GOLD, CASES["synth"])
    f = bookshelf.BookshelfFile(aux)
    d0 = f.design()
    x, y = moved(d0.num_nodes)
    pl = tmp_path / "moved.pl"
    f.write_pl(str(pl), x, y)
    f.read_pl(str(pl))
    d1 = f.design()
    assert np.allclose(d1.node_x, x, atol=2e-6 * np.abs(x).max())
    assert np.allclose(d1.node_y, y, atol=2e-6 * np.abs(y).max())
    assert np.array_equal(d1.node_w, d0.node_w) and np.array_equal(d1.flat_net2pin, d0.flat_net2pin)
    f.close()


def test_pl_skips_unknown_nodes(tmp_path):
    """A .pl line naming no node is skipped (bookshelf_reader.cpp readPlacement)."""
    f = bookshelf.BookshelfFile(os.path.join(GOLD, CASES["tiny"]))
    d0 = f.design()
    pl = tmp_path / "x.pl"
    name = d0.meta["node_names"][0]
    pl.write_text(f"UCLA pl 1.0\n\nnot_a_node 5 5 : N\n{name} 1 2 : N\n")
    f.read_pl(str(pl))
    d1 = f.design()
    assert d1.node_x[0] == np.float32(1) + np.float32(0.5) * d0.node_w[0]
    assert d1.node_y[0] == np.float32(2) + np.float32(0.5) * d0.node_h[0]
    assert np.array_equal(d1.node_x[1:], d0.node_x[1:])


def test_errors(tmp_path):
    with pytest.raises(hip.InvalidArgument):
        bookshelf.read(str(tmp_path / "missing.aux"))
    src = os.path.join(GOLD, "tiny")
    for ext in ("nodes", "pl", "scl", "wts"):
        (tmp_path / f"tiny.{ext}").write_bytes(open(os.path.join(src, f"tiny.{ext}"), "rb").read())
    (tmp_path / "tiny.aux").write_bytes(open(os.path.join(src, "tiny.aux"), "rb").read())
    (tmp_path / "tiny.nets").write_text("UCLA nets 1.0\nNumNets : 1\nNumPins : 1\n"
                                        "NetDegree : 1 n0\n\tghost I : 0 0\n")
    with pytest.raises(hip.InvalidArgument, match="unknown node"):
        bookshelf.read(str(tmp_path / "tiny.aux"))
    d = bookshelf.read(os.path.join(GOLD, CASES["tiny"]))
    d.meta.pop("bookshelf")
    with pytest.raises(ValueError):
        bookshelf.write_pl(str(tmp_path / "o.pl"), d)
