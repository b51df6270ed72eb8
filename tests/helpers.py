"""Small design builders shared by the tests (shapes of the reference's own
test fixtures, test/test_density.cpp:24-54 and test/test_wirelength.cpp:19-57)."""
import numpy as np

from superplace_amd import design as D


def density_db(size, cells):
    """cells: (cx, cy, w, h, flags); one row of height `size` covering the core."""
    order = sorted(range(len(cells)), key=lambda i: (
        2 if cells[i][4] & D.F_FILLER else 1 if cells[i][4] & D.F_FIXED else 0, i))
    cs = [cells[i] for i in order]
    mov = [c for c in cs if not (c[4] & (D.F_FILLER | D.F_FIXED))]
    fix = [c for c in cs if (c[4] & D.F_FIXED) and not (c[4] & D.F_FILLER)]
    fil = [c for c in cs if c[4] & D.F_FILLER]
    base = mov + fix
    rows = (np.array([0.0]), np.array([size]), np.array([0.0]), np.array([size]))
    col = lambda L, k: np.array([c[k] for c in L], dtype=np.float64)
    d = D.make_design(col(base, 0), col(base, 1), col(base, 2), col(base, 3),
                      np.array([c[4] for c in base], np.uint8), len(mov),
                      np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0), np.zeros(0), 0,
                      rows)
    if fil:
        k = len(fil)
        d.node_x = np.concatenate([d.node_x, col(fil, 0).astype(np.float32)])
        d.node_y = np.concatenate([d.node_y, col(fil, 1).astype(np.float32)])
        d.node_w = np.concatenate([d.node_w, col(fil, 2).astype(np.float32)])
        d.node_h = np.concatenate([d.node_h, col(fil, 3).astype(np.float32)])
        d.node_flags = np.concatenate([d.node_flags, np.full(k, D.F_FILLER, np.uint8)])
        d.num_fillers = k
        d.node2pin_start = np.concatenate([d.node2pin_start, np.zeros(k, np.int32)])
    return d


def small_netlist(n_mov, n_fix, nets, seed):
    import importlib.util, os
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(os.path.dirname(__file__), "golden", "make_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m.small_netlist(n_mov, n_fix, nets, seed)
