"""BookShelf placement files (.aux/.nodes/.nets/.pl/.scl/.wts) through the
library's reader and .pl writer (the C ABI's sp_bookshelf_*), mirroring
BookshelfReader::read / readPlacement and BookshelfWriter::writePl
(src/io/bookshelf_reader.h:20-41, src/io/bookshelf_writer.h:15).

read() returns a Design (node order [movable | fixed], centres, no fillers)
whose meta holds the node names and the open file handle that write_pl()
uses.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import hip
from .design import Design


def _array(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype)
    ctype = np.ctypeslib.as_ctypes_type(np.dtype(dtype))
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(n,)).copy()


class BookshelfFile:
    """An open BookShelf design (the reader's database, kept for writing)."""

    def __init__(self, aux_path: str):
        h = hip._P()
        hip.check(hip.lib().sp_bookshelf_read(str(aux_path).encode(), C.byref(h)))
        self._h = h
        self.aux_path = str(aux_path)

    def design(self, name: str | None = None) -> Design:
        v = hip.SpHipDesign()
        hip.check(hip.lib().sp_bookshelf_design(self._h, C.byref(v)))
        nodes, pins, nets, rows = v.num_nodes, v.num_pins, v.num_nets, v.num_rows
        f32, i32 = np.float32, np.int32
        d = Design(
            num_movable=v.num_movable, num_nodes=nodes, num_fillers=0,
            node_x=_array(v.node_x, nodes, f32), node_y=_array(v.node_y, nodes, f32),
            node_w=_array(v.node_w, nodes, f32), node_h=_array(v.node_h, nodes, f32),
            node_flags=_array(v.node_flags, nodes, np.uint8),
            pin_offset_x=_array(v.pin_offset_x, pins, f32),
            pin_offset_y=_array(v.pin_offset_y, pins, f32),
            pin2node=_array(v.pin2node, pins, i32), pin2net=_array(v.pin2net, pins, i32),
            net2pin_start=_array(v.net2pin_start, nets + 1, i32),
            flat_net2pin=_array(v.flat_net2pin, pins, i32),
            node2pin_start=_array(v.node2pin_start, nodes + 1, i32),
            flat_node2pin=_array(v.flat_node2pin, pins, i32),
            row_ly=_array(v.row_ly, rows, f32), row_height=_array(v.row_height, rows, f32),
            row_lx=_array(v.row_lx, rows, f32), row_hx=_array(v.row_hx, rows, f32),
            core=(float(v.core_lx), float(v.core_ly), float(v.core_hx), float(v.core_hy)),
            name=name or self.aux_path.rsplit("/", 1)[-1].rsplit(".", 1)[0])
        orient, wts, step, sites = (hip._P() for _ in range(4))
        rh = C.c_float()
        hip.check(hip.lib().sp_bookshelf_extra(self._h, C.byref(orient), C.byref(wts),
                                               C.byref(step), C.byref(sites), C.byref(rh)))
        d.meta.update(node_names=self.names(0), net_names=self.names(1),
                      node_orient=_array(orient.value, nodes, np.uint8),
                      net_weight=_array(wts.value, nets, f32),
                      row_step=_array(step.value, rows, f32),
                      row_sites=_array(sites.value, rows, i32),
                      row_height=float(rh.value), bookshelf=self)
        return d

    def names(self, which: int) -> list[str]:
        need = C.c_int64()
        hip.check(hip.lib().sp_bookshelf_names(self._h, which, None, 0, C.byref(need)))
        buf = C.create_string_buffer(max(need.value, 1))
        hip.check(hip.lib().sp_bookshelf_names(self._h, which, buf, need.value, C.byref(need)))
        text = buf.raw[:need.value].decode()
        return text.split("\n")[:-1] if text else []

    def read_pl(self, pl_path: str) -> None:
        hip.check(hip.lib().sp_bookshelf_read_pl(self._h, str(pl_path).encode()))

    def write_pl(self, path: str, x, y, flags=None, orient=None) -> None:
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        if flags is None and orient is None:
            hip.check(hip.lib().sp_bookshelf_write_pl(self._h, str(path).encode(),
                                                      x.ctypes.data, y.ctypes.data))
            return
        fl = None if flags is None else np.ascontiguousarray(flags, np.uint8)
        ori = None if orient is None else np.ascontiguousarray(orient, np.uint8)
        hip.check(hip.lib().sp_bookshelf_write_pl_state(
            self._h, str(path).encode(), x.ctypes.data, y.ctypes.data,
            None if fl is None else fl.ctypes.data, None if ori is None else ori.ctypes.data))

    def close(self) -> None:
        if getattr(self, "_h", None):
            hip.lib().sp_bookshelf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read(aux_path: str) -> Design:
    """BookshelfReader::read: the design of `aux_path`."""
    return BookshelfFile(aux_path).design()


def write_pl(path: str, design: Design, x=None, y=None) -> None:
    """BookshelfWriter::writePl for a design from read(): the first num_nodes
    centres (x, y; default the design's) as lower-left corners, with the
    design's current flags (/FIXED) and orientations."""
    bf = design.meta.get("bookshelf")
    if bf is None:
        raise ValueError("write_pl needs a design read from BookShelf files")
    n = design.num_nodes
    x = design.node_x[:n] if x is None else np.asarray(x)[:n]
    y = design.node_y[:n] if y is None else np.asarray(y)[:n]
    orient = design.meta.get("node_orient")
    bf.write_pl(path, x, y, flags=design.node_flags[:n],
                orient=None if orient is None else np.asarray(orient)[:n])


def _num(v) -> str:
    """Shortest decimal that reads back as the same float32."""
    return np.format_float_positional(np.float32(v), unique=True, trim="-")


def write_bookshelf(design: Design, directory: str, name: str, site: float = 1.0) -> str:
    """Writes a design (without fillers) as a BookShelf set: name.aux, .nodes,
    .nets, .pl, .scl, .wts; returns the .aux path.  Fixed nodes are
    terminals; the .pl holds lower-left corners; rows of whole sites of
    width `site` from each row's lx."""
    import os
    os.makedirs(directory, exist_ok=True)
    d = design
    n = d.num_nodes
    nodes = [f"o{i}" for i in range(n)]
    fixed = (d.node_flags[:n] & 2) != 0
    with open(os.path.join(directory, f"{name}.nodes"), "w") as f:
        f.write(f"UCLA nodes 1.0\n\nNumNodes : {n}\nNumTerminals : {int(fixed.sum())}\n")
        for i in range(n):
            f.write(f"\t{nodes[i]}\t{_num(d.node_w[i])}\t{_num(d.node_h[i])}"
                    f"{chr(9) + 'terminal' if fixed[i] else ''}\n")
    with open(os.path.join(directory, f"{name}.nets"), "w") as f:
        f.write(f"UCLA nets 1.0\n\nNumNets : {d.num_nets}\nNumPins : {d.num_pins}\n")
        for k in range(d.num_nets):
            b, e = int(d.net2pin_start[k]), int(d.net2pin_start[k + 1])
            f.write(f"NetDegree : {e - b}   net{k}\n")
            for p in d.flat_net2pin[b:e]:
                f.write(f"\t{nodes[d.pin2node[p]]}\tI : {_num(d.pin_offset_x[p])}\t"
                        f"{_num(d.pin_offset_y[p])}\n")
    with open(os.path.join(directory, f"{name}.pl"), "w") as f:
        f.write("UCLA pl 1.0\n\n")
        for i in range(n):
            llx = np.float32(d.node_x[i]) - np.float32(0.5) * np.float32(d.node_w[i])
            lly = np.float32(d.node_y[i]) - np.float32(0.5) * np.float32(d.node_h[i])
            f.write(f"{nodes[i]}\t{_num(llx)}\t{_num(lly)}\t: N{' /FIXED' if fixed[i] else ''}\n")
    with open(os.path.join(directory, f"{name}.scl"), "w") as f:
        f.write(f"UCLA scl 1.0\n\nNumRows : {d.num_rows}\n\n")
        step = _num(site)
        for r in range(d.num_rows):
            span = float(d.row_hx[r] - d.row_lx[r])
            sites = int(round(span)) if site == 1.0 else int(np.floor(span / float(np.float32(site)) + 1e-6))
            f.write(f"CoreRow Horizontal\n  Coordinate    :   {_num(d.row_ly[r])}\n"
                    f"  Height        :   {_num(d.row_height[r])}\n  Sitewidth     :    {step}\n"
                    f"  Sitespacing   :    {step}\n  Siteorient    :    1\n  Sitesymmetry  :    1\n"
                    f"  SubrowOrigin  :    {_num(d.row_lx[r])}\tNumSites  :  {sites}\nEnd\n")
    with open(os.path.join(directory, f"{name}.wts"), "w") as f:
        f.write("UCLA wts 1.0\n\n")
    aux = os.path.join(directory, f"{name}.aux")
    with open(aux, "w") as f:
        f.write(f"RowBasedPlacement :  {name}.nodes  {name}.nets  {name}.wts  {name}.pl  {name}.scl\n")
    return aux
