"""ctypes bindings of the CPU oracle (oracle/_build/libsp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libsp_oracle.so"
REF_LIB_PATH = HERE / "_ref" / "libsp_ref.so"

_P = C.c_void_p


class OraDesign(C.Structure):
    _fields_ = [("num_movable", C.c_int32), ("num_nodes", C.c_int32),
                ("num_fillers", C.c_int32), ("num_pins", C.c_int32),
                ("num_nets", C.c_int32), ("num_rows", C.c_int32),
                ("node_w", _P), ("node_h", _P), ("node_flags", _P),
                ("pin_offset_x", _P), ("pin_offset_y", _P), ("pin2node", _P), ("pin2net", _P),
                ("net2pin_start", _P), ("flat_net2pin", _P), ("node2pin_start", _P),
                ("flat_node2pin", _P), ("row_ly", _P), ("row_height", _P), ("row_lx", _P),
                ("row_hx", _P), ("core_lx", C.c_float), ("core_ly", C.c_float),
                ("core_hx", C.c_float), ("core_hy", C.c_float)]


class OraGpConfig(C.Structure):
    _fields_ = [("target_density", C.c_float), ("target_overflow", C.c_float),
                ("cgp_target_overflow", C.c_float), ("gp_max_iter", C.c_int32),
                ("ignore_net_degree", C.c_int32), ("gp_bin_dim", C.c_int32),
                ("delta_hpwl_ref", C.c_float), ("lambda_relative", C.c_int32),
                ("delta_hpwl_relative", C.c_double), ("use_bb", C.c_int32),
                ("gp_backtrack", C.c_int32), ("filler_only_iters", C.c_int32),
                ("gp_stagnation_window", C.c_int32), ("gp_stagnation_min_gain", C.c_float),
                ("stage", C.c_int32), ("precision", C.c_int32), ("max_backtrack", C.c_int32)]


class OraIterMetrics(C.Structure):
    _fields_ = [("iter", C.c_int32), ("evaluations", C.c_int32)] + [
        (k, C.c_double) for k in ("hpwl", "overflow", "lambda_", "gamma", "step_size",
                                  "grad_norm_wl", "grad_norm_den", "eval_lambda", "eval_gamma",
                                  "delta_hpwl", "lambda_mu", "elapsed_ms")]


class OraGpResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32),
                ("stagnated", C.c_int32), ("rollback_iter", C.c_int32),
                ("evaluations", C.c_int64), ("hpwl", C.c_double), ("overflow", C.c_double),
                ("initial_hpwl", C.c_double), ("initial_overflow", C.c_double),
                ("initial_lambda", C.c_double), ("seconds_loop", C.c_double)]


def build(quiet: bool = True) -> None:
    """Compile the oracle (and oracle/_ref when the reference tree is present)."""
    r = subprocess.run(["make", "-C", str(HERE), "all"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    if not quiet:
        print(r.stdout)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.ora_grid_create.restype = _P
        L.ora_overflow.restype = C.c_double
        L.ora_total_charge.restype = C.c_double
        L.ora_compute_gamma.restype = C.c_float
        L.ora_compute_gamma.argtypes = [C.c_float, C.c_float]
        L.ora_wa.restype = C.c_double
        L.ora_wa.argtypes = [_P, _P, _P, C.c_float, C.c_int, _P, C.c_int]
        L.ora_hpwl.restype = C.c_double
        L.ora_lambda_multiplier.restype = C.c_float
        L.ora_lambda_multiplier.argtypes = [C.c_double, C.c_double]
        L.ora_nesterov_quadratic.restype = C.c_double
        L.ora_nesterov_quadratic.argtypes = [C.c_int, _P, _P, _P, C.c_int, C.c_int, C.c_int,
                                             C.c_float, C.c_float, _P]
        L.ora_poisson.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _P, _P, _P, _P,
                                  C.c_int]
        L.ora_grid_create.argtypes = [_P, _P, _P, C.c_float, C.c_int, C.c_int]
        L.ora_filler_plan.argtypes = [_P, _P, _P, C.c_float, C.c_float, C.c_double, _P]
        L.ora_accumulate_fixed.argtypes = [_P, _P, _P, _P, C.c_int, _P, _P]
        _lib = L
    return _lib


CHECKER_PATH = HERE / "_ref" / "superplace-check"
ADAPTER_CHECK_PATH = HERE / "_ref" / "adapter_check"


def check_placement(aux_path, pl_path):
    """The reference's standalone legality checker (tools/check/main.cpp,
    built into oracle/_ref): (passed, report text), or None if not built."""
    if not CHECKER_PATH.exists():
        return None
    p = subprocess.run([str(CHECKER_PATH), str(aux_path), str(pl_path)], capture_output=True,
                       text=True, timeout=300)
    if p.returncode not in (0, 1):
        raise RuntimeError(f"superplace-check failed to run: {p.stderr.strip()}")
    return p.returncode == 0, p.stdout


def adapter_check(aux_path, bin_dim, precision=32):
    """The reference-side adapters (makeHipPoissonBackend, runHipStage) built
    against the reference's headers and sources (oracle/_ref/adapter_check):
    mGP of the BookShelf design with the reference's fillers, and a 96 x 64
    Poisson solve; a dict of the results and hashes, or None if not built."""
    import json
    if not ADAPTER_CHECK_PATH.exists():
        return None
    p = subprocess.run([str(ADAPTER_CHECK_PATH), str(aux_path), str(int(bin_dim)),
                        str(int(precision))], capture_output=True, text=True, timeout=600)
    if p.returncode != 0:
        raise RuntimeError(f"adapter_check failed ({p.returncode}): {p.stderr.strip()}")
    return json.loads(p.stdout.strip().splitlines()[-1])


def ref_lib():
    """The reference's own WA/HPWL/Nesterov (oracle/_ref), or None if not built."""
    global _ref
    if _ref is None:
        if not REF_LIB_PATH.exists():
            return None
        L = C.CDLL(str(REF_LIB_PATH))
        L.ref_wa.restype = C.c_double
        L.ref_wa.argtypes = [C.c_int] * 4 + [_P] * 9 + [C.c_float, C.c_int, _P]
        L.ref_hpwl.restype = C.c_double
        L.ref_hpwl.argtypes = [C.c_int] * 4 + [_P] * 9
        L.ref_nesterov_quadratic.restype = C.c_double
        L.ref_nesterov_quadratic.argtypes = [C.c_int, _P, _P, _P, C.c_int, C.c_int, C.c_int,
                                             C.c_float, C.c_float, _P]
        _ref = L
    return _ref


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


class _Keep:
    """Holds the arrays an OraDesign points into."""


def design_struct(d):
    keep = _Keep()
    arrs = dict(node_w=np.ascontiguousarray(d.node_w, np.float32),
                node_h=np.ascontiguousarray(d.node_h, np.float32),
                node_flags=np.ascontiguousarray(d.node_flags, np.uint8),
                pin_offset_x=np.ascontiguousarray(d.pin_offset_x, np.float32),
                pin_offset_y=np.ascontiguousarray(d.pin_offset_y, np.float32),
                pin2node=np.ascontiguousarray(d.pin2node, np.int32),
                pin2net=np.ascontiguousarray(d.pin2net, np.int32),
                net2pin_start=np.ascontiguousarray(d.net2pin_start, np.int32),
                flat_net2pin=np.ascontiguousarray(d.flat_net2pin, np.int32),
                node2pin_start=np.ascontiguousarray(d.node2pin_start, np.int32),
                flat_node2pin=np.ascontiguousarray(d.flat_node2pin, np.int32),
                row_ly=np.ascontiguousarray(d.row_ly, np.float32),
                row_height=np.ascontiguousarray(d.row_height, np.float32),
                row_lx=np.ascontiguousarray(d.row_lx, np.float32),
                row_hx=np.ascontiguousarray(d.row_hx, np.float32))
    keep.arrs = arrs
    s = OraDesign(num_movable=d.num_movable, num_nodes=d.num_nodes, num_fillers=d.num_fillers,
                  num_pins=d.num_pins, num_nets=d.num_nets, num_rows=d.num_rows,
                  core_lx=d.core[0], core_ly=d.core[1], core_hx=d.core[2], core_hy=d.core[3],
                  **{k: _ptr(v) for k, v in arrs.items()})
    keep.s = s
    return keep


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


class Grid:
    """BinGrid restated (bin_grid.cpp); precision 64 (truth) or 32 (reference FP32)."""

    def __init__(self, d, target_density=1.0, bin_dim=0, precision=64, x=None, y=None):
        self.d = d
        self.ds = design_struct(d)
        self.x = _f32(d.node_x if x is None else x)
        self.y = _f32(d.node_y if y is None else y)
        L = lib()
        self.h = L.ora_grid_create(C.byref(self.ds.s), _ptr(self.x), _ptr(self.y),
                                   C.c_float(target_density), int(bin_dim), int(precision))
        info = np.zeros(6)
        L.ora_grid_info(_P(self.h), _ptr(info))
        self.dim = int(info[0])
        self.step_x, self.step_y, self.origin_x, self.origin_y = (float(v) for v in info[1:5])
        self.movable_area_scaled = float(info[5])
        nb = self.dim * self.dim
        self.terminal = np.zeros(nb)
        self.base = np.zeros(nb)
        L.ora_grid_static(_P(self.h), _ptr(self.terminal), _ptr(self.base))

    def __del__(self):
        try:
            lib().ora_grid_destroy(_P(self.h))
        except Exception:
            pass

    def accumulate(self, x, y):
        nb = self.dim * self.dim
        nd, fd = np.zeros(nb), np.zeros(nb)
        x, y = _f32(x), _f32(y)
        lib().ora_accumulate(_P(self.h), C.byref(self.ds.s), _ptr(x), _ptr(y), _ptr(nd), _ptr(fd))
        return nd, fd

    def accumulate_fixed(self, x, y, frac_bits):
        nb = self.dim * self.dim
        nq, fq = np.zeros(nb, np.int64), np.zeros(nb, np.int64)
        x, y = _f32(x), _f32(y)
        lib().ora_accumulate_fixed(_P(self.h), C.byref(self.ds.s), _ptr(x), _ptr(y),
                                   int(frac_bits), _ptr(nq), _ptr(fq))
        return nq, fq

    def overflow(self, node_density):
        nd = np.ascontiguousarray(node_density, np.float64)
        return lib().ora_overflow(_P(self.h), _ptr(nd))

    def rho(self, node_density, filler_density):
        out = np.zeros(self.dim * self.dim)
        nd = np.ascontiguousarray(node_density, np.float64)
        fd = np.ascontiguousarray(filler_density, np.float64)
        lib().ora_density_rho(_P(self.h), _ptr(nd), _ptr(fd), _ptr(out))
        return out

    def gather(self, x, y, ex, ey):
        t = self.d.total_nodes
        force = np.zeros(2 * t)
        x, y = _f32(x), _f32(y)
        ex = np.ascontiguousarray(ex, np.float64)
        ey = np.ascontiguousarray(ey, np.float64)
        lib().ora_gather(_P(self.h), C.byref(self.ds.s), _ptr(x), _ptr(y), _ptr(ex), _ptr(ey),
                         _ptr(force))
        return force

    def total_charge(self, x, y):
        x, y = _f32(x), _f32(y)
        L = lib()
        L.ora_total_charge.argtypes = [_P, _P, _P, _P]
        return L.ora_total_charge(_P(self.h), C.byref(self.ds.s), _ptr(x), _ptr(y))


def poisson(rho, step_x, step_y, precision=64):
    rho = np.ascontiguousarray(rho, np.float64)
    nx, ny = rho.shape
    phi, ex, ey = (np.zeros_like(rho) for _ in range(3))
    lib().ora_poisson(nx, ny, step_x, step_y, _ptr(rho), _ptr(phi), _ptr(ex), _ptr(ey),
                      int(precision))
    return phi, ex, ey


def r2r_1d(kind, x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    L = lib()
    L.ora_r2r_1d.argtypes = [C.c_int, C.c_int, _P, _P]
    L.ora_r2r_1d(int(kind), len(x), _ptr(x), _ptr(out))
    return out


def r2r_1d_fast(kind, x):
    """The FFT path of the same transform (power-of-two length)."""
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    L = lib()
    L.ora_r2r_1d_fast.argtypes = [C.c_int, C.c_int, _P, _P]
    L.ora_r2r_1d_fast(int(kind), len(x), _ptr(x), _ptr(out))
    return out


def set_fast_poisson(on: bool):
    """Selects the FFT transforms (power-of-two grids) for poisson / eplace_run."""
    L = lib()
    L.ora_set_fast_poisson.argtypes = [C.c_int]
    L.ora_set_fast_poisson(1 if on else 0)


def wa(d, gamma, ignore_net_degree=100, precision=64, x=None, y=None):
    ds = design_struct(d)
    x = _f32(d.node_x if x is None else x)
    y = _f32(d.node_y if y is None else y)
    grad = np.zeros(2 * d.num_movable)
    wl = lib().ora_wa(C.byref(ds.s), _ptr(x), _ptr(y), float(gamma), int(ignore_net_degree),
                      _ptr(grad), int(precision))
    return wl, grad


def hpwl(d, precision=64, x=None, y=None):
    ds = design_struct(d)
    x = _f32(d.node_x if x is None else x)
    y = _f32(d.node_y if y is None else y)
    L = lib()
    L.ora_hpwl.argtypes = [_P, _P, _P, C.c_int]
    return L.ora_hpwl(C.byref(ds.s), _ptr(x), _ptr(y), int(precision))


def compute_gamma(tau, step):
    return lib().ora_compute_gamma(float(tau), float(step))


def lambda_multiplier(dh, ref):
    return lib().ora_lambda_multiplier(float(dh), float(ref))


def hpwl_reference(relative, delta_ref, delta_relative, initial_hpwl, filler_only):
    """The lambda schedule's HPWL scale, or None where the reference throws."""
    L = lib()
    L.ora_hpwl_reference.argtypes = [C.c_int, C.c_float, C.c_double, C.c_double, C.c_int, _P]
    out = C.c_double(0)
    rc = L.ora_hpwl_reference(int(relative), float(delta_ref), float(delta_relative),
                              float(initial_hpwl), int(filler_only), C.byref(out))
    return None if rc else out.value


def ref_compute_gamma(tau, step):
    L = ref_lib()
    L.ref_compute_gamma.restype = C.c_float
    L.ref_compute_gamma.argtypes = [C.c_float, C.c_float]
    return L.ref_compute_gamma(float(tau), float(step))


def ref_lambda_multiplier(dh, ref):
    L = ref_lib()
    L.ref_lambda_multiplier.restype = C.c_float
    L.ref_lambda_multiplier.argtypes = [C.c_double, C.c_double]
    return L.ref_lambda_multiplier(float(dh), float(ref))


def ref_hpwl_reference(relative, delta_ref, delta_relative, initial_hpwl, filler_only):
    L = ref_lib()
    L.ref_hpwl_reference.argtypes = [C.c_int, C.c_float, C.c_double, C.c_double, C.c_int, _P]
    out = C.c_double(0)
    rc = L.ref_hpwl_reference(int(relative), float(delta_ref), float(delta_relative),
                              float(initial_hpwl), int(filler_only), C.byref(out))
    return None if rc else out.value


def filler_plan(d, target_density=1.0):
    ds = design_struct(d)
    out = np.zeros(6)
    x, y = _f32(d.node_x), _f32(d.node_y)
    lib().ora_filler_plan(C.byref(ds.s), _ptr(x), _ptr(y), float(target_density),
                          float(d.row_height_mode), float(d.total_row_area), _ptr(out))
    return int(out[0]), float(out[1]), float(out[2]), float(out[3]), float(out[4]), float(out[5])


def nesterov_quadratic(coef, target, pos, steps, use_bb=True, backtrack=False, init_step=0.0,
                       clamp_hi=float("inf"), use_reference=False):
    coef, target = _f32(coef), _f32(target)
    pos = _f32(pos).copy()
    ev = C.c_long(0)
    L = ref_lib() if use_reference else lib()
    fn = L.ref_nesterov_quadratic if use_reference else L.ora_nesterov_quadratic
    obj = fn(len(pos), _ptr(coef), _ptr(target), _ptr(pos), int(steps), int(use_bb),
             int(backtrack), float(init_step), float(clamp_hi), C.byref(ev))
    return obj, pos, ev.value


def default_config(**kw) -> OraGpConfig:
    """configs/default.json values for the GP path."""
    c = OraGpConfig(target_density=1.0, target_overflow=0.10, cgp_target_overflow=0.10,
                    gp_max_iter=2000, ignore_net_degree=100, gp_bin_dim=0,
                    delta_hpwl_ref=3.5e5, lambda_relative=0, delta_hpwl_relative=0.01,
                    use_bb=1, gp_backtrack=1, filler_only_iters=20, gp_stagnation_window=200,
                    gp_stagnation_min_gain=0.02, stage=0, precision=64, max_backtrack=10)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def eplace_run(d, cfg: OraGpConfig, max_rows=4000, threads=0):
    """Runs one GP stage on a copy of the positions; returns (result, rows, x, y)."""
    if threads:
        lib().ora_set_threads(int(threads))
    ds = design_struct(d)
    x, y = _f32(d.node_x).copy(), _f32(d.node_y).copy()
    res = OraGpResult()
    rows = (OraIterMetrics * max_rows)()
    L = lib()
    L.ora_eplace_run.argtypes = [_P, _P, _P, _P, _P, _P, C.c_int]
    L.ora_eplace_run(C.byref(ds.s), _ptr(x), _ptr(y), C.byref(cfg), C.byref(res),
                     C.cast(rows, _P), int(max_rows))
    return res, list(rows[:min(res.iterations, max_rows)]), x, y


def get_threads():
    return lib().ora_get_threads()


def ref_wa(d, gamma, ignore_net_degree=100, x=None, y=None):
    L = ref_lib()
    x = _f32(d.node_x if x is None else x)
    y = _f32(d.node_y if y is None else y)
    n = d.num_nodes
    grad = np.zeros(2 * d.num_movable, np.float32)
    args = [_f32(x[:n]), _f32(y[:n]), _f32(d.node_w[:n]), _f32(d.node_h[:n]),
            np.ascontiguousarray(d.node_flags[:n], np.uint8),
            np.ascontiguousarray(d.pin2node, np.int32), np.ascontiguousarray(d.pin2net, np.int32),
            _f32(d.pin_offset_x), _f32(d.pin_offset_y)]
    wl = L.ref_wa(d.num_movable, n, d.num_pins, d.num_nets, *[_ptr(a) for a in args],
                  float(gamma), int(ignore_net_degree), _ptr(grad))
    return wl, grad


def ref_hpwl(d, x=None, y=None):
    L = ref_lib()
    x = _f32(d.node_x if x is None else x)
    y = _f32(d.node_y if y is None else y)
    n = d.num_nodes
    args = [_f32(x[:n]), _f32(y[:n]), _f32(d.node_w[:n]), _f32(d.node_h[:n]),
            np.ascontiguousarray(d.node_flags[:n], np.uint8),
            np.ascontiguousarray(d.pin2node, np.int32), np.ascontiguousarray(d.pin2net, np.int32),
            _f32(d.pin_offset_x), _f32(d.pin_offset_y)]
    return L.ref_hpwl(d.num_movable, n, d.num_pins, d.num_nets, *[_ptr(a) for a in args])


# ---- quadratic initial placement (src/init/quadratic_placer.cpp) -----------
class OraQpConfig(C.Structure):
    _fields_ = [("max_iter", C.c_int32), ("tol", C.c_float), ("solver_max_iter", C.c_int32),
                ("outer_tol", C.c_float), ("min_distance", C.c_float),
                ("ignore_net_degree", C.c_int32)]


def qp_config(max_iter=20, tol=1e-6, solver_max_iter=100, outer_tol=1e-4, min_distance=1.0,
              ignore_net_degree=100):
    """The QP fields of Config (src/util/config.h:49-74, :83) with their defaults."""
    return OraQpConfig(max_iter, tol, solver_max_iter, outer_tol, min_distance,
                       ignore_net_degree)


def bound_pins(d, x, y):
    L = lib()
    L.ora_bound_pins.argtypes = [_P] * 7
    ds = design_struct(d)
    x, y = _f32(x), _f32(y)
    out = [np.zeros(d.num_nets, np.int32) for _ in range(4)]
    L.ora_bound_pins(C.byref(ds.s), _ptr(x), _ptr(y), *[_ptr(o) for o in out])
    return out


def b2b_system(d, x, y, axis, net_weight=None, min_distance=1.0, ignore_net_degree=100):
    """(diag, rhs, row_start, col, val) of one axis' Bound2Bound system."""
    L = lib()
    L.ora_b2b_system.restype = C.c_int64
    L.ora_b2b_system.argtypes = [_P, _P, _P, _P, C.c_int, C.c_float, C.c_int] + [_P] * 5
    ds = design_struct(d)
    x, y = _f32(x), _f32(y)
    w = None if net_weight is None else _f32(net_weight)
    wp = None if w is None else _ptr(w)
    nnz = L.ora_b2b_system(C.byref(ds.s), _ptr(x), _ptr(y), wp, int(axis), min_distance,
                           int(ignore_net_degree), None, None, None, None, None)
    n = d.num_movable
    diag, rhs = np.zeros(n), np.zeros(n)
    rs = np.zeros(n + 1, np.int64)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1))
    L.ora_b2b_system(C.byref(ds.s), _ptr(x), _ptr(y), wp, int(axis), min_distance,
                     int(ignore_net_degree), _ptr(diag), _ptr(rhs), _ptr(rs), _ptr(col), _ptr(val))
    return diag, rhs, rs, col[:nnz], val[:nnz]


def bicgstab(diag, rhs, rs, col, val, x0, tol=1e-6, max_iter=100):
    L = lib()
    L.ora_bicgstab.restype = C.c_double
    L.ora_bicgstab.argtypes = [C.c_int, _P, _P, _P, _P, _P, _P, C.c_double, C.c_int, _P]
    x = np.ascontiguousarray(x0, np.float64).copy()
    its = C.c_int(0)
    c = np.ascontiguousarray(col, np.int32) if len(col) else np.zeros(1, np.int32)
    v = np.ascontiguousarray(val, np.float64) if len(val) else np.zeros(1)
    err = L.ora_bicgstab(len(diag), _ptr(np.ascontiguousarray(diag, np.float64)),
                         _ptr(np.ascontiguousarray(rs, np.int64)), _ptr(c), _ptr(v),
                         _ptr(np.ascontiguousarray(rhs, np.float64)), _ptr(x), float(tol),
                         int(max_iter), C.byref(its))
    return x, err, its.value


def qp_place(d, cfg=None, net_weight=None, max_rows=64):
    """QuadraticPlacer::place: (x, y, hpwl per iteration, err per iteration)."""
    L = lib()
    L.ora_qp_place.restype = C.c_int
    L.ora_qp_place.argtypes = [_P, _P, _P, _P, _P, _P, _P, C.c_int]
    cfg = cfg or qp_config()
    ds = design_struct(d)
    x = _f32(d.node_x[:d.num_nodes]).copy()
    y = _f32(d.node_y[:d.num_nodes]).copy()
    hp, er = np.zeros(max_rows), np.zeros(max_rows)
    w = None if net_weight is None else _f32(net_weight)
    n = L.ora_qp_place(C.byref(ds.s), None if w is None else _ptr(w), C.byref(cfg), _ptr(x),
                       _ptr(y), _ptr(hp), _ptr(er), max_rows)
    return x, y, hp[:min(n, max_rows)], er[:min(n, max_rows)]
