"""ctypes binding of libsuperplace_hip.so and a Python mirror of the
reference's plug-in interface for the GP path.

* HipPoissonBackend mirrors PoissonBackend (src/numeric/poisson_backend.h:20-46):
  solve(density, phi, field_x, field_y) on nx*ny float32 grids (row-major, y
  fastest), name(), cpu_threads().
* make_poisson_backend mirrors makePoissonBackend(BackendConfig)
  (poisson_backend.h:48-60).  There is no CPU fallback in this package: a
  missing library or GPU raises.
* GpEngine is the device-resident EPlace stage (src/gp/eplace.cpp:205-526).

The library is loaded from this directory only; if it is missing every entry
point raises (no silent fallback).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import pathlib
import re

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libsuperplace_hip.so"
HEADER = HERE.parent / "include" / "superplace_hip.h"

SP_HIP_OK, SP_HIP_ENODEV, SP_HIP_EINVAL, SP_HIP_EOOM, SP_HIP_EDEVICE, SP_HIP_ENONFINITE = range(6)
SP_HIP_FLAG_FP64 = 1
SP_HIP_FLAG_BF16X3 = 2
SP_HIP_FLAG_DENSE = 4
KERNEL_FAMILIES = ["scatter", "finalize", "poisson", "wa_nets", "node_grad", "update", "sort",
                   "reduce"]

_P = C.c_void_p


class HipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"superplace_hip error {code}: {msg}")
        self.code = code


class NoDevice(HipError):
    pass


class InvalidArgument(HipError):
    """An argument or input file the library rejected (SP_HIP_EINVAL)."""


class NonFinite(HipError):
    """InvariantError("non_finite", ...) of the reference (src/util/invariant.h:10-33)."""


class SpAbacusResult(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("precondition_ok", "cells", "failed", "rows", "segments",
                                         "rounded_widths", "selfcheck_overlaps",
                                         "selfcheck_offsite")] + \
        [("disp_mean", C.c_double), ("disp_max", C.c_double)]


class SpMlgConfig(C.Structure):
    _fields_ = [("random_seed", C.c_uint64), ("moves", C.c_int32), ("init_accept", C.c_float),
                ("temp_ratio", C.c_float), ("disp_weight", C.c_float)]


class SpMlgResult(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("num_macros", "moves", "accepted", "repaired",
                                         "chained", "hot_macros")] + \
        [(k, C.c_double) for k in ("overlap_before", "overlap_after_sa", "overlap_after", "sa_ms",
                                   "t0", "typical_worsen", "hpwl_before", "hpwl_after")]


class SpHipJobInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("device_bytes", C.c_int64),
                ("arrival_s", C.c_double), ("admit_s", C.c_double), ("finish_s", C.c_double)]


class SpQpConfig(C.Structure):
    _fields_ = [("max_iter", C.c_int32), ("tol", C.c_float), ("solver_max_iter", C.c_int32),
                ("outer_tol", C.c_float), ("min_distance", C.c_float),
                ("ignore_net_degree", C.c_int32)]


class SpFillerInfo(C.Structure):
    _fields_ = [("count", C.c_int32), ("width", C.c_float), ("height", C.c_float),
                ("whitespace_area", C.c_double), ("node_area_scaled", C.c_double),
                ("total_filler_area", C.c_double)]


class SpHipDesign(C.Structure):
    _fields_ = [("num_movable", C.c_int32), ("num_nodes", C.c_int32),
                ("num_fillers", C.c_int32), ("num_pins", C.c_int32),
                ("num_nets", C.c_int32), ("num_rows", C.c_int32),
                ("node_x", _P), ("node_y", _P), ("node_w", _P), ("node_h", _P),
                ("node_flags", _P), ("pin_offset_x", _P), ("pin_offset_y", _P),
                ("pin2node", _P), ("pin2net", _P), ("net2pin_start", _P), ("flat_net2pin", _P),
                ("node2pin_start", _P), ("flat_node2pin", _P), ("row_ly", _P),
                ("row_height", _P), ("row_lx", _P), ("row_hx", _P),
                ("core_lx", C.c_float), ("core_ly", C.c_float), ("core_hx", C.c_float),
                ("core_hy", C.c_float)]


class SpHipGpConfig(C.Structure):
    _fields_ = [("target_density", C.c_float), ("target_overflow", C.c_float),
                ("cgp_target_overflow", C.c_float), ("gp_max_iter", C.c_int32),
                ("ignore_net_degree", C.c_int32), ("gp_bin_dim", C.c_int32),
                ("delta_hpwl_ref", C.c_float), ("lambda_relative", C.c_int32),
                ("delta_hpwl_relative", C.c_double), ("use_bb", C.c_int32),
                ("gp_backtrack", C.c_int32), ("filler_only_iters", C.c_int32),
                ("gp_stagnation_window", C.c_int32), ("gp_stagnation_min_gain", C.c_float),
                ("stage", C.c_int32), ("precision", C.c_int32), ("max_backtrack", C.c_int32),
                ("loop_mode", C.c_int32), ("graph_slots", C.c_int32)]


class SpHipIterMetrics(C.Structure):
    _fields_ = [("iter", C.c_int32), ("evaluations", C.c_int32)] + [
        (k, C.c_double) for k in ("hpwl", "overflow", "lambda_", "gamma", "step_size",
                                  "grad_norm_wl", "grad_norm_den", "eval_lambda", "eval_gamma",
                                  "delta_hpwl", "lambda_mu", "elapsed_ms")]


class SpHipGpResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32),
                ("stagnated", C.c_int32), ("rollback_iter", C.c_int32),
                ("evaluations", C.c_int64), ("hpwl", C.c_double), ("overflow", C.c_double),
                ("initial_hpwl", C.c_double), ("initial_overflow", C.c_double),
                ("initial_lambda", C.c_double), ("seconds_loop", C.c_double)]


class SpHipEvalScalars(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("tau", "wl", "hpwl", "sum_wl", "sum_den", "dx2",
                                          "dg2", "sy", "gg", "nonfinite")]


class SpHipEvalDebug(C.Structure):
    _fields_ = [(k, _P) for k in ("node_q", "filler_q", "node_density", "filler_density", "rho",
                                  "field_x", "field_y", "force", "grad_wl", "grad")]


class SpHipShapes(C.Structure):
    _fields_ = [("scatter_chunk", C.c_int32), ("sort_period", C.c_int32),
                ("col_pairs", C.c_int32), ("graph_slots", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class SpHipGpInfo(C.Structure):
    _fields_ = [("dim", C.c_int32), ("n_opt", C.c_int32), ("frac_bits", C.c_int32),
                ("precision", C.c_int32), ("step_x", C.c_double), ("step_y", C.c_double),
                ("origin_x", C.c_double), ("origin_y", C.c_double),
                ("movable_area_scaled", C.c_double), ("device_bytes", C.c_int64),
                ("shapes", SpHipShapes)]


_lib = None


def declared_functions() -> list[str]:
    """Function names declared in include/superplace_hip.h."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(sp_(?:hip|bookshelf|filler|qp|metrics|abacus|macro|rng)_\w+)\s*\(", text)))


def lib():
    """Loads the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                           "(python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(str(LIB_PATH))
    L.sp_hip_last_error.restype = C.c_char_p
    L.sp_hip_poisson_name.restype = C.c_char_p
    L.sp_hip_poisson_create.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_int,
                                        C.POINTER(_P)]
    L.sp_hip_poisson_solve.argtypes = [_P] * 5
    L.sp_hip_poisson_solve_dev.argtypes = [_P] * 3
    L.sp_hip_poisson_solve_async.argtypes = [_P] * 4
    L.sp_hip_poisson_name.argtypes = [_P]
    L.sp_hip_poisson_destroy.argtypes = [_P]
    L.sp_hip_gp_create.argtypes = [_P, _P, C.POINTER(_P)]
    L.sp_hip_gp_eval.argtypes = [_P, _P, _P, C.c_float, C.c_float, _P, _P]
    L.sp_hip_gp_run.argtypes = [_P, _P, _P, C.c_int]
    L.sp_hip_gp_get_positions.argtypes = [_P, _P, _P]
    L.sp_hip_gp_info_get.argtypes = [_P, _P]
    L.sp_hip_gp_opt_index.argtypes = [_P, _P]
    L.sp_hip_gp_set_profiling.argtypes = [_P, C.c_int]
    L.sp_hip_gp_kernel_stats.argtypes = [_P, _P, _P, _P, _P]
    L.sp_hip_gp_destroy.argtypes = [_P]
    L.sp_hip_gemm.argtypes = [C.c_int, C.c_int, _P, _P, _P, C.c_int, C.c_int, C.c_int]
    L.sp_hip_device_count.argtypes = [C.POINTER(C.c_int)]
    L.sp_bookshelf_read.argtypes = [C.c_char_p, C.POINTER(_P)]
    L.sp_bookshelf_design.argtypes = [_P, _P]
    L.sp_bookshelf_extra.argtypes = [_P, _P, _P, _P, _P, _P]
    L.sp_bookshelf_names.argtypes = [_P, C.c_int, _P, C.c_int64, C.POINTER(C.c_int64)]
    L.sp_bookshelf_read_pl.argtypes = [_P, C.c_char_p]
    L.sp_bookshelf_write_pl.argtypes = [_P, C.c_char_p, _P, _P]
    L.sp_bookshelf_write_pl_state.argtypes = [_P, C.c_char_p, _P, _P, _P, _P]
    L.sp_bookshelf_destroy.argtypes = [_P]
    L.sp_qp_place.argtypes = [_P, _P, _P, _P, _P, _P, C.c_int, _P]
    L.sp_qp_b2b_system.argtypes = [_P, _P, C.c_int, C.c_float, C.c_int32, _P, _P, _P, _P, _P,
                                   C.c_int64, _P]
    L.sp_abacus_legalize.argtypes = [_P, _P, _P, _P, C.c_int, _P]
    L.sp_abacus_segments.argtypes = [_P, _P, _P, _P, _P, C.c_int32, _P]
    L.sp_hip_gp_footprint.argtypes = [_P, _P, _P]
    L.sp_hip_server_create.argtypes = [C.c_int64, C.c_int32, C.POINTER(_P)]
    L.sp_hip_server_submit.argtypes = [_P, _P, _P, C.c_double, _P]
    L.sp_hip_server_pump.argtypes = [_P, _P, _P, _P, _P]
    L.sp_hip_server_drain.argtypes = [_P, C.c_double]
    L.sp_hip_server_now.argtypes = [_P, _P]
    L.sp_hip_server_job.argtypes = [_P, C.c_int32, _P, _P, _P, _P]
    L.sp_hip_server_release.argtypes = [_P, C.c_int32]
    L.sp_hip_server_destroy.argtypes = [_P]
    L.sp_hip_server_destroy.restype = None
    L.sp_macro_legalize.argtypes = [_P, _P, C.c_float, C.c_float, _P, _P, _P, _P, _P, _P, _P, _P,
                                    _P]
    L.sp_metrics_create.argtypes = [C.c_char_p, C.POINTER(_P)]
    L.sp_metrics_begin_stage.argtypes = [_P, C.c_char_p]
    L.sp_metrics_push.argtypes = [_P, _P, C.c_int, _P]
    L.sp_metrics_end_stage.argtypes = [_P, C.c_double]
    L.sp_metrics_record_duration.argtypes = [_P, C.c_char_p, C.c_double]
    L.sp_metrics_set_number.argtypes = [_P, C.c_char_p, C.c_double]
    L.sp_metrics_set_string.argtypes = [_P, C.c_char_p, C.c_char_p]
    L.sp_metrics_gp_result.argtypes = [_P, C.c_char_p, _P, C.c_double, C.c_int]
    L.sp_metrics_write_summary.argtypes = [_P, C.c_char_p]
    L.sp_metrics_destroy.argtypes = [_P]
    L.sp_filler_plan.argtypes = [_P, C.c_float, _P]
    L.sp_filler_scatter.argtypes = [_P, C.c_int, C.c_float, C.c_float, C.c_uint64, _P, _P]
    L.sp_rng_uniform.argtypes = [C.c_uint64, C.c_int64, C.c_float, C.c_float, _P]
    _lib = L
    return L


def check(code: int) -> None:
    if code == SP_HIP_OK:
        return
    msg = lib().sp_hip_last_error().decode()
    if code == SP_HIP_ENODEV:
        raise NoDevice(code, msg)
    if code == SP_HIP_ENONFINITE:
        raise NonFinite(code, msg)
    if code == SP_HIP_EINVAL:
        raise InvalidArgument(code, msg)
    raise HipError(code, msg)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().sp_hip_device_count(C.byref(n)))
    return n.value


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


# ------------------------------------------------------------------------ Poisson
@dataclasses.dataclass
class BackendConfig:
    """BackendConfig, src/numeric/poisson_backend.h:48-55 (+ fp64 for the FP64 path)."""
    nx: int = 0
    ny: int = 0
    step_x: float = 1.0
    step_y: float = 1.0
    prefer_gpu: bool = True
    fftw_threads: int = 4
    fp64: bool = False
    # the transforms (no counterpart in BackendConfig): dense products on
    # every grid size, and / or those products as split-BF16 products (FP32)
    dense: bool = False
    bf16x3: bool = False


class HipPoissonBackend:
    """PoissonBackend on the GPU (LDS FFT passes, or dense transforms on the
    matrix cores for grids that are not a power of two): solve() reads
    density and writes phi / field_x / field_y (the field xi = -grad phi,
    not a gradient)."""

    def __init__(self, cfg: BackendConfig):
        self.nx, self.ny = int(cfg.nx), int(cfg.ny)
        h = _P()
        flags = ((SP_HIP_FLAG_FP64 if cfg.fp64 else 0) | (SP_HIP_FLAG_DENSE if cfg.dense else 0) |
                 (SP_HIP_FLAG_BF16X3 if cfg.bf16x3 else 0))
        check(lib().sp_hip_poisson_create(self.nx, self.ny, float(cfg.step_x), float(cfg.step_y),
                                          flags, C.byref(h)))
        self._h = h

    def _span(self, a, writable):
        a = np.asarray(a)
        if a.dtype != np.float32 or not a.flags.c_contiguous or a.size != self.nx * self.ny:
            raise ValueError("grid span dimension mismatch")   # cpu_poisson.cpp:220-223
        if writable and not a.flags.writeable:
            raise ValueError("output span is read-only")
        return a

    def solve(self, density, phi, field_x, field_y) -> None:
        spans = [self._span(density, False), self._span(phi, True), self._span(field_x, True),
                 self._span(field_y, True)]
        check(lib().sp_hip_poisson_solve(self._h, *[_ptr(s) for s in spans]))

    def solve_device(self, density_ptr: int, out3_ptr: int) -> None:
        check(lib().sp_hip_poisson_solve_dev(self._h, _P(density_ptr), _P(out3_ptr)))

    def solve_async(self, density_ptr: int, out3_ptr: int, stream: int | None = None) -> None:
        """solve_device ordered on the caller's HIP stream (None: the default
        stream), without waiting (sp_hip_poisson_solve_async)."""
        check(lib().sp_hip_poisson_solve_async(self._h, _P(density_ptr), _P(out3_ptr),
                                               _P(stream) if stream else None))

    def name(self) -> str:
        return lib().sp_hip_poisson_name(self._h).decode()

    def cpu_threads(self) -> int:
        return 0

    def close(self):
        if getattr(self, "_h", None):
            lib().sp_hip_poisson_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_poisson_backend(cfg: BackendConfig) -> HipPoissonBackend:
    """makePoissonBackend for this package: always the HIP backend (raises
    NoDevice without a GPU instead of falling back to a CPU solver)."""
    return HipPoissonBackend(cfg)


def gemm(a_ptrs, b_ptrs, c_ptrs, m, n, k, fp64=False, bf16x3=False):
    """Batched C = A B^T on device pointers (A m x k, B n x k, row-major,
    multiples of 64); bf16x3: FP32 as split-BF16 products."""
    batch = len(a_ptrs)
    arr = lambda ps: (_P * batch)(*[_P(p) for p in ps])
    flags = (SP_HIP_FLAG_FP64 if fp64 else 0) | (SP_HIP_FLAG_BF16X3 if bf16x3 else 0)
    check(lib().sp_hip_gemm(flags, batch, arr(a_ptrs), arr(b_ptrs), arr(c_ptrs), m, n, k))


# ------------------------------------------------------------------------ GP engine
@dataclasses.dataclass
class GpConfig:
    """GP fields of Config (src/util/config.h:77-118), default.json values."""
    target_density: float = 1.0
    target_overflow: float = 0.10
    cgp_target_overflow: float = 0.10
    gp_max_iter: int = 2000
    ignore_net_degree: int = 100
    gp_bin_dim: int = 0
    delta_hpwl_ref: float = 3.5e5
    lambda_relative: int = 0
    delta_hpwl_relative: float = 0.01
    use_bb: int = 1
    gp_backtrack: int = 1
    filler_only_iters: int = 20
    gp_stagnation_window: int = 200
    gp_stagnation_min_gain: float = 0.02
    stage: int = 0
    precision: int = 32
    max_backtrack: int = 10
    loop_mode: int = 2        # 0 host decisions, 1 device decisions, 2 device decisions + hipGraph
    graph_slots: int = 0      # iterations per graph launch in loop_mode 2 (0: the tuned shapes)

    def struct(self) -> SpHipGpConfig:
        return SpHipGpConfig(**dataclasses.asdict(self))


class _DesignArrays:
    def __init__(self, d):
        f32 = lambda a: np.ascontiguousarray(a, np.float32)
        i32 = lambda a: np.ascontiguousarray(a, np.int32)
        self.arrs = dict(node_x=f32(d.node_x), node_y=f32(d.node_y), node_w=f32(d.node_w),
                         node_h=f32(d.node_h),
                         node_flags=np.ascontiguousarray(d.node_flags, np.uint8),
                         pin_offset_x=f32(d.pin_offset_x), pin_offset_y=f32(d.pin_offset_y),
                         pin2node=i32(d.pin2node), pin2net=i32(d.pin2net),
                         net2pin_start=i32(d.net2pin_start), flat_net2pin=i32(d.flat_net2pin),
                         node2pin_start=i32(d.node2pin_start), flat_node2pin=i32(d.flat_node2pin),
                         row_ly=f32(d.row_ly), row_height=f32(d.row_height), row_lx=f32(d.row_lx),
                         row_hx=f32(d.row_hx))
        self.s = SpHipDesign(num_movable=d.num_movable, num_nodes=d.num_nodes,
                             num_fillers=d.num_fillers, num_pins=d.num_pins,
                             num_nets=d.num_nets, num_rows=d.num_rows,
                             core_lx=d.core[0], core_ly=d.core[1], core_hx=d.core[2],
                             core_hy=d.core[3],
                             **{k: _ptr(v) for k, v in self.arrs.items()})


def run_batch(designs, cfgs=None, slots=8, positions=True):
    """Many FP32 stages advanced together in batched launches
    (sp_hip_gp_run_batch): one result dict per design, and its positions
    (x, y of all nodes) when `positions`; equal to each design's own run."""
    n = len(designs)
    cfgs = list(cfgs) if cfgs is not None else [GpConfig() for _ in designs]
    arrays = [_DesignArrays(d) for d in designs]
    ds = (SpHipDesign * max(n, 1))(*[a.s for a in arrays])
    cs = (SpHipGpConfig * max(n, 1))(*[c.struct() for c in cfgs])
    res = (SpHipGpResult * max(n, 1))()
    xs = [np.zeros(d.total_nodes, np.float32) for d in designs] if positions else []
    ys = [np.zeros(d.total_nodes, np.float32) for d in designs] if positions else []
    px = (_P * max(n, 1))(*[_ptr(a) for a in xs]) if positions else None
    py = (_P * max(n, 1))(*[_ptr(a) for a in ys]) if positions else None
    L = lib()
    L.sp_hip_gp_run_batch.argtypes = [C.c_int, _P, _P, C.c_int, _P, _P, _P]
    check(L.sp_hip_gp_run_batch(n, C.cast(ds, _P), C.cast(cs, _P), int(slots), C.cast(res, _P),
                                C.cast(px, _P) if px is not None else None,
                                C.cast(py, _P) if py is not None else None))
    out = [{k: getattr(res[i], k) for k, _ in SpHipGpResult._fields_} for i in range(n)]
    return (out, xs, ys) if positions else out


class GpEngine:
    """One placement stage on the GPU (EPlace constructed for one stage)."""

    def __init__(self, design, cfg: GpConfig | None = None):
        self.design = design
        self.cfg = cfg or GpConfig()
        da = _DesignArrays(design)
        cs = self.cfg.struct()
        h = _P()
        check(lib().sp_hip_gp_create(C.byref(da.s), C.byref(cs), C.byref(h)))
        self._h = h
        info = SpHipGpInfo()
        check(lib().sp_hip_gp_info_get(self._h, C.byref(info)))
        self.info = {k: getattr(info, k) for k, _ in SpHipGpInfo._fields_}
        self.dim = info.dim
        self.n_opt = info.n_opt
        self.frac_bits = info.frac_bits
        idx = np.zeros(max(self.n_opt, 1), np.int32)
        check(lib().sp_hip_gp_opt_index(self._h, _ptr(idx)))
        self.opt_idx = idx[:self.n_opt]

    def eval(self, x=None, y=None, lam=1.0, gamma=1.0, debug=False):
        """One objAndGrad at positions x, y of all nodes (default: the design's)."""
        d = self.design
        x = np.ascontiguousarray(d.node_x if x is None else x, np.float32)
        y = np.ascontiguousarray(d.node_y if y is None else y, np.float32)
        sc = SpHipEvalScalars()
        out = None
        dbg = None
        if debug:
            nb, n2 = self.dim * self.dim, 2 * self.n_opt
            out = dict(node_q=np.zeros(nb, np.int64), filler_q=np.zeros(nb, np.int64),
                       node_density=np.zeros(nb), filler_density=np.zeros(nb), rho=np.zeros(nb),
                       field_x=np.zeros(nb), field_y=np.zeros(nb), force=np.zeros(n2),
                       grad_wl=np.zeros(n2), grad=np.zeros(n2))
            dbg = SpHipEvalDebug(**{k: _ptr(v) for k, v in out.items()})
        check(lib().sp_hip_gp_eval(self._h, _ptr(x), _ptr(y), float(lam), float(gamma),
                                   C.byref(sc), C.byref(dbg) if dbg is not None else None))
        scal = {k: getattr(sc, k) for k, _ in SpHipEvalScalars._fields_}
        return (scal, out) if debug else scal

    def run(self, max_rows=None):
        """EPlace::run(stage) -> (result dict, metrics rows as list of dicts)."""
        max_rows = max_rows or (max(self.cfg.gp_max_iter, self.cfg.filler_only_iters) + 1)
        rows = (SpHipIterMetrics * max_rows)()
        res = SpHipGpResult()
        check(lib().sp_hip_gp_run(self._h, C.byref(res), C.cast(rows, _P), max_rows))
        r = {k: getattr(res, k) for k, _ in SpHipGpResult._fields_}
        m = [{k: getattr(row, k) for k, _ in SpHipIterMetrics._fields_}
             for row in rows[:min(res.iterations, max_rows)]]
        return r, m

    def device_bytes(self) -> int:
        """Device memory the instance has allocated (set-up and loop buffers)."""
        info = SpHipGpInfo()
        check(lib().sp_hip_gp_info_get(self._h, C.byref(info)))
        return int(info.device_bytes)

    def positions(self):
        t = self.design.total_nodes
        x, y = np.zeros(t, np.float32), np.zeros(t, np.float32)
        check(lib().sp_hip_gp_get_positions(self._h, _ptr(x), _ptr(y)))
        return x, y

    def set_profiling(self, mask=(1 << len(KERNEL_FAMILIES)) - 1):
        """Time the kernel families in `mask` (bit f = KERNEL_FAMILIES[f]) with HIP events."""
        check(lib().sp_hip_gp_set_profiling(self._h, int(mask)))

    def kernel_stats(self):
        n = len(KERNEL_FAMILIES)
        ms, cnt = np.zeros(n), np.zeros(n, np.int64)
        by, fl = np.zeros(n), np.zeros(n)
        check(lib().sp_hip_gp_kernel_stats(self._h, _ptr(ms), _ptr(cnt), _ptr(by), _ptr(fl)))
        return {f: dict(ms=float(ms[i]), launches=int(cnt[i]), bytes=float(by[i]),
                        flops=float(fl[i])) for i, f in enumerate(KERNEL_FAMILIES)}

    def close(self):
        if getattr(self, "_h", None):
            lib().sp_hip_gp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
