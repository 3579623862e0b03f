"""Python mirror of the solver-slot interface of include/bda_bicgstab.h.

The reference's host library (src/sda_app/common/fpga_functions_bicgstab.hpp:30-137)
is driven in this order by its caller: set up buffers once, copy the matrix,
set kernel parameters (max iterations, precision), run, read back the status
and map the results.  ``BicgstabSolver`` keeps that order:

    s = BicgstabSolver(bs, n, nnzb_max)       # fpga_setup_*_datamem
    s.upload_bcsr(rowptr, colidx, values)     # fpga_copy_host_datamem
    s.analyse(); s.factor()                   # colouring / ILU0 (caller's job in the reference)
    x, res = s.solve(b, tol, maxit)           # fpga_set_kernel_parameters + fpga_kernel_run
                                              # + fpga_copy_from_device_debugbuf + fpga_map_results

Errors: every call whose C entry point returns non-zero raises
``BdaError`` with the library's message (the reference returns 1 and prints
"ERROR: <function>: ..." instead).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from ._lib import BdaGridSpec, BdaLimits, BdaResult, lib


class BdaError(RuntimeError):
    pass


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


@dataclass
class SolveResult:
    iterations: float
    half_iterations: int
    last_norm_idx: int
    norms: list
    final_norm: float
    rel_residual: float
    converged: bool
    no_results: bool
    max_iter_reached: bool
    breakdown: bool
    nan_detected: bool
    aborted: bool
    outer_iterations: int
    inner_half_iterations: int
    solve_ms: float

    @staticmethod
    def from_c(r: BdaResult) -> "SolveResult":
        return SolveResult(
            iterations=r.iterations, half_iterations=r.half_iterations,
            last_norm_idx=r.last_norm_idx, norms=list(r.norms), final_norm=r.final_norm,
            rel_residual=r.rel_residual, converged=bool(r.converged),
            no_results=bool(r.no_results), max_iter_reached=bool(r.max_iter_reached),
            breakdown=bool(r.breakdown), nan_detected=bool(r.nan_detected),
            aborted=bool(r.aborted), outer_iterations=r.outer_iterations,
            inner_half_iterations=r.inner_half_iterations, solve_ms=r.solve_ms)


# --------------------------------------------------------------------------
@dataclass
class GridSpec:
    nx: int
    ny: int
    nz: int
    bs: int = 3
    seed: int = 0
    perm_sigma: float = 1.0
    layer_corr: float = 0.0
    inactive_frac: float = 0.0
    nnc_frac: float = 0.0
    nnc_dist: int = 3
    pert_max: float = 0.2
    upwind: float = 0.3
    diag_noise: float = 0.05

    def to_c(self) -> BdaGridSpec:
        return BdaGridSpec(self.nx, self.ny, self.nz, self.bs, self.seed, self.perm_sigma,
                           self.layer_corr, self.inactive_frac, self.nnc_frac, self.nnc_dist,
                           self.pert_max, self.upwind, self.diag_noise)


@dataclass
class BlockSystem:
    bs: int
    n: int
    rowptr: np.ndarray
    colidx: np.ndarray
    values: np.ndarray   # (nnzb, bs, bs)
    b: np.ndarray        # (n*bs,)

    @property
    def nnzb(self) -> int:
        return int(self.rowptr[-1])

    @property
    def N(self) -> int:
        return self.n * self.bs


def generate(spec: GridSpec) -> BlockSystem:
    """Synthetic reservoir-like block system (see fpga_amd/csrc/grid_gen.cpp)."""
    g = C.c_void_p()
    cs = spec.to_c()
    if lib.bda_grid_create(C.byref(cs), C.byref(g)) != 0:
        raise BdaError(f"grid generation failed for {spec}")
    try:
        n, nnzb, bs = C.c_int32(), C.c_int32(), C.c_int32()
        lib.bda_grid_sizes(g, C.byref(n), C.byref(nnzb), C.byref(bs))
        rowptr = np.empty(n.value + 1, np.int32)
        colidx = np.empty(nnzb.value, np.int32)
        values = np.empty((nnzb.value, bs.value, bs.value), np.float64)
        b = np.empty(n.value * bs.value, np.float64)
        lib.bda_grid_copy(g, _p(rowptr), _p(colidx), _p(values), _p(b))
    finally:
        lib.bda_grid_destroy(g)
    return BlockSystem(bs.value, n.value, rowptr, colidx, values, b)


def colour_pattern(rowptr, colidx, level_scheduling: bool = False):
    """Host-only colouring: (colour[n], perm[n], colour_ptr[C+1])."""
    rowptr = _i32(rowptr)
    colidx = _i32(colidx)
    n = rowptr.size - 1
    colour = np.empty(n, np.int32)
    perm = np.empty(n, np.int32)
    cptr = np.empty(n + 1, np.int32)
    nc = C.c_int32()
    rc = lib.bda_colour_pattern(n, _p(rowptr), _p(colidx), int(level_scheduling), _p(colour),
                                _p(perm), _p(cptr), C.byref(nc))
    if rc != 0:
        raise BdaError("colouring failed")
    return colour, perm, cptr[: nc.value + 1].copy()


def set_debug_level(level: int) -> int:
    """Set the library's diagnostic level (0 silent, 1 summaries and errors,
    2 per-colour detail); returns the previous level.  The initial level is
    the environment variable BDA_DEBUG_LEVEL."""
    prev = int(lib.bda_get_debug_level())
    if lib.bda_set_debug_level(int(level)) != 0:
        raise BdaError(f"debug level must be >= 0, got {level}")
    return prev


def debug_level() -> int:
    return int(lib.bda_get_debug_level())


# --------------------------------------------------------------------------
class BicgstabSolver:
    """One solver handle: one HIP stream, device buffers owned by the handle."""

    def __init__(self, bs: int, n: int, nnzb_max: int):
        self._h = C.c_void_p()
        rc = lib.bda_create(bs, n, nnzb_max, C.byref(self._h))
        if rc != 0:
            raise BdaError(f"bda_create(bs={bs}, n={n}, nnzb_max={nnzb_max}) failed ({rc})")
        self.bs, self.n, self.nnzb_max = bs, n, nnzb_max

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = lib.bda_last_error(self._h)
            raise BdaError(f"{what}: {msg.decode() if msg else 'error'}")

    def close(self):
        if self._h:
            lib.bda_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self) -> int:
        return self._h.value

    @property
    def N(self) -> int:
        return self.n * self.bs

    # -- lifecycle --------------------------------------------------------
    def query(self) -> BdaLimits:
        lim = BdaLimits()
        self._check(lib.bda_query(self._h, C.byref(lim)), "bda_query")
        return lim

    def upload_bcsr(self, rowptr, colidx, values):
        self._rowptr, self._colidx = _i32(rowptr), _i32(colidx)
        v = _f64(values)
        self._check(lib.bda_upload_bcsr(self._h, _p(self._rowptr), _p(self._colidx), _p(v)),
                    "bda_upload_bcsr")

    def upload_system(self, sysm: BlockSystem):
        self.upload_bcsr(sysm.rowptr, sysm.colidx, sysm.values)

    def upload_factored(self, colour_ptr, rowptr, colidx, values, lu_values, dinv):
        """A colour-ordered system with the caller's own L, U (on A's pattern)
        and D^-1; the handle then solves with exactly these factors."""
        cp = _i32(colour_ptr)
        self._rowptr, self._colidx = _i32(rowptr), _i32(colidx)
        v, lu, d = _f64(values), _f64(lu_values), _f64(dinv)
        self._check(lib.bda_upload_factored(self._h, cp.size - 1, _p(cp), _p(self._rowptr), _p(self._colidx),
                                            _p(v), _p(lu), _p(d)), "bda_upload_factored")

    def update_values(self, values):
        v = _f64(values)
        self._check(lib.bda_update_values(self._h, _p(v)), "bda_update_values")

    def update_values_dev(self, ptr: int):
        self._check(lib.bda_update_values_dev(self._h, C.c_void_p(ptr)), "bda_update_values_dev")

    @property
    def values_generation(self) -> int:
        g = C.c_uint64()
        self._check(lib.bda_values_generation(self._h, C.byref(g)), "bda_values_generation")
        return g.value

    def analyse(self, level_scheduling: bool = False):
        self._check(lib.bda_analyse(self._h, int(level_scheduling)), "bda_analyse")

    @property
    def num_colours(self) -> int:
        nc = C.c_int32()
        self._check(lib.bda_num_colours(self._h, C.byref(nc)), "bda_num_colours")
        return nc.value

    def colouring(self):
        nc = self.num_colours
        colour = np.empty(self.n, np.int32)
        perm = np.empty(self.n, np.int32)
        cptr = np.empty(nc + 1, np.int32)
        self._check(lib.bda_get_colouring(self._h, _p(colour), _p(perm), _p(cptr)), "bda_get_colouring")
        return colour, perm, cptr

    def factor(self):
        self._check(lib.bda_factor(self._h), "bda_factor")

    def prepare_mixed(self):
        """Make the precision form's FP32 copies of the factored system now."""
        self._check(lib.bda_prepare_mixed(self._h), "bda_prepare_mixed")

    def set_options(self, timeout_s: float = 0.0, abs_tol: bool = False, poll_iters: int = 1):
        self._check(lib.bda_set_options(self._h, timeout_s, int(abs_tol), poll_iters), "bda_set_options")

    # -- solves (original ordering) ----------------------------------------
    def solve(self, b, x0=None, tol: float = 1e-6, maxit: int = 1000):
        b = _f64(b)
        x = np.zeros(self.N) if x0 is None else _f64(x0).copy()
        r = BdaResult()
        self._check(lib.bda_solve(self._h, _p(b), _p(x), tol, maxit, C.byref(r)), "bda_solve")
        return x, SolveResult.from_c(r)

    def solve_dev(self, b_ptr: int, x_ptr: int, tol: float, maxit: int) -> SolveResult:
        r = BdaResult()
        self._check(lib.bda_solve_dev(self._h, C.c_void_p(b_ptr), C.c_void_p(x_ptr), tol, maxit, C.byref(r)),
                    "bda_solve_dev")
        return SolveResult.from_c(r)

    def residual(self) -> np.ndarray:
        out = np.empty(self.N)
        self._check(lib.bda_get_residual(self._h, _p(out)), "bda_get_residual")
        return out

    def solve_mixed(self, b, x0=None, tol: float = 1e-6, max_outer: int = 20, inner_tol: float = 1e-4,
                    inner_maxit: int = 1000, vec32: bool = True):
        b = _f64(b)
        x = np.zeros(self.N) if x0 is None else _f64(x0).copy()
        r = BdaResult()
        self._check(lib.bda_solve_mixed(self._h, _p(b), _p(x), tol, max_outer, inner_tol, inner_maxit,
                                        int(vec32), C.byref(r)), "bda_solve_mixed")
        return x, SolveResult.from_c(r)

    def solve_mixed_dev(self, b_ptr: int, x_ptr: int, tol: float = 1e-6, max_outer: int = 20,
                        inner_tol: float = 1e-4, inner_maxit: int = 1000, vec32: bool = True) -> SolveResult:
        r = BdaResult()
        self._check(lib.bda_solve_mixed_dev(self._h, C.c_void_p(b_ptr), C.c_void_p(x_ptr), tol, max_outer,
                                            inner_tol, inner_maxit, int(vec32), C.byref(r)),
                    "bda_solve_mixed_dev")
        return SolveResult.from_c(r)

    def set_refinement(self, adaptive: bool = True, eta: float = 0.9):
        self._check(lib.bda_set_refinement(self._h, int(adaptive), eta), "bda_set_refinement")

    def adjoint_solve(self, w, lam0=None, tol: float = 1e-10, maxit: int = 1000):
        w = _f64(w)
        lam = np.zeros(self.N) if lam0 is None else _f64(lam0).copy()
        r = BdaResult()
        self._check(lib.bda_adjoint_solve(self._h, _p(w), _p(lam), tol, maxit, C.byref(r)), "bda_adjoint_solve")
        return lam, SolveResult.from_c(r)

    def adjoint_solve_dev(self, w_ptr: int, lam_ptr: int, tol: float, maxit: int) -> SolveResult:
        r = BdaResult()
        self._check(lib.bda_adjoint_solve_dev(self._h, C.c_void_p(w_ptr), C.c_void_p(lam_ptr), tol, maxit,
                                              C.byref(r)), "bda_adjoint_solve_dev")
        return SolveResult.from_c(r)

    def adjoint_solve_mixed(self, w, lam0=None, tol: float = 1e-10, max_outer: int = 20, inner_tol: float = 1e-4,
                            inner_maxit: int = 1000, vec32: bool = True):
        """A^T lambda = w in the precision form (FP64 refinement around FP32-value inner solves)."""
        w = _f64(w)
        lam = np.zeros(self.N) if lam0 is None else _f64(lam0).copy()
        r = BdaResult()
        self._check(lib.bda_adjoint_solve_mixed(self._h, _p(w), _p(lam), tol, max_outer, inner_tol, inner_maxit,
                                                int(vec32), C.byref(r)), "bda_adjoint_solve_mixed")
        return lam, SolveResult.from_c(r)

    def adjoint_solve_mixed_dev(self, w_ptr: int, lam_ptr: int, tol: float = 1e-10, max_outer: int = 20,
                                inner_tol: float = 1e-4, inner_maxit: int = 1000, vec32: bool = True) -> SolveResult:
        r = BdaResult()
        self._check(lib.bda_adjoint_solve_mixed_dev(self._h, C.c_void_p(w_ptr), C.c_void_p(lam_ptr), tol, max_outer,
                                                    inner_tol, inner_maxit, int(vec32), C.byref(r)),
                    "bda_adjoint_solve_mixed_dev")
        return SolveResult.from_c(r)

    def grad_values(self, lam, x) -> np.ndarray:
        lam, x = _f64(lam), _f64(x)
        nnzb = int(self._rowptr[-1])
        g = np.empty((nnzb, self.bs, self.bs))
        self._check(lib.bda_grad_values(self._h, _p(lam), _p(x), _p(g)), "bda_grad_values")
        return g

    def grad_values_dev(self, lam_ptr: int, x_ptr: int, grad_ptr: int):
        self._check(lib.bda_grad_values_dev(self._h, C.c_void_p(lam_ptr), C.c_void_p(x_ptr),
                                            C.c_void_p(grad_ptr)), "bda_grad_values_dev")

    # -- components (colour-ordered numbering) -----------------------------
    def spmv_perm(self, x, transpose: bool = False) -> np.ndarray:
        x = _f64(x)
        y = np.empty(self.N)
        self._check(lib.bda_spmv_perm(self._h, _p(x), _p(y), int(transpose)), "bda_spmv_perm")
        return y

    def precond_apply_perm(self, p, transpose: bool = False) -> np.ndarray:
        p = _f64(p)
        z = np.empty(self.N)
        self._check(lib.bda_precond_apply_perm(self._h, _p(p), _p(z), int(transpose)), "bda_precond_apply_perm")
        return z

    _PREC_DTYPE = {0: np.float64, 1: np.float32, 2: np.float64}

    def spmv_perm_ex(self, x, prec: int = 0, transpose: bool = False, w=None, red: int = 0):
        """y = A x in kernel form ``prec`` (0 FP64, 1 FP32 values and vectors,
        2 FP32 values with FP64 vectors); red = 1 or 3 also returns the fused
        dots (y.w; y.w and y.y) of the solve's SpMV instances."""
        dt = self._PREC_DTYPE[prec]
        x = np.ascontiguousarray(x, dt)
        y = np.empty(self.N, dt)
        wv = None if w is None else np.ascontiguousarray(w, dt)
        dots = (C.c_double * 2)()
        self._check(lib.bda_spmv_perm_ex(self._h, prec, _p(x), _p(y), int(transpose), _p(wv), red, dots),
                    "bda_spmv_perm_ex")
        return (y, (dots[0], dots[1])) if red else y

    def spmv_dev(self, x_ptr: int, y_ptr: int, prec: int = 0, transpose: bool = False, w_ptr: int = 0,
                 red: int = 0):
        """y = A x on device vectors (colour-ordered); returns the fused dots."""
        dots = (C.c_double * 2)()
        self._check(lib.bda_spmv_dev(self._h, prec, C.c_void_p(x_ptr), C.c_void_p(y_ptr), int(transpose),
                                     C.c_void_p(w_ptr) if w_ptr else None, red, dots), "bda_spmv_dev")
        return dots[0], dots[1]

    def precond_apply_perm_ex(self, p, prec: int = 0, transpose: bool = False) -> np.ndarray:
        dt = self._PREC_DTYPE[prec]
        p = np.ascontiguousarray(p, dt)
        z = np.empty(self.N, dt)
        self._check(lib.bda_precond_apply_perm_ex(self._h, prec, _p(p), _p(z), int(transpose)),
                    "bda_precond_apply_perm_ex")
        return z

    def precond_launches(self, transpose: bool = False) -> int:
        """Kernel launches of one preconditioner application."""
        k = C.c_int32()
        self._check(lib.bda_precond_launches(self._h, int(transpose), C.byref(k)), "bda_precond_launches")
        return k.value

    def factors(self):
        nl, nu = C.c_int32(), C.c_int32()
        self._check(lib.bda_factor_sizes(self._h, C.byref(nl), C.byref(nu)), "bda_factor_sizes")
        bs = self.bs
        lr, ur = np.empty(self.n + 1, np.int32), np.empty(self.n + 1, np.int32)
        lc, uc = np.empty(nl.value, np.int32), np.empty(nu.value, np.int32)
        lv, uv = np.empty((nl.value, bs, bs)), np.empty((nu.value, bs, bs))
        dinv = np.empty((self.n, bs, bs))
        self._check(lib.bda_get_factors(self._h, _p(lr), _p(lc), _p(lv), _p(ur), _p(uc), _p(uv), _p(dinv)),
                    "bda_get_factors")
        return (lr, lc, lv), (ur, uc, uv), dinv

    # -- measurement -----------------------------------------------------
    def set_profiling(self, on: bool):
        self._check(lib.bda_set_profiling(self._h, int(on)), "bda_set_profiling")

    def kernel_stats(self, kind: int):
        n, ms = C.c_int64(), C.c_double()
        self._check(lib.bda_get_kernel_stats(self._h, kind, C.byref(n), C.byref(ms)), "bda_get_kernel_stats")
        return n.value, ms.value

    def kernel_times(self, kind: int, capacity: int = 1 << 16) -> np.ndarray:
        """Durations (ms) of every launch of one kind counted since profiling was switched on."""
        out = np.empty(max(capacity, 1), dtype=np.float32)
        n = C.c_int32()
        self._check(lib.bda_get_kernel_times(self._h, kind, out.ctypes.data_as(C.c_void_p), capacity, C.byref(n)),
                    "bda_get_kernel_times")
        return out[:min(n.value, capacity)].astype(np.float64)

    def set_history(self, capacity: int):
        """Record ||r0|| and every half-iteration's residual norm (up to capacity values)."""
        self._check(lib.bda_set_history(self._h, capacity), "bda_set_history")

    def history(self, capacity: int = 1 << 16) -> np.ndarray:
        out = np.empty(max(capacity, 1))
        n = C.c_int32()
        self._check(lib.bda_get_history(self._h, out.ctypes.data_as(C.c_void_p), capacity, C.byref(n)),
                    "bda_get_history")
        return out[:min(n.value, capacity)].copy()

    def time_spmv(self, reps: int = 20, fp32: bool = False) -> float:
        ms = C.c_double()
        self._check(lib.bda_time_spmv(self._h, reps, int(fp32), C.byref(ms)), "bda_time_spmv")
        return ms.value

    def time_precond(self, reps: int = 20, fp32: bool = False) -> float:
        ms = C.c_double()
        self._check(lib.bda_time_precond(self._h, reps, int(fp32), C.byref(ms)), "bda_time_precond")
        return ms.value

    def measure_bandwidth(self, nbytes: int = 1 << 30, reps: int = 10):
        """(read GB/s, copy GB/s) of streaming kernels over nbytes."""
        r, c = C.c_double(), C.c_double()
        self._check(lib.bda_measure_bandwidth(self._h, nbytes, reps, C.byref(r), C.byref(c)), "bda_measure_bandwidth")
        return r.value, c.value

    def algorithmic_bytes(self, kind: int, fp32: bool = False) -> float:
        v = C.c_double()
        self._check(lib.bda_algorithmic_bytes(self._h, kind, int(fp32), C.byref(v)), "bda_algorithmic_bytes")
        return v.value


KIND_SPMV, KIND_LOWER, KIND_UPPER, KIND_VEC = 0, 1, 2, 3
