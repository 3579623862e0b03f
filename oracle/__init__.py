"""CPU oracle of the ILU0-BiCGSTAB path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
use this package, and only as the checker or the CPU baseline.  The solver
library never calls it.

Parity pinning (what each oracle function is checked against, in tests/):
  orc_colour / permutation  closed form (i+j+k) mod 2 on the 7-point grid,
                            brute-force validity and first-fit property,
                            committed SHA-256 of the 32x32x16 and 64x64x32
                            colourings (tests/golden/)
  spmv                      dense numpy A @ x on small systems and
                            scipy.sparse.bsr_matrix @ x fixtures (tests/golden/)
  factor (block ILU0)       (L U)_ij = A_ij on the pattern; red-black closed
                            form D_b = A_bb - sum A_br A_rr^-1 A_rb; dense LU
                            on block-tridiagonal chains (ILU0 = LU there)
  sweeps                    L U z = p with dense triangular solves
  bicgstab                  scipy.sparse.linalg.bicgstab with a right
                            preconditioner (same step structure, strict
                            tolerance compare) on tiny fixtures; dense solve
  mixed / adjoint           error vs the FP64 solution; dense A^-T w and
                            -lambda x^T gradients; central differences
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_i32, _f64, _vp = C.c_int32, C.c_double, C.c_void_p


class OrcResult(C.Structure):
    _fields_ = [
        ("iterations", _f64), ("half_iterations", _i32), ("last_norm_idx", _i32),
        ("norms", _f64 * 4), ("final_norm", _f64), ("rel_residual", _f64),
        ("converged", _i32), ("no_results", _i32), ("max_iter_reached", _i32),
        ("breakdown", _i32), ("nan_detected", _i32), ("aborted", _i32),
        ("outer_iterations", _i32), ("inner_half_iterations", _i32), ("solve_ms", _f64),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["norms"] = list(self.norms)
        return d


def build(verbose: bool = False) -> Path:
    """Compile liboracle.so with g++ (OpenMP, -ffp-contract=off)."""
    cmd = ["g++", "-O3", "-march=x86-64-v2", "-ffp-contract=off", "-fopenmp", "-std=c++17", "-fPIC", "-shared",
           "-Wall", str(HERE / "oracle.cpp"), "-o", str(LIB) + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stderr}")
    Path(str(LIB) + ".tmp").replace(LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


REF_DIR = HERE / "_ref"
REF_STATUS_LIB = REF_DIR / "libbda_status_ref.so"
REFERENCE_SRC = Path("/root/reference/linearalgebra/ilu0bicgstab/xilinx/src/sda_app/common")


def build_ref(verbose: bool = False):
    """Build the reference's status decoder into oracle/_ref (needs the
    reference sources, i.e. only where /root/reference exists)."""
    if not REFERENCE_SRC.exists():
        return None
    res = subprocess.run(["make", "-s", "-f", str(HERE / "ref.mk"), f"OUT={REF_DIR}", f"REF={REFERENCE_SRC}"],
                         capture_output=True, text=True, cwd=str(HERE.parent))
    if res.returncode != 0:
        raise RuntimeError(f"reference decoder build failed:\n{res.stderr}")
    if verbose:
        print(f"built {REF_STATUS_LIB}")
    return REF_STATUS_LIB


def ref_status_decoder():
    """ctypes handles of the reference's decode_debuginfo_bicgstab and
    decode_debuginfo_bicgstab_query (C++ linkage, called by mangled name)."""
    if not REF_STATUS_LIB.exists():
        return None
    L = C.CDLL(str(REF_STATUS_LIB))
    dec = getattr(L, "_Z25decode_debuginfo_bicgstabbbPmjjjPjS0_PdPhPbS3_S3_S3_S3_S3_")
    q = getattr(L, "_Z31decode_debuginfo_bicgstab_querybPmjjPjS0_S0_S0_PtS0_PbS2_S1_PhS3_S3_S3_S3_S3_S1_S1_")
    dec.restype = q.restype = _i32
    return dec, q


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.orc_colour.restype = _i32
        L.orc_colour.argtypes = [_i32, _vp, _vp, _i32, _vp]
        for name, args in {
            "orc_create": [_i32, _i32, _vp, _vp, _vp, _i32, C.POINTER(_vp)],
            "orc_destroy": [_vp], "orc_num_colours": [_vp, C.POINTER(_i32)],
            "orc_get_colouring": [_vp, _vp, _vp, _vp], "orc_update_values": [_vp, _vp],
            "orc_factor": [_vp], "orc_factor_sizes": [_vp, C.POINTER(_i32), C.POINTER(_i32)],
            "orc_get_factors": [_vp] * 8, "orc_spmv_perm": [_vp, _vp, _vp, _i32],
            "orc_precond_perm": [_vp, _vp, _vp, _i32],
            "orc_spmv_perm_ex": [_vp, _i32, _vp, _vp, _i32],
            "orc_precond_perm_ex": [_vp, _i32, _vp, _vp, _i32],
            "orc_solve": [_vp, _vp, _vp, _f64, _i32, _i32, C.POINTER(OrcResult)],
            "orc_solve_mixed": [_vp, _vp, _vp, _f64, _i32, _f64, _i32, _i32, _i32, _f64, C.POINTER(OrcResult)],
            "orc_adjoint_solve": [_vp, _vp, _vp, _f64, _i32, C.POINTER(OrcResult)],
            "orc_adjoint_solve_mixed": [_vp, _vp, _vp, _f64, _i32, _f64, _i32, _i32, _i32, _f64, C.POINTER(OrcResult)],
            "orc_grad": [_vp, _vp, _vp, _vp],
            "orc_get_history": [_vp, _vp, _i32, C.POINTER(_i32)],
        }.items():
            f = getattr(L, name)
            f.restype = _i32
            f.argtypes = args
        L.orc_set_threads.argtypes = [_i32]
        L.orc_set_threads.restype = None
        L.orc_get_threads.restype = _i32
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_vp)


def set_threads(n: int):
    lib().orc_set_threads(n)


def get_threads() -> int:
    return lib().orc_get_threads()


def colour(rowptr, colidx, level_scheduling=False) -> np.ndarray:
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colidx = np.ascontiguousarray(colidx, np.int32)
    n = rowptr.size - 1
    out = np.empty(n, np.int32)
    nc = lib().orc_colour(n, _p(rowptr), _p(colidx), int(level_scheduling), _p(out))
    if nc <= 0:
        raise RuntimeError("oracle colouring failed")
    return out


class OracleSystem:
    """CPU restatement of one solver handle (colour, factor, solve)."""

    def __init__(self, bs, rowptr, colidx, values, level_scheduling=False):
        self.rowptr = np.ascontiguousarray(rowptr, np.int32)
        self.colidx = np.ascontiguousarray(colidx, np.int32)
        vals = np.ascontiguousarray(values, np.float64)
        self.bs, self.n = bs, self.rowptr.size - 1
        self._h = _vp()
        rc = lib().orc_create(bs, self.n, _p(self.rowptr), _p(self.colidx), _p(vals), int(level_scheduling),
                              C.byref(self._h))
        if rc != 0:
            raise RuntimeError("orc_create failed")

    def close(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass    # at interpreter exit the module's globals may already be gone

    @property
    def N(self):
        return self.n * self.bs

    def colouring(self):
        nc = _i32()
        lib().orc_num_colours(self._h, C.byref(nc))
        c, p, cp = np.empty(self.n, np.int32), np.empty(self.n, np.int32), np.empty(nc.value + 1, np.int32)
        lib().orc_get_colouring(self._h, _p(c), _p(p), _p(cp))
        return c, p, cp

    def update_values(self, values):
        v = np.ascontiguousarray(values, np.float64)
        lib().orc_update_values(self._h, _p(v))

    def factor(self):
        if lib().orc_factor(self._h) != 0:
            raise RuntimeError("oracle factorisation failed (singular or missing diagonal block)")

    def factors(self):
        nl, nu = _i32(), _i32()
        lib().orc_factor_sizes(self._h, C.byref(nl), C.byref(nu))
        bs = self.bs
        lr, ur = np.empty(self.n + 1, np.int32), np.empty(self.n + 1, np.int32)
        lc, uc = np.empty(nl.value, np.int32), np.empty(nu.value, np.int32)
        lv, uv = np.empty((nl.value, bs, bs)), np.empty((nu.value, bs, bs))
        dinv = np.empty((self.n, bs, bs))
        lib().orc_get_factors(self._h, _p(lr), _p(lc), _p(lv), _p(ur), _p(uc), _p(uv), _p(dinv))
        return (lr, lc, lv), (ur, uc, uv), dinv

    def spmv_perm(self, x, transpose=False):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.N)
        if lib().orc_spmv_perm(self._h, _p(x), _p(y), int(transpose)) != 0:
            raise RuntimeError("orc_spmv_perm failed")
        return y

    def precond_perm(self, p, transpose=False):
        p = np.ascontiguousarray(p, np.float64)
        z = np.empty(self.N)
        if lib().orc_precond_perm(self._h, _p(p), _p(z), int(transpose)) != 0:
            raise RuntimeError("orc_precond_perm failed")
        return z

    _PREC_DTYPE = {0: np.float64, 1: np.float32, 2: np.float64}

    def spmv_perm_ex(self, x, prec=0, transpose=False):
        """y = A x in kernel form prec (0 FP64; 1 FP32 values and vectors;
        2 FP32 values, FP64 vectors)."""
        dt = self._PREC_DTYPE[prec]
        x = np.ascontiguousarray(x, dt)
        y = np.empty(self.N, dt)
        if lib().orc_spmv_perm_ex(self._h, prec, _p(x), _p(y), int(transpose)) != 0:
            raise RuntimeError("orc_spmv_perm_ex failed")
        return y

    def precond_perm_ex(self, p, prec=0, transpose=False):
        dt = self._PREC_DTYPE[prec]
        p = np.ascontiguousarray(p, dt)
        z = np.empty(self.N, dt)
        if lib().orc_precond_perm_ex(self._h, prec, _p(p), _p(z), int(transpose)) != 0:
            raise RuntimeError("orc_precond_perm_ex failed")
        return z

    def solve(self, b, x0=None, tol=1e-6, maxit=1000, abs_tol=False):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.N) if x0 is None else np.array(x0, np.float64)
        r = OrcResult()
        if lib().orc_solve(self._h, _p(b), _p(x), tol, maxit, int(abs_tol), C.byref(r)) != 0:
            raise RuntimeError("orc_solve failed")
        return x, r.as_dict()

    def solve_mixed(self, b, x0=None, tol=1e-6, max_outer=20, inner_tol=1e-4, inner_maxit=1000, vec32=True,
                    adaptive=True, eta=0.9):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.N) if x0 is None else np.array(x0, np.float64)
        r = OrcResult()
        if lib().orc_solve_mixed(self._h, _p(b), _p(x), tol, max_outer, inner_tol, inner_maxit, int(vec32),
                                 int(adaptive), eta, C.byref(r)) != 0:
            raise RuntimeError("orc_solve_mixed failed")
        return x, r.as_dict()

    def history(self, capacity=1 << 16):
        out = np.empty(capacity)
        n = _i32()
        lib().orc_get_history(self._h, _p(out), capacity, C.byref(n))
        return out[:min(n.value, capacity)].copy()

    def adjoint_solve(self, w, lam0=None, tol=1e-10, maxit=1000):
        w = np.ascontiguousarray(w, np.float64)
        lam = np.zeros(self.N) if lam0 is None else np.array(lam0, np.float64)
        r = OrcResult()
        if lib().orc_adjoint_solve(self._h, _p(w), _p(lam), tol, maxit, C.byref(r)) != 0:
            raise RuntimeError("orc_adjoint_solve failed")
        return lam, r.as_dict()

    def adjoint_solve_mixed(self, w, lam0=None, tol=1e-10, max_outer=20, inner_tol=1e-4, inner_maxit=1000,
                            vec32=True, adaptive=True, eta=0.9):
        w = np.ascontiguousarray(w, np.float64)
        lam = np.zeros(self.N) if lam0 is None else np.array(lam0, np.float64)
        r = OrcResult()
        if lib().orc_adjoint_solve_mixed(self._h, _p(w), _p(lam), tol, max_outer, inner_tol, inner_maxit, int(vec32),
                                         int(adaptive), eta, C.byref(r)) != 0:
            raise RuntimeError("orc_adjoint_solve_mixed failed")
        return lam, r.as_dict()

    def grad(self, lam, x):
        lam = np.ascontiguousarray(lam, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        g = np.empty((int(self.rowptr[-1]), self.bs, self.bs))
        lib().orc_grad(self._h, _p(lam), _p(x), _p(g))
        return g
