"""Parity at the full 256x256x64 size with 4x4 blocks (3.77M block rows,
24.1M blocks): colouring, SpMV, factors and preconditioner bit for bit
against the oracle on the host cores, the FP64 solve to 1e-6 within one
iteration and 1e-8 relative in x, and the true residual on the host."""
import hashlib

import numpy as np
import pytest

import fpga_amd as fa
import oracle
from fpga_amd.configs import GRID_256x256x64

pytestmark = pytest.mark.gpu


def h(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def big4():
    print("case grid256x256x64 bs4", flush=True)
    s = fa.generate(GRID_256x256x64)
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    g.upload_system(s)
    g.analyse()
    g.factor()
    o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, s.values)
    o.factor()
    yield s, g, o
    g.close()


def test_full_bs4_components_bit_exact(big4):
    s, g, o = big4
    assert s.bs == 4
    # one launch per colour and sweep: the last colour (84 rows) needs two
    # tiles, so neither tail launch applies, and colour 2's launch (6,260
    # tiles) streams at full rate, so colours 2 and 3 keep separate launches
    assert g.precond_launches() == 8
    for a, b in zip(g.colouring(), o.colouring()):
        np.testing.assert_array_equal(a, b)
    x = np.random.default_rng(11).uniform(-1, 1, s.N)
    assert h(g.spmv_perm(x)) == h(o.spmv_perm(x))
    assert h(g.precond_apply_perm(x)) == h(o.precond_perm(x))
    (gl, gu, gd), (ol, ou, od) = [(f[0][2], f[1][2], f[2]) for f in (g.factors(), o.factors())]
    assert h(gl) == h(ol) and h(gu) == h(ou) and h(gd) == h(od)


def test_full_bs4_solve_parity(big4):
    s, g, o = big4
    xg, rg = g.solve(s.b, tol=1e-6, maxit=1000)
    xo, ro = o.solve(s.b, tol=1e-6, maxit=1000)
    print("grid256x256x64 bs4 half iterations gpu", rg.half_iterations, "oracle", ro["half_iterations"])
    assert rg.converged and ro["converged"]
    assert abs(rg.half_iterations - ro["half_iterations"]) <= 2
    if rg.half_iterations == ro["half_iterations"]:
        assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-8
    import scipy.sparse as sp
    A = sp.bsr_matrix((s.values, s.colidx, s.rowptr), shape=(s.N, s.N))
    assert np.linalg.norm(s.b - A @ xg) / np.linalg.norm(s.b) < 1.1e-6


def test_full_bs4_precision_form_and_adjoint(big4):
    """The 256x256x64 4x4 workload end to end: the precision form to 1e-6
    (FP32 values and inner vectors, FP64 refinement) with an x error no worse
    than the FP64 solve's at the same tolerance, both measured against a
    1e-10 FP64 solve; the adjoint A^T lambda = w to 1e-10 within 1e-8 of the
    oracle's and within one iteration of its count; the gradient -lambda x^T
    on the pattern bit for bit."""
    s, g, o = big4
    x_ref, r_ref = g.solve(s.b, tol=1e-10, maxit=2000)
    assert r_ref.converged
    xm, rm = g.solve_mixed(s.b, tol=1e-6, inner_tol=1e-4, vec32=True)
    xf, rf = g.solve(s.b, tol=1e-6, maxit=1000)
    nref = np.linalg.norm(x_ref)
    err_m, err_f = np.linalg.norm(xm - x_ref) / nref, np.linalg.norm(xf - x_ref) / nref
    print(f"grid256x256x64 bs4: mixed {rm.outer_iterations} outer / {rm.iterations} inner it, x err {err_m:.3e}; "
          f"fp64 {rf.iterations} it, x err {err_f:.3e}", flush=True)
    assert rm.converged and rf.converged
    assert err_m <= err_f
    w = np.random.default_rng(12).standard_normal(s.N)
    lg, rg = g.adjoint_solve(w, tol=1e-10, maxit=1000)
    lo, ro = o.adjoint_solve(w, tol=1e-10, maxit=1000)
    print(f"adjoint half iterations gpu {rg.half_iterations} oracle {ro['half_iterations']}", flush=True)
    assert rg.converged and ro["converged"]
    assert abs(rg.half_iterations - ro["half_iterations"]) <= 2
    assert np.linalg.norm(lg - lo) / np.linalg.norm(lo) < 1e-8
    assert h(g.grad_values(lo, x_ref)) == h(o.grad(lo, x_ref))
