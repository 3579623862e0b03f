"""Level scheduling at the full 128x128x64 size (943,550 block rows, 421
level sets of the lower-triangular graph instead of 5 colours).

The host API's level_scheduling flag (fpga_functions_bicgstab.cpp:84)
orders the rows by level sets; every sweep then takes one launch per level.
Against the oracle on the same ordering: the level sets, SpMV and
preconditioner in FP64 and FP32 bit for bit (SHA-256), the ILU0 factors bit
for bit, the FP64 solve to 1e-6 within one iteration, and the precision
form's error against a 1e-10 solve no worse than the FP64 form's."""
import hashlib

import numpy as np
import pytest

import fpga_amd as fa
import oracle
from fpga_amd.configs import GRID_128x128x64

pytestmark = pytest.mark.gpu


def h(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def levelled():
    print("case grid128x128x64, level scheduling", flush=True)
    s = fa.generate(GRID_128x128x64)
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    g.upload_system(s)
    g.analyse(level_scheduling=True)
    g.factor()
    o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, s.values, True)
    o.factor()
    yield s, g, o
    g.close()


def test_level_components_bit_exact(levelled):
    s, g, o = levelled
    for a, b in zip(g.colouring(), o.colouring()):
        np.testing.assert_array_equal(a, b)
    nlev = len(g.colouring()[2]) - 1
    assert nlev == 421
    # one launch per level and sweep, less the merged launches (the first
    # level's lower sweep is a copy, two lower levels share a launch, the
    # last level and two upper levels share the tail launch)
    assert g.precond_launches() == 2 * nlev - 5
    x = np.random.default_rng(17).uniform(-1, 1, s.N)
    assert h(g.spmv_perm(x)) == h(o.spmv_perm(x))
    assert h(g.precond_apply_perm(x)) == h(o.precond_perm(x))
    x32 = x.astype(np.float32)
    assert h(g.spmv_perm_ex(x32, 1)) == h(o.spmv_perm_ex(x32, 1))
    assert h(g.precond_apply_perm_ex(x32, 1)) == h(o.precond_perm_ex(x32, 1))
    (gl, gu, gd), (ol, ou, od) = [(f[0][2], f[1][2], f[2]) for f in (g.factors(), o.factors())]
    assert h(gl) == h(ol) and h(gu) == h(ou) and h(gd) == h(od)


def test_level_solves(levelled):
    s, g, o = levelled
    xg, rg = g.solve(s.b, tol=1e-6, maxit=1000)
    xo, ro = o.solve(s.b, tol=1e-6, maxit=1000)
    print("level scheduling half iterations gpu", rg.half_iterations, "oracle", ro["half_iterations"], flush=True)
    assert rg.converged and ro["converged"]
    assert abs(rg.half_iterations - ro["half_iterations"]) <= 2
    if rg.half_iterations == ro["half_iterations"]:
        assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-8
    x64, _ = g.solve(s.b, tol=1e-10, maxit=2000)
    xm, rm = g.solve_mixed(s.b, tol=1e-6, inner_tol=1e-4, vec32=True)
    err_m = np.linalg.norm(xm - x64) / np.linalg.norm(x64)
    err_f = np.linalg.norm(xg - x64) / np.linalg.norm(x64)
    print(f"level scheduling x error vs 1e-10 solve: mixed {err_m:.3e}, fp64 {err_f:.3e}", flush=True)
    assert rm.converged and err_m < 1e-5
    assert err_m <= err_f
