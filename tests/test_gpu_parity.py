"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Component kernels (SpMV, block ILU0, L/U sweeps, their transposes) use the
same per-row summation order as the oracle and must match it bit for bit.
Solves: iteration counts within one full iteration, x within 1e-8 relative
at equal iteration counts (FP64).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import fpga_amd as fa
import oracle
from fpga_amd.configs import GRID_32x32x16, GRID_128x128x64, GRID_256x256x64, scaled

pytestmark = pytest.mark.gpu

CASES = {
    "grid32x32x16": GRID_32x32x16,
    "grid20x18x10_nnc": scaled(GRID_128x128x64, 20, 18, 10),
    "grid12x10x6_bs4_nnc": fa.GridSpec(12, 10, 6, bs=4, seed=3, perm_sigma=2.0, layer_corr=0.5,
                                       inactive_frac=0.1, nnc_frac=0.05, nnc_dist=3),
}


def _pair(spec, level=False):
    print(f"case {spec}", flush=True)
    s = fa.generate(spec)
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    g.upload_system(s)
    g.analyse(level)
    g.factor()
    o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, s.values, level)
    o.factor()
    return s, g, o


@pytest.fixture(scope="module", params=list(CASES))
def system(request):
    s, g, o = _pair(CASES[request.param])
    yield request.param, s, g, o
    g.close()


def test_colouring_bit_exact(system):
    _, s, g, o = system
    for a, b in zip(g.colouring(), o.colouring()):
        np.testing.assert_array_equal(a, b)


def test_spmv_bit_exact(system):
    _, s, g, o = system
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, s.N)
    np.testing.assert_array_equal(g.spmv_perm(x), o.spmv_perm(x))
    np.testing.assert_array_equal(g.spmv_perm(x, transpose=True), o.spmv_perm(x, transpose=True))


def test_factors_bit_exact(system):
    _, s, g, o = system
    (glr, glc, glv), (gur, guc, guv), gd = g.factors()
    (olr, olc, olv), (our, ouc, ouv), od = o.factors()
    np.testing.assert_array_equal(glr, olr)
    np.testing.assert_array_equal(glc, olc)
    np.testing.assert_array_equal(gur, our)
    np.testing.assert_array_equal(guc, ouc)
    np.testing.assert_array_equal(glv, olv)
    np.testing.assert_array_equal(guv, ouv)
    np.testing.assert_array_equal(gd, od)


def test_precond_bit_exact(system):
    _, s, g, o = system
    rng = np.random.default_rng(2)
    p = rng.uniform(-1, 1, s.N)
    np.testing.assert_array_equal(g.precond_apply_perm(p), o.precond_perm(p))
    np.testing.assert_array_equal(g.precond_apply_perm(p, transpose=True), o.precond_perm(p, transpose=True))


@pytest.mark.parametrize("maxit", [0, 1, 2, 5])
def test_fixed_iterations(system, maxit):
    """tol = 0 never converges: both stop at the same half step; x agrees."""
    _, s, g, o = system
    xg, rg = g.solve(s.b, tol=0.0, maxit=maxit)
    xo, ro = o.solve(s.b, tol=0.0, maxit=maxit)
    assert rg.max_iter_reached and ro["max_iter_reached"]
    assert rg.half_iterations == ro["half_iterations"] == 2 * maxit
    np.testing.assert_allclose(xg, xo, rtol=0, atol=1e-10 * np.abs(xo).max())
    np.testing.assert_allclose(rg.norms[0], ro["norms"][0], rtol=1e-13)


def test_solve_parity(system):
    name, s, g, o = system
    tol = 1e-2 if name == "grid32x32x16" else 1e-6
    xg, rg = g.solve(s.b, tol=tol, maxit=200)
    xo, ro = o.solve(s.b, tol=tol, maxit=200)
    print(name, rg.half_iterations, ro["half_iterations"], rg.rel_residual, ro["rel_residual"])
    assert rg.converged and ro["converged"]
    assert abs(rg.half_iterations - ro["half_iterations"]) <= 2
    if rg.half_iterations == ro["half_iterations"]:
        err = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
        assert err < 1e-8, err
    assert rg.last_norm_idx == ro["last_norm_idx"] or rg.half_iterations != ro["half_iterations"]


def test_residual_vector(system):
    _, s, g, o = system
    xg, rg = g.solve(s.b, tol=1e-6, maxit=200)
    r = g.residual()
    assert np.isclose(np.linalg.norm(r), rg.final_norm, rtol=1e-12)


def test_mixed_precision(system):
    name, s, g, o = system
    x64, _ = o.solve(s.b, tol=1e-12, maxit=1000)
    for vec32 in (False, True):
        xm, rm = g.solve_mixed(s.b, tol=1e-6, inner_tol=1e-4, vec32=vec32)
        xom, rom = o.solve_mixed(s.b, tol=1e-6, inner_tol=1e-4, vec32=vec32)
        assert rm.converged and rom["converged"]
        err = np.linalg.norm(xm - x64) / np.linalg.norm(x64)
        print(name, "vec32" if vec32 else "vec64", rm.outer_iterations, rm.inner_half_iterations, err)
        assert err < 1e-5
        assert abs(rm.outer_iterations - rom["outer_iterations"]) <= 1


def test_adjoint_parity(system):
    name, s, g, o = system
    rng = np.random.default_rng(5)
    w = rng.standard_normal(s.N)
    lg, rg = g.adjoint_solve(w, tol=1e-10, maxit=500)
    lo, ro = o.adjoint_solve(w, tol=1e-10, maxit=500)
    assert rg.converged and ro["converged"]
    assert abs(rg.half_iterations - ro["half_iterations"]) <= 2
    err = np.linalg.norm(lg - lo) / np.linalg.norm(lo)
    assert err < 1e-8, err
    x, _ = o.solve(s.b, tol=1e-12, maxit=500)
    np.testing.assert_array_equal(g.grad_values(lo, x), o.grad(lo, x))


@pytest.mark.parametrize("bs", [3, 4])
def test_level_scheduling_solve(bs):
    base = GRID_128x128x64 if bs == 3 else GRID_256x256x64
    s, g, o = _pair(scaled(base, 12, 10, 8), level=True)
    assert s.bs == bs
    for a, b in zip(g.colouring(), o.colouring()):
        np.testing.assert_array_equal(a, b)
    xg, rg = g.solve(s.b, tol=1e-8, maxit=200)
    xo, ro = o.solve(s.b, tol=1e-8, maxit=200)
    assert rg.converged and ro["converged"] and abs(rg.half_iterations - ro["half_iterations"]) <= 2
    if rg.half_iterations == ro["half_iterations"]:
        assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-8
    # the level-set sweeps bit for bit, as for the colouring, in every kernel
    # form and transposed
    p = np.random.default_rng(9).uniform(-1, 1, s.N)
    np.testing.assert_array_equal(g.precond_apply_perm(p), o.precond_perm(p))
    np.testing.assert_array_equal(g.spmv_perm(p), o.spmv_perm(p))
    for prec, dt in ((1, np.float32), (2, np.float64)):
        for tr in (False, True):
            np.testing.assert_array_equal(g.precond_apply_perm_ex(p.astype(dt), prec, transpose=tr),
                                          o.precond_perm_ex(p.astype(dt), prec, transpose=tr))
    # the precision form and both adjoints on the level-set ordering
    xm, rm = g.solve_mixed(s.b, tol=1e-8)
    assert rm.converged and np.linalg.norm(xm - xo) / np.linalg.norm(xo) < 1e-6
    w = np.random.default_rng(10).standard_normal(s.N)
    lo, ro = o.adjoint_solve(w, tol=1e-10, maxit=500)
    lg, rg = g.adjoint_solve(w, tol=1e-10, maxit=500)
    lm, rma = g.adjoint_solve_mixed(w, tol=1e-10)
    assert rg.converged and rma.converged
    assert np.linalg.norm(lg - lo) / np.linalg.norm(lo) < 1e-8
    assert np.linalg.norm(lm - lo) / np.linalg.norm(lo) < 1e-8
    g.close()


def test_mixed_inner_breakdown_ends_refinement():
    """A right-hand side finite in FP64 but beyond the FP32 range: the first
    inner run meets Inf/NaN before its first half-iteration and changes
    nothing, so the refinement stops after that step (it would otherwise
    repeat the same inner run max_outer times), on the GPU as in the oracle."""
    s, g, o = _pair(scaled(GRID_32x32x16, 6, 5, 4))
    b = s.b * 1e39
    xg, rg = g.solve_mixed(b, tol=1e-6, max_outer=20, vec32=True)
    xo, ro = o.solve_mixed(b, tol=1e-6, max_outer=20, vec32=True)
    assert rg.nan_detected and ro["nan_detected"] and not rg.converged and not ro["converged"]
    assert rg.outer_iterations == ro["outer_iterations"] == 1
    assert rg.inner_half_iterations == ro["inner_half_iterations"] == 0
    assert np.all(xg == 0) and np.all(xo == 0)
    g.close()


def test_edge_zero_rhs():
    s, g, o = _pair(scaled(GRID_32x32x16, 6, 5, 4))
    x, r = g.solve(np.zeros(s.N), tol=1e-6, maxit=10)
    assert r.no_results and not r.converged and r.half_iterations == 0
    assert np.all(x == 0)
    x0 = np.ones(s.N)
    x, r = g.solve(np.zeros(s.N), x0=x0, tol=1e-6, maxit=10)
    assert not r.no_results  # b - A x0 != 0
    g.close()


def test_edge_abs_tolerance():
    s, g, o = _pair(scaled(GRID_32x32x16, 6, 5, 4))
    g.set_options(abs_tol=True)
    x, r = g.solve(s.b, tol=1e-3, maxit=100)
    assert r.converged and r.final_norm <= 1e-3
    big = g.solve(s.b, tol=1e9, maxit=100)[1]
    assert big.no_results
    g.close()


def test_edge_nan_values():
    s = fa.generate(scaled(GRID_32x32x16, 6, 5, 4))
    vals = s.values.copy()
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    g.upload_bcsr(s.rowptr, s.colidx, vals)
    g.analyse()
    g.factor()
    b = s.b.copy()
    b[7] = np.nan
    x, r = g.solve(b, tol=1e-6, maxit=10)
    assert r.nan_detected and not r.converged
    g.close()


def test_update_values_reuses_handle():
    spec = scaled(GRID_32x32x16, 8, 8, 4)
    s = fa.generate(spec)
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    g.upload_system(s)
    g.analyse()
    g.factor()
    x1, _ = g.solve(s.b, tol=1e-10, maxit=200)
    v2 = s.values * 1.5
    g.update_values(v2)
    g.factor()
    x2, r2 = g.solve(s.b, tol=1e-10, maxit=200)
    o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, v2)
    o.factor()
    xo, _ = o.solve(s.b, tol=1e-10, maxit=200)
    assert np.linalg.norm(x2 - xo) / np.linalg.norm(xo) < 1e-8
    assert np.linalg.norm(x2 * 1.5 - x1) / np.linalg.norm(x1) < 1e-8
    g.close()


def test_mixed_edge_cases():
    """Stop rules of the device-side refinement loop against the oracle's."""
    s, g, o = _pair(scaled(GRID_32x32x16, 8, 7, 5))
    # zero right-hand side: nothing to do
    x, r = g.solve_mixed(np.zeros(s.N), tol=1e-6)
    xo, ro = o.solve_mixed(np.zeros(s.N), tol=1e-6)
    assert r.no_results and ro["no_results"] and not r.converged
    assert r.outer_iterations == 0 and np.all(x == 0)
    # no outer step allowed: the residual of x0 only
    x, r = g.solve_mixed(s.b, tol=1e-6, max_outer=0)
    xo, ro = o.solve_mixed(s.b, tol=1e-6, max_outer=0)
    assert r.max_iter_reached and ro["max_iter_reached"] and r.outer_iterations == 0
    assert np.all(x == 0) and np.isclose(r.norms[0], ro["norms"][0], rtol=1e-12)
    # one outer step towards an unreachable target
    x, r = g.solve_mixed(s.b, tol=1e-14, max_outer=1)
    xo, ro = o.solve_mixed(s.b, tol=1e-14, max_outer=1)
    assert r.max_iter_reached and ro["max_iter_reached"] and not r.converged
    assert r.outer_iterations == ro["outer_iterations"] == 1 and r.last_norm_idx == ro["last_norm_idx"] == 1
    assert np.isclose(r.final_norm, ro["final_norm"], rtol=1e-3)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-3
    # runs of both kinds back to back on one handle keep their own results
    for _ in range(2):
        x64, r64 = g.solve(s.b, tol=1e-8, maxit=300)
        xm, rm = g.solve_mixed(s.b, tol=1e-8)
        assert r64.converged and rm.converged and rm.outer_iterations >= 1
        assert np.linalg.norm(xm - x64) / np.linalg.norm(x64) < 1e-6
    g.close()


def test_norm_history(system):
    """Device-recorded norm history: h + 2 values, ||r0|| first, the final
    norm last, the norm ring in its last three values; against the oracle's
    history value by value (dots are summed in different orders: 1e-6)."""
    name, s, g, o = system
    g.set_history(4096)
    for tol, maxit in [(1e-8, 300), (0.0, 4)]:
        xg, rg = g.solve(s.b, tol=tol, maxit=maxit)
        hg = g.history()
        assert len(hg) == rg.half_iterations + 2
        assert hg[0] == rg.norms[0] and hg[-1] == rg.final_norm
        ring = [rg.norms[(rg.last_norm_idx - 1 - k) % 3 + 1] for k in range(3)]
        assert list(hg[::-1][:3]) == ring
        xo, ro = o.solve(s.b, tol=tol, maxit=maxit)
        ho = o.history()
        k = min(len(hg), len(ho), 12)
        np.testing.assert_allclose(hg[:k], ho[:k], rtol=1e-6)
    g.set_history(0)


def test_edge_watchdog_abort():
    """The host watchdog (the counterpart of param0's abort cycles): a tiny
    timeout aborts the solve, the result says so, and the handle stays
    usable for the next solve."""
    s, g, o = _pair(scaled(GRID_32x32x16, 12, 10, 8))
    g.set_options(timeout_s=1e-9)
    x, r = g.solve(s.b, tol=1e-14, maxit=1000)
    assert r.aborted and not r.converged
    # the deadline is checked on the device: the run stops at its first
    # scalar step, and what the host had queued does no work
    assert r.half_iterations == 0 and np.all(x == 0)
    xm, rm = g.solve_mixed(s.b, tol=1e-14)
    assert rm.aborted and not rm.converged
    assert rm.outer_iterations == 0 and np.all(xm == 0)
    g.set_options(timeout_s=0.0)
    x, r = g.solve(s.b, tol=1e-8, maxit=300)
    xo, ro = o.solve(s.b, tol=1e-8, maxit=300)
    assert r.converged and not r.aborted
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8
    g.close()


def test_adjoint_mixed_parity(system):
    """The adjoint in the precision form: lambda against the oracle's FP64
    adjoint (both to 1e-10) and the refinement step count against the
    oracle's precision-form adjoint; the gradient it gives against the FP64
    one."""
    name, s, g, o = system
    rng = np.random.default_rng(5)
    w = rng.standard_normal(s.N)
    lo, ro = o.adjoint_solve(w, tol=1e-10, maxit=500)
    x, _ = o.solve(s.b, tol=1e-12, maxit=500)
    for vec32 in (True, False):
        lg, rg = g.adjoint_solve_mixed(w, tol=1e-10, vec32=vec32)
        lom, rom = o.adjoint_solve_mixed(w, tol=1e-10, vec32=vec32)
        print(name, "vec32" if vec32 else "vec64", rg.outer_iterations, rg.inner_half_iterations,
              rom["outer_iterations"], rom["inner_half_iterations"], flush=True)
        assert rg.converged and rom["converged"]
        assert abs(rg.outer_iterations - rom["outer_iterations"]) <= 1
        err = np.linalg.norm(lg - lo) / np.linalg.norm(lo)
        assert err < 1e-8, err
        gg, go = g.grad_values(lg, x), o.grad(lo, x)
        assert np.linalg.norm(gg - go) <= 1e-8 * np.linalg.norm(go)


def test_adjoint_mixed_edge_cases():
    """Stop rules of the precision-form adjoint against the oracle's, and
    forward and adjoint runs of both forms back to back on one handle."""
    s, g, o = _pair(scaled(GRID_32x32x16, 8, 7, 5))
    w = np.random.default_rng(9).standard_normal(s.N)
    lam, r = g.adjoint_solve_mixed(np.zeros(s.N), tol=1e-10)
    assert r.no_results and not r.converged and r.outer_iterations == 0 and np.all(lam == 0)
    lam, r = g.adjoint_solve_mixed(w, tol=1e-10, max_outer=0)
    lo, ro = o.adjoint_solve_mixed(w, tol=1e-10, max_outer=0)
    assert r.max_iter_reached and ro["max_iter_reached"] and r.outer_iterations == 0
    assert np.all(lam == 0) and np.isclose(r.norms[0], ro["norms"][0], rtol=1e-12)
    lam, r = g.adjoint_solve_mixed(w, tol=1e-14, max_outer=1)
    lo, ro = o.adjoint_solve_mixed(w, tol=1e-14, max_outer=1)
    assert r.max_iter_reached and ro["max_iter_reached"] and r.outer_iterations == ro["outer_iterations"] == 1
    assert np.linalg.norm(lam - lo) / np.linalg.norm(lo) < 1e-3
    l64, _ = o.adjoint_solve(w, tol=1e-12, maxit=500)
    for _ in range(2):
        xm, rm = g.solve_mixed(s.b, tol=1e-8)
        lm, ra = g.adjoint_solve_mixed(w, tol=1e-10)
        lf, rf = g.adjoint_solve(w, tol=1e-10, maxit=500)
        assert rm.converged and ra.converged and rf.converged
        assert np.linalg.norm(lm - l64) / np.linalg.norm(l64) < 1e-8
        assert np.linalg.norm(lf - l64) / np.linalg.norm(l64) < 1e-8
        assert np.linalg.norm(g.residual() - (w - sp.bsr_matrix((s.values, s.colidx, s.rowptr),
                                                                  shape=(s.N, s.N)).T @ lf)) < 1e-6 * np.linalg.norm(w)
    g.close()
