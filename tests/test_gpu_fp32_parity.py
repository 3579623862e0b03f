"""GPU parity of the precision form's kernels, component by component.

The timed solve of the precision form runs FP32 instances of the streaming
SpMV (both reduction twins), of the L and U sweeps (narrow, tall and merged
tail launches) and FP32 vector updates.  Here each of them is run on its own
through bda_spmv_perm_ex / bda_precond_apply_perm_ex and compared bit for bit
with the oracle's FP32 path (FP32 values, FP32 products, FP64 row sums,
D^-1 in FP64, outputs rounded to the vector type), in both vector forms:
prec 1 = float vectors (vec32), prec 2 = double vectors.

Row semantics anchored on the reference kernel: spmvp.vhd:240-383 (SpMV
rows), ilu0.vhd:385-416 (L: y = p - L y; U: z = D^-1 (y - U z)).

The fused dot products are checked against exact sums (math.fsum of the
products): an FP64 sum of n terms in the solve's fixed order is within a few
hundred units of 2^-53 of the exact value relative to sum |y_i w_i|; 1e-12 is
far above that and far below the 1e-8..1e-7 an FP32 accumulation leaves.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import fpga_amd as fa
import oracle
from fpga_amd.configs import GRID_32x32x16, GRID_128x128x64, scaled

pytestmark = pytest.mark.gpu

CASES = {
    "grid32x32x16": GRID_32x32x16,
    "grid20x18x10_nnc": scaled(GRID_128x128x64, 20, 18, 10),
    "grid12x10x6_bs4_nnc": fa.GridSpec(12, 10, 6, bs=4, seed=3, perm_sigma=2.0, layer_corr=0.5,
                                       inactive_frac=0.1, nnc_frac=0.05, nnc_dist=3),
}
DT = {0: np.float64, 1: np.float32, 2: np.float64}


@pytest.fixture(scope="module", params=list(CASES))
def system(request):
    spec = CASES[request.param]
    print(f"case {request.param} {spec}", flush=True)
    s = fa.generate(spec)
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    g.upload_system(s)
    g.analyse()
    g.factor()
    o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, s.values)
    o.factor()
    yield request.param, s, g, o
    g.close()


def exact_dot(a, b):
    return math.fsum((np.asarray(a, np.float64) * np.asarray(b, np.float64)).tolist())


def check_dot(got, a, b, tol=1e-12):
    ex = exact_dot(a, b)
    scale = float(np.sum(np.abs(np.asarray(a, np.float64) * np.asarray(b, np.float64))))
    assert abs(got - ex) <= tol * scale, (got, ex, scale)


@pytest.mark.parametrize("prec", [1, 2])
def test_spmv_fp32_bit_exact(system, prec):
    _, s, g, o = system
    x = np.random.default_rng(21).uniform(-1, 1, s.N).astype(DT[prec])
    for tr in (False, True):
        np.testing.assert_array_equal(g.spmv_perm_ex(x, prec, transpose=tr), o.spmv_perm_ex(x, prec, transpose=tr))


@pytest.mark.parametrize("prec", [1, 2])
def test_precond_fp32_bit_exact(system, prec):
    _, s, g, o = system
    p = np.random.default_rng(22).uniform(-1, 1, s.N).astype(DT[prec])
    for tr in (False, True):
        np.testing.assert_array_equal(g.precond_apply_perm_ex(p, prec, transpose=tr),
                                      o.precond_perm_ex(p, prec, transpose=tr))


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_spmv_reduction_twins(system, prec):
    """The two SpMV instances of a solve with their fused sums: v = A phat
    with v.rt (red 1) and t = A shat with t.s and t.t (red 3)."""
    _, s, g, o = system
    rng = np.random.default_rng(23)
    x = rng.uniform(-1, 1, s.N).astype(DT[prec])
    w = rng.uniform(-1, 1, s.N).astype(DT[prec])
    yref = o.spmv_perm_ex(x, prec)
    y1, d1 = g.spmv_perm_ex(x, prec, w=w, red=1)
    y3, d3 = g.spmv_perm_ex(x, prec, w=w, red=3)
    np.testing.assert_array_equal(y1, yref)
    np.testing.assert_array_equal(y3, yref)
    check_dot(d1[0], yref, w)
    check_dot(d3[0], yref, w)
    check_dot(d3[1], yref, yref)
    assert d1[0] == d3[0]  # one fixed summation order for both instances


def test_mixed_fixed_iterations(system):
    """One refinement step with a fixed number of inner iterations (inner
    tolerance 0) against the oracle's precision form.  GPU and oracle differ
    only in the order of the dot products' sums (a few units of 2^-53 in
    alpha, omega, beta).  After one inner iteration that leaves x equal to
    1e-12 -- the discriminating check: an FP32 accumulation in any kernel
    moves every entry by 1e-8..1e-7.  Later iterations amplify the last-bit
    differences through BiCGSTAB's recurrences, and with FP32 inner vectors
    a difference also flips an entry's FP32 rounding now and then (measured
    on grid20x18x10_nnc: 0, 2.4e-10, 9.6e-9 after 1, 3, 6 iterations with FP32
    vectors; 7e-16, 3e-14, 1.6e-12 with FP64 vectors), so those are bounded
    by 1e-8 and 1e-7."""
    _, s, g, o = system
    A = sp.bsr_matrix((s.values, s.colidx, s.rowptr), shape=(s.N, s.N))
    bounds = {1: 1e-12, 3: 1e-8, 6: 1e-7}
    for vec32 in (False, True):
        for k, bound in bounds.items():
            xg, rg = g.solve_mixed(s.b, tol=0.0, max_outer=1, inner_tol=0.0, inner_maxit=k, vec32=vec32)
            xo, ro = o.solve_mixed(s.b, tol=0.0, max_outer=1, inner_tol=0.0, inner_maxit=k, vec32=vec32)
            assert rg.max_iter_reached and ro["max_iter_reached"]
            assert rg.outer_iterations == ro["outer_iterations"] == 1
            assert rg.inner_half_iterations == ro["inner_half_iterations"] == 2 * k
            err = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
            print(f"vec32={vec32} inner_maxit={k}: x rel diff {err:.3e}", flush=True)
            assert err <= bound, err
            # both final norms are ||b - A x|| in FP64, so they differ by at
            # most ||A (xg - xo)|| plus the rounding of the two residuals
            dr = np.linalg.norm(A @ (xg - xo))
            assert abs(rg.final_norm - ro["final_norm"]) <= dr + 1e-12 * ro["norms"][0]


def test_adjoint_mixed_fixed_iterations(system):
    """The precision-form adjoint, one refinement step with a fixed number of
    inner iterations (inner tolerance 0), against the oracle's: the same
    bounds as the forward form (the transposed FP32 kernels are bit-exact,
    only the dot products' summation order differs)."""
    _, s, g, o = system
    A = sp.bsr_matrix((s.values, s.colidx, s.rowptr), shape=(s.N, s.N))
    w = np.random.default_rng(26).standard_normal(s.N)
    bounds = {1: 1e-12, 3: 1e-8, 6: 1e-7}
    for vec32 in (False, True):
        for k, bound in bounds.items():
            lg, rg = g.adjoint_solve_mixed(w, tol=0.0, max_outer=1, inner_tol=0.0, inner_maxit=k, vec32=vec32)
            lo, ro = o.adjoint_solve_mixed(w, tol=0.0, max_outer=1, inner_tol=0.0, inner_maxit=k, vec32=vec32)
            assert rg.max_iter_reached and ro["max_iter_reached"]
            assert rg.outer_iterations == ro["outer_iterations"] == 1
            assert rg.inner_half_iterations == ro["inner_half_iterations"] == 2 * k
            err = np.linalg.norm(lg - lo) / np.linalg.norm(lo)
            print(f"adjoint vec32={vec32} inner_maxit={k}: lambda rel diff {err:.3e}", flush=True)
            assert err <= bound, err
            dr = np.linalg.norm(A.T @ (lg - lo))
            assert abs(rg.final_norm - ro["final_norm"]) <= dr + 1e-12 * ro["norms"][0]


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_spmv_dev_on_caller_vectors(system, prec):
    """bda_spmv_dev on torch device tensors equals the oracle's product and
    the host-array access, including the fused sums."""
    import torch
    _, s, g, o = system
    rng = np.random.default_rng(24)
    x = rng.uniform(-1, 1, s.N).astype(DT[prec])
    w = rng.uniform(-1, 1, s.N).astype(DT[prec])
    dev = torch.device("cuda:0")
    xd, wd = torch.from_numpy(x).to(dev), torch.from_numpy(w).to(dev)
    yd = torch.empty_like(xd)
    torch.cuda.synchronize()
    d3 = g.spmv_dev(xd.data_ptr(), yd.data_ptr(), prec, w_ptr=wd.data_ptr(), red=3)
    np.testing.assert_array_equal(yd.cpu().numpy(), o.spmv_perm_ex(x, prec))
    _, dh = g.spmv_perm_ex(x, prec, w=w, red=3)
    assert d3 == dh
    g.spmv_dev(xd.data_ptr(), yd.data_ptr(), prec, transpose=True)
    np.testing.assert_array_equal(yd.cpu().numpy(), o.spmv_perm_ex(x, prec, transpose=True))
