"""The step-0 residual with a zero start vector takes a short path (r = b
without the product; the full product adds only signed zeros to a row sum
that starts from +0.0, so it gives b exactly).  A start vector that is zero
except for one denormal entry takes the full path: its products round away
against b and the entry is absorbed by the first update, so both solves must
agree bit for bit, in the FP64 form and in the precision form."""
import numpy as np
import pytest

import fpga_amd as fa

pytestmark = pytest.mark.gpu


SPECS = {
    "grid64x64x32": fa.configs.GRID_64x64x32,
    "grid48x48x24_bs4_nnc": fa.GridSpec(48, 48, 24, bs=4, seed=2, perm_sigma=2.0, layer_corr=0.5,
                                        inactive_frac=0.10, nnc_frac=0.01, nnc_dist=3),
}


@pytest.fixture(scope="module", params=sorted(SPECS))
def handle(request):
    print(f"case {request.param} zero start", flush=True)
    s = fa.generate(SPECS[request.param])
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    g.upload_system(s)
    g.analyse()
    g.factor()
    yield s, g
    g.close()


def tiny_start(s):
    x0 = np.zeros(s.N)
    x0[s.N // 3] = 5e-324
    return x0


def test_fp64_zero_start_matches_full_path(handle):
    s, g = handle
    xa, ra = g.solve(s.b, tol=1e-8, maxit=500)
    xb, rb = g.solve(s.b, x0=tiny_start(s), tol=1e-8, maxit=500)
    assert ra.converged
    assert ra.half_iterations == rb.half_iterations
    assert list(ra.norms) == list(rb.norms) and ra.final_norm == rb.final_norm
    np.testing.assert_array_equal(xa, xb)


def test_mixed_zero_start_matches_full_path(handle):
    s, g = handle
    xa, ra = g.solve_mixed(s.b, tol=1e-8)
    xb, rb = g.solve_mixed(s.b, x0=tiny_start(s), tol=1e-8)
    assert ra.converged
    assert ra.inner_half_iterations == rb.inner_half_iterations
    assert ra.outer_iterations == rb.outer_iterations
    np.testing.assert_array_equal(xa, xb)
