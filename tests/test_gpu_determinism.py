"""Run-to-run determinism: every reduction sums per-tile or per-block
partials in a fixed order (each tile writes its own slot), so a solve
repeated on one handle, or on a second handle built from the same matrix,
returns bit-identical x and the same iteration counts, whatever the
scheduling of workgroups on the GPU."""
import numpy as np
import pytest

import fpga_amd as fa

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    print("case grid64x64x32 determinism", flush=True)
    return fa.generate(fa.configs.GRID_64x64x32)


def make(s):
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    g.upload_system(s)
    g.analyse()
    g.factor()
    return g


def test_fp64_solve_repeats_bit_identical(system):
    s = system
    g1, g2 = make(s), make(s)
    try:
        x1, r1 = g1.solve(s.b, tol=1e-8, maxit=500)
        x1b, r1b = g1.solve(s.b, tol=1e-8, maxit=500)
        x2, r2 = g2.solve(s.b, tol=1e-8, maxit=500)
        assert r1.converged
        assert r1.half_iterations == r1b.half_iterations == r2.half_iterations
        np.testing.assert_array_equal(x1, x1b)
        np.testing.assert_array_equal(x1, x2)
        assert r1.final_norm == r1b.final_norm == r2.final_norm
    finally:
        g1.close()
        g2.close()


def test_mixed_solve_repeats_bit_identical(system):
    s = system
    g = make(s)
    try:
        xa, ra = g.solve_mixed(s.b, tol=1e-8)
        xb, rb = g.solve_mixed(s.b, tol=1e-8)
        assert ra.converged
        assert ra.inner_half_iterations == rb.inner_half_iterations
        assert ra.outer_iterations == rb.outer_iterations
        np.testing.assert_array_equal(xa, xb)
    finally:
        g.close()


@pytest.mark.parametrize("poll_iters", [2, 3])
def test_host_lead_does_not_change_results(system, poll_iters):
    """How far the host queues ahead (bda_set_options' poll_iters) changes
    only how many launches are pending, never the results, including the
    no-op launches left queued when a run stops in either half-iteration."""
    s = system
    g = make(s)
    try:
        x1, r1 = g.solve(s.b, tol=1e-8, maxit=500)
        xm1, rm1 = g.solve_mixed(s.b, tol=1e-8)
        g.set_options(poll_iters=poll_iters)
        x2, r2 = g.solve(s.b, tol=1e-8, maxit=500)
        xm2, rm2 = g.solve_mixed(s.b, tol=1e-8)
        x3, r3 = g.solve(s.b, tol=0.0, maxit=7)  # stops at a first half (max-iteration exit)
        g.set_options(poll_iters=1)
        x4, r4 = g.solve(s.b, tol=0.0, maxit=7)
        assert r1.half_iterations == r2.half_iterations
        np.testing.assert_array_equal(x1, x2)
        assert rm1.inner_half_iterations == rm2.inner_half_iterations
        np.testing.assert_array_equal(xm1, xm2)
        assert r3.max_iter_reached and r3.half_iterations == r4.half_iterations == 14
        np.testing.assert_array_equal(x3, x4)
    finally:
        g.close()
