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


def test_handles_on_two_threads():
    """One handle per thread, as the reference's callers use it: two threads
    solving two different systems at the same time (each on its own handle
    and stream; the ctypes calls release the GIL) get the same iteration
    counts and bit-identical x as the same solves run alone."""
    import threading
    from fpga_amd.configs import GRID_128x128x64, scaled
    specs = [fa.configs.GRID_32x32x16, scaled(GRID_128x128x64, 20, 18, 10)]
    systems = [fa.generate(sp) for sp in specs]
    handles = [make(s) for s in systems]
    try:
        solo = []
        for s, g in zip(systems, handles):
            x, r = g.solve(s.b, tol=1e-8, maxit=500)
            xm, rm = g.solve_mixed(s.b, tol=1e-8)
            solo.append((x, r.half_iterations, xm, rm.inner_half_iterations))
        out = [[] for _ in handles]
        errors = []

        def work(k):
            try:
                s, g = systems[k], handles[k]
                for _ in range(3):
                    x, r = g.solve(s.b, tol=1e-8, maxit=500)
                    xm, rm = g.solve_mixed(s.b, tol=1e-8)
                    out[k].append((x, r.half_iterations, xm, rm.inner_half_iterations))
            except Exception as e:  # reported below, on the main thread
                errors.append(e)

        print("case two threads, two handles", flush=True)
        ts = [threading.Thread(target=work, args=(k,)) for k in range(2)]
        for t in ts:
            t.start()
        for t in ts:
            t.join(timeout=300)
        assert not any(t.is_alive() for t in ts)
        assert not errors, errors
        for k in range(2):
            assert len(out[k]) == 3
            for x, h, xm, hm in out[k]:
                assert h == solo[k][1] and hm == solo[k][3]
                np.testing.assert_array_equal(x, solo[k][0])
                np.testing.assert_array_equal(xm, solo[k][2])
    finally:
        for g in handles:
            g.close()
