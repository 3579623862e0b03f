"""Solving on the caller's own colour-ordered system and factors.

The reference's boundary takes L, U and the inverted diagonal blocks from its
caller, ready-made, with the colour sizes (fpga_copy_host_datamem,
fpga_functions_bicgstab.cpp:606-629; rtl/bicgstab_solver_2r-3r3w/
sparstition.vhd:673-717).  bda_upload_factored takes the same: here a handle's
system goes out through the reference's per-colour device format
(bda_wire_export), comes back in (bda_wire_import) and is solved on a second
handle that neither colours nor factorises.  With the same colour boundaries
the second handle must reproduce the first one's solve bit for bit (same
factors, same tiles, same fixed-order sums); with the device format's
partitions as colours (each a part of one colour, so the same sweeps) only
the reduction tiles differ, and the solve agrees to rounding.
"""
import numpy as np
import pytest

import fpga_amd as fa
from fpga_amd import wire
from fpga_amd.configs import GRID_128x128x64, scaled

pytestmark = pytest.mark.gpu


def native(spec):
    s = fa.generate(spec)
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    g.upload_system(s)
    g.analyse()
    g.factor()
    return s, g


@pytest.mark.parametrize("spec", [scaled(GRID_128x128x64, 20, 18, 10),
                                  fa.GridSpec(10, 9, 6, bs=4, seed=5, perm_sigma=2.0, layer_corr=0.5,
                                              inactive_frac=0.1, nnc_frac=0.05, nnc_dist=3)],
                         ids=["bs3_nnc", "bs4_nnc"])
def test_wire_roundtrip_solve_bit_identical(spec):
    print(f"case {spec}", flush=True)
    s, g = native(spec)
    h2 = h3 = None
    try:
        _, perm, cptr = g.colouring()
        sp = lambda v: np.asarray(v).reshape(s.n, s.bs)[perm].ravel()
        w = wire.export_handle(g, s.b, None)
        rowptr, colidx, vals, lu, dinv, part_rows, b_w, x0_w = wire.import_arrays(s.bs, w)
        np.testing.assert_array_equal(b_w, sp(s.b))
        h2 = fa.BicgstabSolver(s.bs, s.n, len(colidx))
        h2.upload_factored(cptr, rowptr, colidx, vals, lu, dinv)
        assert h2.num_colours == g.num_colours
        # the same factors, as the handle holds them
        for a, b in zip(h2.factors(), g.factors()):
            for u, v in zip(a if isinstance(a, tuple) else (a,), b if isinstance(b, tuple) else (b,)):
                np.testing.assert_array_equal(u, v)
        # FP64 solve, precision form and adjoint: bit for bit
        xn, rn = g.solve(s.b, tol=1e-8, maxit=500)
        xi, ri = h2.solve(b_w, tol=1e-8, maxit=500)
        assert rn.converged and ri.half_iterations == rn.half_iterations
        np.testing.assert_array_equal(xi, sp(xn))
        xn, rn = g.solve_mixed(s.b, tol=1e-8)
        xi, ri = h2.solve_mixed(b_w, tol=1e-8)
        assert rn.converged and ri.inner_half_iterations == rn.inner_half_iterations
        np.testing.assert_array_equal(xi, sp(xn))
        wv = np.random.default_rng(3).standard_normal(s.N)
        ln, rn = g.adjoint_solve(wv, tol=1e-10, maxit=500)
        li, ri = h2.adjoint_solve(sp(wv), tol=1e-10, maxit=500)
        assert rn.converged and ri.half_iterations == rn.half_iterations
        np.testing.assert_array_equal(li, sp(ln))
        # the device format's partitions as the colours
        parts = part_rows // s.bs
        assert len(parts) - 1 >= g.num_colours
        h3 = fa.BicgstabSolver(s.bs, s.n, len(colidx))
        h3.upload_factored(parts, rowptr, colidx, vals, lu, dinv)
        xn, rn = g.solve(s.b, tol=1e-8, maxit=500)
        xp, rp = h3.solve(b_w, tol=1e-8, maxit=500)
        assert rp.converged and abs(rp.half_iterations - rn.half_iterations) <= 2
        if rp.half_iterations == rn.half_iterations:
            assert np.linalg.norm(xp - sp(xn)) / np.linalg.norm(xn) < 1e-8
        np.testing.assert_array_equal(h3.precond_apply_perm(sp(wv)), g.precond_apply_perm(sp(wv)))
    finally:
        for hh in (g, h2, h3):
            if hh is not None:
                hh.close()


def test_upload_factored_rejects_bad_colouring():
    """Rows of one colour must not be coupled (the sweeps run a colour's rows
    in parallel); the check runs before anything is copied."""
    s, g = native(scaled(GRID_128x128x64, 8, 7, 5))
    h2 = None
    try:
        _, perm, cptr = g.colouring()
        w = wire.export_handle(g, s.b, None)
        rowptr, colidx, vals, lu, dinv, part_rows, b_w, x0_w = wire.import_arrays(s.bs, w)
        h2 = fa.BicgstabSolver(s.bs, s.n, len(colidx))
        merged = np.array([0, cptr[2], *cptr[3:]], np.int32)   # colours 0 and 1 as one
        with pytest.raises(fa.BdaError, match="coupled rows share a colour"):
            h2.upload_factored(merged, rowptr, colidx, vals, lu, dinv)
        with pytest.raises(fa.BdaError, match="colour_ptr"):
            h2.upload_factored(np.array([0, s.n - 1], np.int32), rowptr, colidx, vals, lu, dinv)
        h2.upload_factored(cptr, rowptr, colidx, vals, lu, dinv)   # and the handle is still usable
        x, r = h2.solve(b_w, tol=1e-8, maxit=300)
        assert r.converged
    finally:
        g.close()
        if h2 is not None:
            h2.close()
