"""GPU parity on edge-case shapes: one block row, two block rows, a matrix
of diagonal blocks only (one colour, empty L and U), a row holding exactly
the most blocks a tile takes (168 for 3x3, 128 for 4x4) next to rows of
two to four blocks, and a 10-row clique inside a grid (ten colours, most of
them one or two rows).  For each: colouring, factors, SpMV and the
preconditioner in every kernel form (FP64; FP32 values with FP32 or FP64
vectors, forward and transposed) bit for bit against the oracle; the FP64,
precision-form and adjoint solves against the oracle's.  A row one block
wider than a tile takes is refused by the analysis with a clear message.

The SpMV and sweep row semantics follow spmvp.vhd:240-383 and
ilu0.vhd:385-416 (L: y = p - L y; U: z = D^-1 (y - U z)).
"""
import numpy as np
import pytest

import fpga_amd as fa
import oracle

pytestmark = pytest.mark.gpu

DT = {0: np.float64, 1: np.float32, 2: np.float64}


def from_neighbours(bs, nb, seed):
    """Block CSR of the (symmetric) neighbour sets nb, diagonally dominant."""
    rng = np.random.default_rng(seed)
    n = len(nb)
    rowptr = np.zeros(n + 1, np.int32)
    cols = []
    for i in range(n):
        c = sorted(nb[i] | {i})
        cols.extend(c)
        rowptr[i + 1] = rowptr[i] + len(c)
    colidx = np.array(cols, np.int32)
    vals = rng.uniform(-0.5, 0.5, (len(cols), bs, bs))
    for i in range(n):
        q = np.arange(rowptr[i], rowptr[i + 1])
        qd = q[colidx[q] == i][0]
        off = sum((np.abs(vals[p]).sum(axis=1) for p in q if p != qd), np.zeros(bs))
        vals[qd] = np.diag(off + 1.0) + rng.normal(0.0, 0.05, (bs, bs))
    b = rng.uniform(-1, 1, n * bs)
    return rowptr, colidx, vals, b


def couple(nb, i, j):
    nb[i].add(j)
    nb[j].add(i)


def arrow(bs, n=400):
    """Row 0 couples to rows 1..w-1 (w = the tile capacity, so row 0 fills a
    tile on its own); rows 1..n-1 form a chain."""
    w = 168 if bs == 3 else 128
    nb = [set() for _ in range(n)]
    for j in range(1, w):
        couple(nb, 0, j)
    for i in range(1, n - 1):
        couple(nb, i, i + 1)
    return nb


def clique_in_grid(nx=10, ny=10, nz=4, k=10):
    n = nx * ny * nz
    nb = [set() for _ in range(n)]
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = x + nx * (y + ny * z)
                if x + 1 < nx:
                    couple(nb, i, i + 1)
                if y + 1 < ny:
                    couple(nb, i, i + nx)
                if z + 1 < nz:
                    couple(nb, i, i + nx * ny)
    members = [(n // k) * m + 3 for m in range(k)]
    for a in members:
        for c in members:
            if a < c:
                couple(nb, a, c)
    return nb


CASES = {
    "one_row_bs3": (3, lambda: [set()]),
    "one_row_bs4": (4, lambda: [set()]),
    "two_rows_bs4": (4, lambda: [{1}, {0}]),
    "diagonal_only_bs3": (3, lambda: [set() for _ in range(500)]),
    "arrow_full_tile_bs3": (3, lambda: arrow(3)),
    "arrow_full_tile_bs4": (4, lambda: arrow(4)),
    "clique_in_grid_bs3": (3, clique_in_grid),
}


@pytest.fixture(scope="module", params=list(CASES))
def system(request):
    bs, make = CASES[request.param]
    rowptr, colidx, vals, b = from_neighbours(bs, make(), seed=len(request.param))
    n = len(rowptr) - 1
    print(f"case {request.param}: bs={bs} n={n} nnzb={len(colidx)}", flush=True)
    g = fa.BicgstabSolver(bs, n, len(colidx))
    g.upload_bcsr(rowptr, colidx, vals)
    g.analyse()
    g.factor()
    o = oracle.OracleSystem(bs, rowptr, colidx, vals)
    o.factor()
    yield request.param, bs, b, g, o
    g.close()


def test_colouring_and_factors(system):
    name, bs, b, g, o = system
    for a, c in zip(g.colouring(), o.colouring()):
        np.testing.assert_array_equal(a, c)
    nc = len(g.colouring()[2]) - 1
    expect = {"one_row_bs3": 1, "one_row_bs4": 1, "two_rows_bs4": 2, "diagonal_only_bs3": 1}
    if name in expect:
        assert nc == expect[name]
    if name.startswith("clique"):
        assert nc >= 10
    (glr, glc, glv), (gur, guc, guv), gd = g.factors()
    (olr, olc, olv), (our, ouc, ouv), od = o.factors()
    for a, c in ((glr, olr), (glc, olc), (gur, our), (guc, ouc), (glv, olv), (guv, ouv), (gd, od)):
        np.testing.assert_array_equal(a, c)


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_spmv_and_precond_bit_exact(system, prec):
    _, bs, b, g, o = system
    rng = np.random.default_rng(40 + prec)
    x = rng.uniform(-1, 1, b.size).astype(DT[prec])
    for tr in (False, True):
        np.testing.assert_array_equal(g.spmv_perm_ex(x, prec, transpose=tr), o.spmv_perm_ex(x, prec, transpose=tr))
        np.testing.assert_array_equal(g.precond_apply_perm_ex(x, prec, transpose=tr),
                                      o.precond_perm_ex(x, prec, transpose=tr))


def test_solves(system):
    name, bs, b, g, o = system
    xg, rg = g.solve(b, tol=1e-10, maxit=500)
    xo, ro = o.solve(b, tol=1e-10, maxit=500)
    print(name, rg.half_iterations, ro["half_iterations"], flush=True)
    assert rg.converged and ro["converged"]
    assert abs(rg.half_iterations - ro["half_iterations"]) <= 2
    assert np.linalg.norm(xg - xo) <= 1e-8 * np.linalg.norm(xo)
    xm, rm = g.solve_mixed(b, tol=1e-8)
    assert rm.converged
    assert np.linalg.norm(xm - xo) <= 1e-6 * np.linalg.norm(xo)
    w = np.random.default_rng(7).standard_normal(b.size)
    lg, ra = g.adjoint_solve(w, tol=1e-10, maxit=500)
    lo, rao = o.adjoint_solve(w, tol=1e-10, maxit=500)
    assert ra.converged and rao["converged"]
    assert np.linalg.norm(lg - lo) <= 1e-8 * np.linalg.norm(lo)


@pytest.mark.parametrize("bs", [3, 4])
def test_row_wider_than_a_tile_is_refused(bs):
    nb = arrow(bs)
    nb[0].add(len(nb) - 1)      # one block more than a tile takes
    nb[len(nb) - 1].add(0)
    rowptr, colidx, vals, _ = from_neighbours(bs, nb, seed=3)
    print(f"case row wider than a tile: bs={bs} row 0 has {rowptr[1] - rowptr[0]} blocks", flush=True)
    g = fa.BicgstabSolver(bs, len(rowptr) - 1, len(colidx))
    try:
        g.upload_bcsr(rowptr, colidx, vals)
        with pytest.raises(fa.BdaError, match="at most"):
            g.analyse()
    finally:
        g.close()
