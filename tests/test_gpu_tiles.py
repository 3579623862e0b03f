"""GPU parity on a matrix built so that the sweeps take the tall tiling.

Most block rows are uncoupled; a few couple to rows far away, which gives
three colours whose rows carry at most one or two lower or upper blocks.
With 64-row tiles such a sweep colour needs twice the workgroups of the tall
(128-row) tiling, so every sweep launch of the preconditioner except the
last colour's takes the 128-row kernel: the lower sweep of colour 1, and the
upper sweeps of colours 1 and 0.  The results must stay bit-identical to the
oracle, as for the other tilings.
"""
import numpy as np
import pytest

import fpga_amd as fa
import oracle

pytestmark = pytest.mark.gpu


def sparse_coupled(bs, n=2000, seed=0):
    """Block rows i < 200 couple to 1000 + i, and 1000 + i (i < 50) to 1500 + i,
    which also couples to i; rows 200 + i (i < 40) couple to 1000 + i and
    1001 + i, so that ILU0 drops fill.  Greedy colouring in natural order gives
    colour 0 to 1,750 rows, colour 1 to rows 1000..1199 and colour 2 to
    1500..1549."""
    rng = np.random.default_rng(seed)
    nb = [set([i]) for i in range(n)]
    for i in range(200):
        nb[i].add(1000 + i); nb[1000 + i].add(i)
    for i in range(50):
        nb[1000 + i].add(1500 + i); nb[1500 + i].add(1000 + i)
        nb[i].add(1500 + i); nb[1500 + i].add(i)
    for i in range(40):  # fill that ILU0 drops (between rows 1000 + i and 1001 + i)
        for j in (1000 + i, 1001 + i):
            nb[200 + i].add(j); nb[j].add(200 + i)
    rowptr = np.zeros(n + 1, np.int32)
    cols = []
    for i in range(n):
        c = sorted(nb[i])
        cols.extend(c)
        rowptr[i + 1] = rowptr[i] + len(c)
    colidx = np.array(cols, np.int32)
    vals = rng.uniform(-0.5, 0.5, (len(cols), bs, bs))
    for i in range(n):
        for q in range(rowptr[i], rowptr[i + 1]):
            if colidx[q] == i:
                off = sum(np.abs(vals[p]).sum(axis=1) for p in range(rowptr[i], rowptr[i + 1]) if colidx[p] != i)
                off = off if np.ndim(off) else np.zeros(bs)
                vals[q] = np.diag(off + 1.0) + rng.normal(0.0, 0.05, (bs, bs))
    b = rng.uniform(-1, 1, n * bs)
    return rowptr, colidx, vals, b


@pytest.fixture(scope="module", params=[3, 4])
def pair(request):
    bs = request.param
    print(f"case sparse_coupled bs={bs}", flush=True)
    rowptr, colidx, vals, b = sparse_coupled(bs)
    g = fa.BicgstabSolver(bs, len(rowptr) - 1, len(colidx))
    g.upload_bcsr(rowptr, colidx, vals)
    g.analyse()
    g.factor()
    o = oracle.OracleSystem(bs, rowptr, colidx, vals)
    o.factor()
    yield bs, b, g, o
    g.close()


def test_colours_as_built(pair):
    bs, b, g, o = pair
    colour, perm, cptr = g.colouring()
    assert g.num_colours == 3
    np.testing.assert_array_equal(np.diff(cptr), [1750, 200, 50])
    # the last colour and colour 1's upper sweep share a launch; colour 0 has
    # 90 rows that reach colour 2 or a colour-1 row with blocks, too many for
    # the extended tail, so its upper sweep keeps its own launch
    assert g.precond_launches() == 3


def test_precond_bit_exact_tall_tiles(pair):
    bs, b, g, o = pair
    p = np.random.default_rng(5).uniform(-1, 1, len(b))
    np.testing.assert_array_equal(g.precond_apply_perm(p), o.precond_perm(p))
    np.testing.assert_array_equal(g.precond_apply_perm(p, transpose=True), o.precond_perm(p, transpose=True))
    # the precision form's instances of the same launches (FP32 values; float
    # and double vectors)
    for prec, dt in ((1, np.float32), (2, np.float64)):
        for tr in (False, True):
            np.testing.assert_array_equal(g.precond_apply_perm_ex(p.astype(dt), prec, transpose=tr),
                                          o.precond_perm_ex(p.astype(dt), prec, transpose=tr))


def test_spmv_bit_exact(pair):
    bs, b, g, o = pair
    x = np.random.default_rng(6).uniform(-1, 1, len(b))
    np.testing.assert_array_equal(g.spmv_perm(x), o.spmv_perm(x))
    for prec, dt in ((1, np.float32), (2, np.float64)):
        np.testing.assert_array_equal(g.spmv_perm_ex(x.astype(dt), prec), o.spmv_perm_ex(x.astype(dt), prec))


def test_solve_parity(pair):
    bs, b, g, o = pair
    xg, rg = g.solve(b, tol=1e-10, maxit=200)
    xo, ro = o.solve(b, tol=1e-10, maxit=200)
    assert rg.converged and ro["converged"]
    assert abs(rg.half_iterations - ro["half_iterations"]) <= 2
    np.testing.assert_allclose(xg, xo, rtol=0, atol=1e-8 * np.abs(xo).max())
    xm, rm = g.solve_mixed(b, tol=1e-8)
    assert rm.converged
    np.testing.assert_allclose(xm, xo, rtol=0, atol=1e-6 * np.abs(xo).max())


def fan_coupled(bs, nlast, fan, n=2000, seed=1):
    """Rows 1000 + q (q < 200) couple to q and take colour 1; rows 1500 + i
    (i < nlast) couple to i and to the `fan` colour-1 rows 1000 + i*fan + j,
    and take colour 2.  Colour 1 then has nlast * fan rows with blocks into
    the last colour: 64 of them still fit the merged launch of the last
    colour with colour 1's upper sweep, 70 take the separate launches."""
    rng = np.random.default_rng(seed)
    nb = [set([i]) for i in range(n)]
    for q in range(200):
        nb[1000 + q].add(q); nb[q].add(1000 + q)
    for i in range(nlast):
        nb[1500 + i].add(i); nb[i].add(1500 + i)
        for j in range(fan):
            c = 1000 + i * fan + j
            nb[1500 + i].add(c); nb[c].add(1500 + i)
    rowptr = np.zeros(n + 1, np.int32)
    cols = []
    for i in range(n):
        c = sorted(nb[i])
        cols.extend(c)
        rowptr[i + 1] = rowptr[i] + len(c)
    colidx = np.array(cols, np.int32)
    vals = rng.uniform(-0.5, 0.5, (len(cols), bs, bs))
    for i in range(n):
        qd = [q for q in range(rowptr[i], rowptr[i + 1]) if colidx[q] == i][0]
        off = sum((np.abs(vals[q]).sum(axis=1) for q in range(rowptr[i], rowptr[i + 1]) if q != qd), np.zeros(bs))
        vals[qd] = np.diag(off + 1.0) + rng.normal(0.0, 0.05, (bs, bs))
    return rowptr, colidx, vals


@pytest.mark.parametrize("bs", [3, 4])
@pytest.mark.parametrize("nlast,fan", [(8, 8), (10, 7)])
def test_precond_bit_exact_tail_limit(bs, nlast, fan):
    print(f"case fan_coupled bs={bs} nlast={nlast} fan={fan}", flush=True)
    rowptr, colidx, vals = fan_coupled(bs, nlast, fan)
    g = fa.BicgstabSolver(bs, len(rowptr) - 1, len(colidx))
    try:
        g.upload_bcsr(rowptr, colidx, vals)
        g.analyse()
        g.factor()
        _, _, cptr = g.colouring()
        np.testing.assert_array_equal(np.diff(cptr), [2000 - 200 - nlast, 200, nlast])
        # 64 colour-1 rows with blocks: one extended tail launch after the lower
        # sweep of colour 1 (colour 0's 64 rows that reach them fit it too); 70:
        # every colour in a launch of its own
        assert g.precond_launches() == (2 if nlast * fan <= 64 else 4)
        o = oracle.OracleSystem(bs, rowptr, colidx, vals)
        o.factor()
        p = np.random.default_rng(7).uniform(-1, 1, (len(rowptr) - 1) * bs)
        np.testing.assert_array_equal(g.precond_apply_perm(p), o.precond_perm(p))
        np.testing.assert_array_equal(g.precond_apply_perm(p, transpose=True), o.precond_perm(p, transpose=True))
        for prec, dt in ((1, np.float32), (2, np.float64)):
            for tr in (False, True):
                np.testing.assert_array_equal(g.precond_apply_perm_ex(p.astype(dt), prec, transpose=tr),
                                              o.precond_perm_ex(p.astype(dt), prec, transpose=tr))
        b = np.random.default_rng(8).uniform(-1, 1, len(p))
        xg, rg = g.solve(b, tol=1e-10, maxit=200)
        xo, ro = o.solve(b, tol=1e-10, maxit=200)
        assert rg.converged and ro["converged"]
        np.testing.assert_allclose(xg, xo, rtol=0, atol=1e-8 * np.abs(xo).max())
    finally:
        g.close()


def chain4(bs, n=2000, seed=2):
    """Four colours: rows 1000 + q (q < 200) couple to q (colour 1); rows
    1500 + i (i < 50) to 300 + i and 1000 + i (colour 2); rows 1800 + j
    (j < 10) to 400 + j, 1100 + j and 1500 + j (colour 3).  Rows 1100 + j also
    couple to 1500 + j and 1510 + j.  Colour 2's rows 1500..1509 then carry
    blocks into colour 3, and colour 1's rows 1000..1009 and 1100..1109 reach
    them or colour 3: the extended tail computes those 20 rows in its first
    workgroup, rows 1100 + j with one block each whose z comes from colour 3
    (LDS), from a colour-2 row with blocks (LDS) and from an ordinary colour-2
    row (memory, written by the lower sweep's launch of colour 2)."""
    rng = np.random.default_rng(seed)
    nb = [set([i]) for i in range(n)]

    def link(a, c):
        nb[a].add(c); nb[c].add(a)

    for q in range(200):
        link(1000 + q, q)
    for i in range(50):
        link(1500 + i, 300 + i); link(1500 + i, 1000 + i)
    for j in range(10):
        link(1800 + j, 400 + j); link(1800 + j, 1100 + j); link(1800 + j, 1500 + j)
        link(1100 + j, 1500 + j); link(1100 + j, 1510 + j)
    rowptr = np.zeros(n + 1, np.int32)
    cols = []
    for i in range(n):
        c = sorted(nb[i])
        cols.extend(c)
        rowptr[i + 1] = rowptr[i] + len(c)
    colidx = np.array(cols, np.int32)
    vals = rng.uniform(-0.5, 0.5, (len(cols), bs, bs))
    for i in range(n):
        qd = [q for q in range(rowptr[i], rowptr[i + 1]) if colidx[q] == i][0]
        off = sum((np.abs(vals[q]).sum(axis=1) for q in range(rowptr[i], rowptr[i + 1]) if q != qd), np.zeros(bs))
        vals[qd] = np.diag(off + 1.0) + rng.normal(0.0, 0.05, (bs, bs))
    return rowptr, colidx, vals


@pytest.mark.parametrize("bs", [3, 4])
def test_precond_bit_exact_extended_tail(bs):
    print(f"case chain4 bs={bs}", flush=True)
    rowptr, colidx, vals = chain4(bs)
    g = fa.BicgstabSolver(bs, len(rowptr) - 1, len(colidx))
    try:
        g.upload_bcsr(rowptr, colidx, vals)
        g.analyse()
        g.factor()
        _, _, cptr = g.colouring()
        np.testing.assert_array_equal(np.diff(cptr), [1740, 200, 50, 10])
        # the lower sweeps of colours 1 and 2 in one launch (colour 2's tiles
        # recompute the colour-1 rows they read), the extended tail, the upper
        # sweep of colour 0
        assert g.precond_launches() == 3
        o = oracle.OracleSystem(bs, rowptr, colidx, vals)
        o.factor()
        p = np.random.default_rng(9).uniform(-1, 1, (len(rowptr) - 1) * bs)
        np.testing.assert_array_equal(g.precond_apply_perm(p), o.precond_perm(p))
        np.testing.assert_array_equal(g.precond_apply_perm(p, transpose=True), o.precond_perm(p, transpose=True))
        for prec, dt in ((1, np.float32), (2, np.float64)):
            for tr in (False, True):
                np.testing.assert_array_equal(g.precond_apply_perm_ex(p.astype(dt), prec, transpose=tr),
                                              o.precond_perm_ex(p.astype(dt), prec, transpose=tr))
        b = np.random.default_rng(10).uniform(-1, 1, len(p))
        xg, rg = g.solve(b, tol=1e-10, maxit=200)
        xo, ro = o.solve(b, tol=1e-10, maxit=200)
        assert rg.converged and ro["converged"]
        assert abs(rg.half_iterations - ro["half_iterations"]) <= 2
        np.testing.assert_allclose(xg, xo, rtol=0, atol=1e-8 * np.abs(xo).max())
        xm, rm = g.solve_mixed(b, tol=1e-8)
        assert rm.converged
        np.testing.assert_allclose(xm, xo, rtol=0, atol=1e-6 * np.abs(xo).max())
    finally:
        g.close()
