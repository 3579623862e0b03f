"""The oracle's FP32 path (the precision form's arithmetic) pinned by cases
the arithmetic fixes exactly, so that the GPU's bit-for-bit FP32 parity
tests check against the right thing:

  - FP32 products: (1 + 2^-23)^2 rounds to 1 + 2^-22 in FP32, while an FP64
    product keeps the 2^-46 term;
  - FP64 sums: 2^25 + (1 - 2^25) is 1 in FP64 and 0 in FP32 (2^25 - 1 needs
    25 bits);
  - integer-valued values and vectors (every product and sum exact) give
    the exact A x, which pins indexing, block orientation and the
    transpose;
  - a planted FP32 accumulation (numpy, on the CPU) differs from the oracle.
"""
import numpy as np
import pytest

import oracle


def two_row_system(a01_row0, diag0_00=2.0 ** 25):
    """Block rows 0 and 1 (3x3 blocks), both coupled."""
    rowptr = np.array([0, 2, 4], np.int32)
    colidx = np.array([0, 1, 0, 1], np.int32)
    v = np.zeros((4, 3, 3))
    v[0] = np.diag([diag0_00, 1.0, 1.0])
    v[1][0] = a01_row0
    v[1][1] = [0.0, 2.0, 0.0]
    v[2] = np.eye(3) * 0.5
    v[3] = np.eye(3) * 4.0
    return rowptr, colidx, v


def oracle_sys(rowptr, colidx, v):
    o = oracle.OracleSystem(3, rowptr, colidx, v)
    o.factor()
    assert list(o.colouring()[1]) == [0, 1]  # identity permutation
    return o


def test_fp64_sums_in_fp32_path():
    o = oracle_sys(*two_row_system([1.0, -(2.0 ** 25), 0.0]))
    for prec in (1, 2):
        dt = np.float32 if prec == 1 else np.float64
        y = o.spmv_perm_ex(np.ones(6, dt), prec)
        assert y[0] == 1.0, y  # FP32 accumulation would give 0
    # the planted mistake: an FP32 accumulator loses the 1
    acc = np.float32(2.0 ** 25)
    acc = np.float32(acc + np.float32(np.float32(1.0) + np.float32(-(2.0 ** 25))))
    assert acc == 0.0


def test_fp32_products():
    e = 1.0 + 2.0 ** -23
    o = oracle_sys(*two_row_system([e, 0.0, 0.0], diag0_00=0.0 + 1.0))
    x = np.zeros(6)
    x[3] = e                       # x of block row 1, component 0
    for prec in (1, 2):
        dt = np.float32 if prec == 1 else np.float64
        y = o.spmv_perm_ex(x.astype(dt), prec)
        assert y[0] == 1.0 + 2.0 ** -22, y[0]           # FP32 product
    y64 = o.spmv_perm_ex(x, 0)
    assert y64[0] == 1.0 + 2.0 ** -22 + 2.0 ** -46      # FP64 product


@pytest.mark.parametrize("bs", [3, 4])
def test_exact_integer_products(bs):
    rng = np.random.default_rng(bs)
    n = 30
    nb = [set([i]) for i in range(n)]
    for i in range(n):
        for j in rng.choice(n, 3, replace=False):
            nb[i].add(int(j)); nb[int(j)].add(i)
    rowptr = np.zeros(n + 1, np.int32)
    cols = []
    for i in range(n):
        c = sorted(nb[i]); cols.extend(c); rowptr[i + 1] = rowptr[i] + len(c)
    colidx = np.array(cols, np.int32)
    vals = rng.integers(-8, 9, (len(cols), bs, bs)).astype(np.float64)
    for i in range(n):
        q = [k for k in range(rowptr[i], rowptr[i + 1]) if colidx[k] == i][0]
        vals[q] = np.eye(bs) * 200.0 + rng.integers(-3, 4, (bs, bs))
    o = oracle.OracleSystem(bs, rowptr, colidx, vals)
    o.factor()
    _, perm, _ = o.colouring()
    iperm = np.argsort(perm)
    # dense permuted A
    N = n * bs
    D = np.zeros((N, N))
    for i in range(n):
        for q in range(rowptr[i], rowptr[i + 1]):
            pi, pj = iperm[i], iperm[colidx[q]]
            D[pi * bs:(pi + 1) * bs, pj * bs:(pj + 1) * bs] = vals[q]
    x = rng.integers(-50, 51, N).astype(np.float64)
    for prec in (0, 1, 2):
        dt = np.float32 if prec == 1 else np.float64
        np.testing.assert_array_equal(o.spmv_perm_ex(x.astype(dt), prec), (D @ x).astype(dt))
        np.testing.assert_array_equal(o.spmv_perm_ex(x.astype(dt), prec, transpose=True), (D.T @ x).astype(dt))


def test_fp32_precond_solves_fp32_factors():
    """z = (L32 U32)^-1 p up to the FP32 rounding of the sweeps' outputs:
    the residual of the FP32 factors' dense system stays at FP32 level, and
    the FP32 path is not the FP64 one."""
    rng = np.random.default_rng(4)
    rowptr, colidx, v = two_row_system([0.3, -0.2, 0.1], diag0_00=3.0)
    v = v + rng.uniform(-0.01, 0.01, v.shape)
    o = oracle_sys(rowptr, colidx, v)
    (lr, lc, lv), (ur, uc, uv), dinv = o.factors()
    n, bs = 2, 3
    L = np.eye(6)
    U = np.zeros((6, 6))
    for i in range(n):
        for q in range(lr[i], lr[i + 1]):
            L[i * 3:i * 3 + 3, lc[q] * 3:lc[q] * 3 + 3] = lv[q].astype(np.float32)
        for q in range(ur[i], ur[i + 1]):
            U[i * 3:i * 3 + 3, uc[q] * 3:uc[q] * 3 + 3] = uv[q].astype(np.float32)
        U[i * 3:i * 3 + 3, i * 3:i * 3 + 3] = np.linalg.inv(dinv[i])
    p = rng.uniform(-1, 1, 6)
    z32 = o.precond_perm_ex(p.astype(np.float32), 1).astype(np.float64)
    z64 = o.precond_perm_ex(p, 0)
    assert np.linalg.norm(L @ U @ z32 - p.astype(np.float32)) <= 1e-6 * np.linalg.norm(p)
    assert not np.array_equal(z32, z64)


def test_refinement_stops_when_inner_run_makes_no_progress():
    """The oracle's precision form: an inner run that meets Inf/NaN before
    its first half-iteration leaves x unchanged, and the refinement ends
    instead of repeating it up to max_outer times."""
    import fpga_amd as fa
    s = fa.generate(fa.configs.scaled(fa.configs.GRID_32x32x16, 6, 5, 4))
    o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, s.values)
    o.factor()
    x, r = o.solve_mixed(s.b * 1e39, tol=1e-6, max_outer=20, vec32=True)
    assert r["nan_detected"] and not r["converged"] and r["outer_iterations"] == 1
    assert np.all(x == 0)
    x, r = o.solve_mixed(s.b, tol=1e-6, max_outer=20, vec32=True)   # a normal run is unaffected
    assert r["converged"] and not r["nan_detected"] and r["outer_iterations"] >= 1
