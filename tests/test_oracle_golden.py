"""Pin the CPU oracle against numpy/scipy fixtures and closed forms (CPU only).

Each oracle function is checked against something that does not share its
code: scipy.sparse products, dense numpy solves, the closed-form red-black
block ILU0 of a 7-point grid, scipy.sparse.linalg.bicgstab with a right
preconditioner (the same step structure as the reference microprogram,
pkg/solver_pkg.vhd:120-141), and invariants (LU = A on the pattern, first-fit
colouring).  A planted error in the oracle (wrong sign, dropped term,
transposed block) fails at least one of these.
"""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

import fpga_amd as fa
import oracle

GOLD = Path(__file__).resolve().parent / "golden"
SYS = sorted(GOLD.glob("sys_*.npz"))
RB = sorted(GOLD.glob("redblack_*.npz"))


def load(p):
    d = np.load(p, allow_pickle=False)
    return {k: d[k] for k in d.files}


def dense(d):
    n = d["rowptr"].size - 1
    bs = int(d["bs"])
    return sp.bsr_matrix((d["values"], d["colidx"], d["rowptr"]), shape=(n * bs, n * bs)).toarray()


def perm_matrix_idx(perm, bs):
    return np.concatenate([np.arange(k * bs, (k + 1) * bs) for k in perm])


@pytest.fixture(scope="module", params=SYS, ids=lambda p: p.stem)
def sysfix(request):
    d = load(request.param)
    o = oracle.OracleSystem(int(d["bs"]), d["rowptr"], d["colidx"], d["values"])
    o.factor()
    return d, o


def test_spmv_vs_scipy(sysfix):
    d, o = sysfix
    bs = int(d["bs"])
    _, perm, _ = o.colouring()
    P = perm_matrix_idx(perm, bs)
    y = o.spmv_perm(d["x"][P])
    np.testing.assert_allclose(y, d["y"][P], rtol=1e-13, atol=1e-13 * np.abs(d["y"]).max())
    # transpose: A^T x in the permuted numbering
    yt = o.spmv_perm(d["x"][P], transpose=True)
    np.testing.assert_allclose(yt, (dense(d).T @ d["x"])[P], rtol=1e-13, atol=1e-12)


def test_ilu0_matches_pattern(sysfix):
    """(L U)_ij = A_ij on every block of the colour-ordered pattern."""
    d, o = sysfix
    bs = int(d["bs"])
    _, perm, _ = o.colouring()
    P = perm_matrix_idx(perm, bs)
    Ap = dense(d)[np.ix_(P, P)]
    (lr, lc, lv), (ur, uc, uv), dinv = o.factors()
    n = lr.size - 1
    L = sp.bsr_matrix((lv, lc, lr), shape=Ap.shape).toarray() + np.eye(Ap.shape[0])
    D = sp.block_diag([np.linalg.inv(dinv[i]) for i in range(n)]).toarray()
    U = sp.bsr_matrix((uv, uc, ur), shape=Ap.shape).toarray() + D
    LU = L @ U
    mask = np.kron((np.abs(Ap).reshape(n, bs, n, bs).sum(axis=(1, 3)) > 0), np.ones((bs, bs))) > 0
    np.testing.assert_allclose(LU[mask], Ap[mask], rtol=1e-11, atol=1e-11 * np.abs(Ap).max())


def test_precond_solves_lu(sysfix):
    d, o = sysfix
    bs = int(d["bs"])
    (lr, lc, lv), (ur, uc, uv), dinv = o.factors()
    n = lr.size - 1
    N = n * bs
    L = sp.bsr_matrix((lv, lc, lr), shape=(N, N)).toarray() + np.eye(N)
    D = sp.block_diag([np.linalg.inv(dinv[i]) for i in range(n)]).toarray()
    U = sp.bsr_matrix((uv, uc, ur), shape=(N, N)).toarray() + D
    p = np.random.default_rng(3).uniform(-1, 1, N)
    z = o.precond_perm(p)
    np.testing.assert_allclose(L @ U @ z, p, rtol=1e-10, atol=1e-10)
    zt = o.precond_perm(p, transpose=True)
    np.testing.assert_allclose((L @ U).T @ zt, p, rtol=1e-10, atol=1e-10)


def test_solve_vs_dense(sysfix):
    d, o = sysfix
    x, r = o.solve(d["b"], tol=1e-12, maxit=500)
    assert r["converged"]
    err = np.linalg.norm(x - d["x_solve"]) / np.linalg.norm(d["x_solve"])
    assert err < 1e-9, err


def test_adjoint_and_gradient_vs_dense(sysfix):
    d, o = sysfix
    lam, r = o.adjoint_solve(d["w"], tol=1e-12, maxit=500)
    assert r["converged"]
    assert np.linalg.norm(lam - d["lam"]) / np.linalg.norm(d["lam"]) < 1e-9
    g = o.grad(d["lam"], d["x_solve"])
    np.testing.assert_allclose(g, d["grad"], rtol=1e-13, atol=1e-14)


def test_mixed_error_vs_dense(sysfix):
    d, o = sysfix
    for vec32 in (False, True):
        x, r = o.solve_mixed(d["b"], tol=1e-8, inner_tol=1e-4, vec32=vec32)
        assert r["converged"]
        err = np.linalg.norm(x - d["x_solve"]) / np.linalg.norm(d["x_solve"])
        assert err < 1e-6, (vec32, err)


# --- closed-form red-black ILU0 and scipy BiCGSTAB ---------------------------
@pytest.fixture(scope="module", params=RB, ids=lambda p: p.stem)
def rbfix(request):
    d = load(request.param)
    o = oracle.OracleSystem(int(d["bs"]), d["rowptr"], d["colidx"], d["values"])
    o.factor()
    return d, o


def test_redblack_colouring_closed_form(rbfix):
    d, o = rbfix
    colour, perm, cptr = o.colouring()
    nx, ny, nz = d["grid"]
    i = np.arange(nx * ny * nz)
    np.testing.assert_array_equal(colour, ((i % nx) + (i // nx) % ny + i // (nx * ny)) % 2)
    np.testing.assert_array_equal(perm, d["perm"])


def test_redblack_ilu0_closed_form(rbfix):
    d, o = rbfix
    bs = int(d["bs"])
    (lr, lc, lv), (ur, uc, uv), dinv = o.factors()
    n = lr.size - 1
    N = n * bs
    L = sp.bsr_matrix((lv, lc, lr), shape=(N, N)).toarray() + np.eye(N)
    U = sp.bsr_matrix((uv, uc, ur), shape=(N, N)).toarray()
    Uexp = d["U"].copy()
    for k in range(n):
        Uexp[k * bs:(k + 1) * bs, k * bs:(k + 1) * bs] = 0.0
    np.testing.assert_allclose(L, d["L"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(U, Uexp, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(dinv, d["dinv"], rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(o.precond_perm(d["p"]), d["z"], rtol=1e-10, atol=1e-12)


def test_redblack_bicgstab_vs_scipy(rbfix):
    """Same step structure as scipy.sparse.linalg.bicgstab (right
    preconditioned, rtilde = r0, early exit on ||s||); scipy stops on a
    strict <, after `scipy_iters` full iterations or half way through the
    next one."""
    d, o = rbfix
    bs = int(d["bs"])
    P = perm_matrix_idx(d["perm"], bs)
    x, r = o.solve(d["b"], tol=1e-8, maxit=200)
    k = int(d["scipy_iters"])
    assert int(d["scipy_info"]) == 0 and r["converged"]
    assert k <= r["iterations"] <= k + 1, (k, r["iterations"])
    np.testing.assert_allclose(x[P], d["x_bicgstab"], rtol=0, atol=1e-9 * np.abs(d["x_bicgstab"]).max())
    assert np.linalg.norm(x[P] - d["x_dense"]) / np.linalg.norm(d["x_dense"]) < 1e-6


# --- microprogram semantics ---------------------------------------------------
def small_oracle():
    s = fa.generate(fa.GridSpec(6, 5, 4, bs=3, seed=2, perm_sigma=1.0, nnc_frac=0.1, nnc_dist=2))
    o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, s.values)
    o.factor()
    return s, o


def test_maxit_counts_half_iterations():
    """maxit M: M full iterations, then the first half of iteration M+1
    (solver.vhd:557 compares iter_num[15:1] with max_iters at steps 7/8)."""
    s, o = small_oracle()
    for m in range(4):
        _, r = o.solve(s.b, tol=0.0, maxit=m)
        assert r["max_iter_reached"] and r["half_iterations"] == 2 * m and r["iterations"] == m + 0.5


def test_norm_ring_and_no_results():
    s, o = small_oracle()
    x, r = o.solve(s.b, tol=0.0, maxit=3)
    h = r["half_iterations"]
    assert r["last_norm_idx"] == (h % 3) + 1          # bicgstab_utils.cpp:256
    assert r["norms"][0] == pytest.approx(np.linalg.norm(s.b), rel=1e-13)
    x, r = o.solve(np.zeros(s.N), tol=1e-6, maxit=10)
    assert r["no_results"] and not r["converged"] and np.all(x == 0)
    x, r = o.solve(s.b, tol=1e9, maxit=10, abs_tol=True)
    assert r["no_results"]
    x, r = o.solve(s.b, tol=1e-3, maxit=100, abs_tol=True)
    assert r["converged"] and r["final_norm"] <= 1e-3


def test_nan_flag():
    s, o = small_oracle()
    b = s.b.copy()
    b[3] = np.nan
    _, r = o.solve(b, tol=1e-6, maxit=10)
    assert r["nan_detected"] and not r["converged"]


def test_planted_errors_are_caught():
    """A corrupted matrix (one transposed off-diagonal block) changes the
    oracle's SpMV away from scipy's result: the fixtures do discriminate."""
    d = load(SYS[0])
    vals = d["values"].copy()
    q = int(np.nonzero(d["colidx"] != np.repeat(np.arange(d["rowptr"].size - 1), np.diff(d["rowptr"])))[0][0])
    vals[q] = vals[q].T
    o = oracle.OracleSystem(int(d["bs"]), d["rowptr"], d["colidx"], vals)
    bs = int(d["bs"])
    _, perm, _ = o.colouring()
    P = perm_matrix_idx(perm, bs)
    assert not np.allclose(o.spmv_perm(d["x"][P]), d["y"][P], rtol=1e-10)


# --- colouring hashes ---------------------------------------------------------
def test_colouring_hashes():
    import hashlib
    gold = json.loads((GOLD / "colouring.json").read_text())
    for name, spec in [("grid32x32x16", fa.configs.GRID_32x32x16), ("grid64x64x32", fa.configs.GRID_64x64x32)]:
        s = fa.generate(spec)
        colour = oracle.colour(s.rowptr, s.colidx)
        o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, s.values)
        _, perm, cptr = o.colouring()
        h = lambda a: hashlib.sha256(np.ascontiguousarray(a, np.int32).tobytes()).hexdigest()
        assert h(colour) == gold[name]["colour_sha256"]
        assert h(perm) == gold[name]["perm_sha256"]
        assert list(cptr) == gold[name]["colour_ptr"]
    s = fa.generate(fa.configs.GRID_32x32x16)
    mh = hashlib.sha256(s.rowptr.tobytes() + s.colidx.tobytes() + s.values.tobytes() + s.b.tobytes()).hexdigest()
    assert mh == gold["grid32x32x16"]["matrix_sha256"] and s.nnzb == gold["grid32x32x16"]["nnzb"] == 110592


def test_oracle_norm_history_matches_status():
    """The recorded history starts with ||r0||, ends with the final norm,
    holds h + 2 values after h half iterations, and its last three values are
    the norm ring (newest at (h % 3) + 1)."""
    s = fa.generate(fa.configs.scaled(fa.configs.GRID_32x32x16, 10, 8, 6))
    o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, s.values)
    o.factor()
    for tol, maxit in [(1e-8, 200), (0.0, 3)]:
        x, r = o.solve(s.b, tol=tol, maxit=maxit)
        h = o.history()
        assert len(h) == r["half_iterations"] + 2
        assert h[0] == r["norms"][0] and h[-1] == r["final_norm"]
        last = r["last_norm_idx"]
        ring = [r["norms"][(last - 1 - k) % 3 + 1] for k in range(3)]   # newest first
        assert list(h[::-1][:3]) == ring
