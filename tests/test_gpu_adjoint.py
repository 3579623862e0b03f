"""The hand-written adjoint behind the LibTorch autograd Function (GPU).

- dense check: dL/db = A^{-T} w and dL/dA = -lambda x^T on the block pattern,
  against numpy on small systems;
- central finite differences (h = 1e-6, FP64) on 64 random matrix entries and
  64 random right-hand-side entries of the 64x64x32 grid, L = w^T x with
  w ~ N(0,1), adjoint solved to 1e-10; required agreement 1e-6 relative.
  Each perturbed solve restarts from the unperturbed solution and runs to
  1e-11 relative to the O(h) change of the residual, and the difference
  quotient is w^T (x+ - x-) / 2h, formed from the difference vector, so that
  neither the solve tolerance nor the cancellation of two O(1) losses adds
  noise of the order of the check;
- a backward pass after a second forward call on the same handle uses the
  values of its own forward call.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import fpga_amd as fa
from fpga_amd.autograd import DifferentiableSolver
from fpga_amd.configs import GRID_64x64x32

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spec", [fa.GridSpec(6, 5, 4, bs=3, seed=1, nnc_frac=0.1, nnc_dist=2),
                                  fa.GridSpec(5, 4, 3, bs=4, seed=2, perm_sigma=2.0, inactive_frac=0.1,
                                              nnc_frac=0.2)], ids=["bs3", "bs4"])
def test_autograd_vs_dense(spec):
    s = fa.generate(spec)
    dev = torch.device("cuda:0")
    ds = DifferentiableSolver(s.bs, s.rowptr, s.colidx, s.values)
    vals = torch.tensor(s.values, device=dev, requires_grad=True)
    b = torch.tensor(s.b, device=dev, requires_grad=True)
    w = torch.tensor(np.random.default_rng(0).standard_normal(s.N), device=dev)
    x = ds(vals, b, tol=1e-13, adj_tol=1e-13)
    (w * x).sum().backward()
    A = sp.bsr_matrix((s.values, s.colidx, s.rowptr), shape=(s.N, s.N)).toarray()
    xd = np.linalg.solve(A, s.b)
    lam = np.linalg.solve(A.T, w.cpu().numpy())
    rows = np.repeat(np.arange(s.n), np.diff(s.rowptr))
    gd = -lam.reshape(s.n, s.bs)[rows][:, :, None] * xd.reshape(s.n, s.bs)[s.colidx][:, None, :]
    assert np.linalg.norm(x.detach().cpu().numpy() - xd) / np.linalg.norm(xd) < 1e-10
    assert np.linalg.norm(b.grad.cpu().numpy() - lam) / np.linalg.norm(lam) < 1e-10
    np.testing.assert_allclose(vals.grad.cpu().numpy(), gd, rtol=1e-8, atol=1e-10 * np.abs(gd).max())
    ds.close()


def test_adjoint_vs_finite_differences_grid64():
    s = fa.generate(GRID_64x64x32)
    dev = torch.device("cuda:0")
    ds = DifferentiableSolver(s.bs, s.rowptr, s.colidx, s.values)
    vals = torch.tensor(s.values, device=dev, requires_grad=True)
    b = torch.tensor(s.b, device=dev, requires_grad=True)
    rng = np.random.default_rng(4)
    w = torch.tensor(rng.standard_normal(s.N), device=dev)
    x = ds(vals, b, tol=1e-13, maxit=2000, adj_tol=1e-10)
    (w * x).sum().backward()
    gv, gb = vals.grad.cpu().numpy(), b.grad.cpu().numpy()
    xb = x.detach().clone()

    g = ds.solver
    h = 1e-6
    nnzb = s.nnzb
    qs = rng.choice(nnzb, 64, replace=False)
    es = rng.integers(0, s.bs * s.bs, 64)
    bi = rng.choice(s.N, 64, replace=False)
    base_vals = torch.tensor(s.values, device=dev)

    def x_with(values_dev, b_dev):
        torch.cuda.synchronize()
        g.update_values_dev(values_dev.data_ptr())
        g.factor()
        xp = xb.clone()
        torch.cuda.synchronize()
        r = g.solve_dev(b_dev.data_ptr(), xp.data_ptr(), 1e-11, 500)
        assert r.converged and not (r.breakdown or r.nan_detected)
        return xp

    def quotient(xp, xm):
        return float(torch.dot(w, xp - xm)) / (2 * h)

    fd_v, fd_b = [], []
    bdev = torch.tensor(s.b, device=dev)
    for q, e in zip(qs, es):
        r_, c_ = divmod(int(e), s.bs)
        vp = base_vals.clone()
        vp[q, r_, c_] += h
        xp = x_with(vp, bdev)
        vp[q, r_, c_] -= 2 * h
        fd_v.append(quotient(xp, x_with(vp, bdev)))
    g.update_values_dev(base_vals.data_ptr())
    g.factor()
    for i in bi:
        bp = bdev.clone()
        bp[i] += h
        xp = x_with(base_vals, bp)
        bp[i] -= 2 * h
        fd_b.append(quotient(xp, x_with(base_vals, bp)))
    adj_v = np.array([gv[q, int(e) // s.bs, int(e) % s.bs] for q, e in zip(qs, es)])
    adj_b = gb[bi]
    fd_v, fd_b = np.array(fd_v), np.array(fd_b)
    err_v = np.linalg.norm(adj_v - fd_v) / np.linalg.norm(fd_v)
    err_b = np.linalg.norm(adj_b - fd_b) / np.linalg.norm(fd_b)
    print(f"FD check grid64x64x32: values rel err {err_v:.3e}, rhs rel err {err_b:.3e}, "
          f"max entry rel err values {np.max(np.abs(adj_v - fd_v) / np.abs(fd_v)):.3e}")
    assert err_v < 1e-6 and err_b < 1e-6
    ds.close()


def dense_grads(s, values, w):
    A = sp.bsr_matrix((values, s.colidx, s.rowptr), shape=(s.N, s.N)).toarray()
    xd = np.linalg.solve(A, s.b)
    lam = np.linalg.solve(A.T, w)
    rows = np.repeat(np.arange(s.n), np.diff(s.rowptr))
    gd = -lam.reshape(s.n, s.bs)[rows][:, :, None] * xd.reshape(s.n, s.bs)[s.colidx][:, None, :]
    return xd, lam, gd


def test_backward_after_second_forward():
    """x1 = ds(v1, b); x2 = ds(v2, b): the backward of x1 must use A(v1),
    although the handle holds A(v2) by then; likewise after a value update
    made directly on the handle between forward and backward."""
    s = fa.generate(fa.GridSpec(6, 5, 4, bs=3, seed=7, nnc_frac=0.1, nnc_dist=2))
    dev = torch.device("cuda:0")
    ds = DifferentiableSolver(s.bs, s.rowptr, s.colidx, s.values)
    rng = np.random.default_rng(3)
    w = rng.standard_normal(s.N)
    wt = torch.tensor(w, device=dev)
    vals2 = s.values * 1.25 + rng.uniform(-0.05, 0.05, s.values.shape)
    v1 = torch.tensor(s.values, device=dev, requires_grad=True)
    v2 = torch.tensor(vals2, device=dev, requires_grad=True)
    b1 = torch.tensor(s.b, device=dev, requires_grad=True)
    b2 = torch.tensor(s.b, device=dev, requires_grad=True)
    x1 = ds(v1, b1, tol=1e-13, adj_tol=1e-13)
    x2 = ds(v2, b2, tol=1e-13, adj_tol=1e-13)
    (wt * x1).sum().backward()
    _, lam1, g1 = dense_grads(s, s.values, w)
    assert np.linalg.norm(b1.grad.cpu().numpy() - lam1) / np.linalg.norm(lam1) < 1e-10
    np.testing.assert_allclose(v1.grad.cpu().numpy(), g1, rtol=1e-8, atol=1e-10 * np.abs(g1).max())
    # a value update straight on the handle before the second backward
    ds.solver.update_values(s.values * 3.0)
    ds.solver.factor()
    (wt * x2).sum().backward()
    _, lam2, g2 = dense_grads(s, vals2, w)
    assert np.linalg.norm(b2.grad.cpu().numpy() - lam2) / np.linalg.norm(lam2) < 1e-10
    np.testing.assert_allclose(v2.grad.cpu().numpy(), g2, rtol=1e-8, atol=1e-10 * np.abs(g2).max())
    ds.close()


@pytest.mark.parametrize("bs", [3, 4])
def test_autograd_precision_form_matches_fp64(bs):
    """The autograd Function in the precision form (forward and adjoint with
    FP64 refinement around FP32-value inner solves, both to 1e-10): x and
    both gradients within 1e-8 of the FP64 form's."""
    from fpga_amd.configs import GRID_128x128x64, scaled
    spec = scaled(GRID_128x128x64, 20, 18, 10) if bs == 3 else \
        fa.GridSpec(14, 12, 8, bs=4, seed=2, perm_sigma=2.0, inactive_frac=0.1, nnc_frac=0.05, nnc_dist=3)
    print(f"case {spec}", flush=True)
    s = fa.generate(spec)
    dev = torch.device("cuda:0")
    ds = DifferentiableSolver(s.bs, s.rowptr, s.colidx, s.values)
    w = torch.tensor(np.random.default_rng(3).standard_normal(s.N), device=dev)
    out = []
    for pf in (False, True):
        vals = torch.tensor(s.values, device=dev, requires_grad=True)
        b = torch.tensor(s.b, device=dev, requires_grad=True)
        x = ds(vals, b, tol=1e-10, adj_tol=1e-10, precision_form=pf)
        (w * x).sum().backward()
        out.append((x.detach().cpu().numpy(), b.grad.cpu().numpy(), vals.grad.cpu().numpy()))
    for ref, got in zip(out[0], out[1]):
        assert np.linalg.norm(got - ref) <= 1e-8 * np.linalg.norm(ref)
    ds.close()
