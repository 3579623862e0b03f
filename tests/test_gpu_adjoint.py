"""The hand-written adjoint behind the LibTorch autograd Function (GPU).

- dense check: dL/db = A^{-T} w and dL/dA = -lambda x^T on the block pattern,
  against numpy on small systems;
- central finite differences (h = 1e-6, FP64) on 64 random matrix entries and
  64 random right-hand-side entries of the 64x64x32 grid, L = w^T x with
  w ~ N(0,1), adjoint solved to 1e-10; required agreement 1e-6 relative.
  Each perturbed solve restarts from the unperturbed solution, so its
  tolerance is relative to the O(h) change of the residual.
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

    def loss_with(values_dev, b_dev):
        torch.cuda.synchronize()
        g.update_values_dev(values_dev.data_ptr())
        g.factor()
        xp = xb.clone()
        torch.cuda.synchronize()
        r = g.solve_dev(b_dev.data_ptr(), xp.data_ptr(), 1e-9, 2000)
        assert not (r.breakdown or r.nan_detected)
        return float(torch.dot(w, xp))

    fd_v, fd_b = [], []
    bdev = torch.tensor(s.b, device=dev)
    for q, e in zip(qs, es):
        r_, c_ = divmod(int(e), s.bs)
        vp = base_vals.clone()
        vp[q, r_, c_] += h
        lp = loss_with(vp, bdev)
        vp[q, r_, c_] -= 2 * h
        lm = loss_with(vp, bdev)
        fd_v.append((lp - lm) / (2 * h))
    g.update_values_dev(base_vals.data_ptr())
    g.factor()
    for i in bi:
        bp = bdev.clone()
        bp[i] += h
        lp = loss_with(base_vals, bp)
        bp[i] -= 2 * h
        lm = loss_with(base_vals, bp)
        fd_b.append((lp - lm) / (2 * h))
    adj_v = np.array([gv[q, int(e) // s.bs, int(e) % s.bs] for q, e in zip(qs, es)])
    adj_b = gb[bi]
    fd_v, fd_b = np.array(fd_v), np.array(fd_b)
    err_v = np.linalg.norm(adj_v - fd_v) / np.linalg.norm(fd_v)
    err_b = np.linalg.norm(adj_b - fd_b) / np.linalg.norm(fd_b)
    print(f"FD check grid64x64x32: values rel err {err_v:.3e}, rhs rel err {err_b:.3e}, "
          f"max entry rel err values {np.max(np.abs(adj_v - fd_v) / np.abs(fd_v)):.3e}")
    assert err_v < 1e-6 and err_b < 1e-6
    ds.close()
