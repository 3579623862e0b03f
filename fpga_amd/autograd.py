"""Differentiable ILU0-BiCGSTAB solve for PyTorch.

The backward pass is hand-written in C++ (fpga_amd/csrc/torch_adjoint.cpp, a
torch::autograd::Function): it solves A^T lambda = dL/dx with the transposed
colour-ordered ILU0 on the GPU and returns dL/db = lambda and
dL/dvalues = -lambda_i x_j^T on the block pattern.

    ds = DifferentiableSolver(bs, rowptr, colidx, values0)
    x = ds(values, b)           # values (nnzb, bs, bs), b (n*bs,): float64 on cuda
    (w * x).sum().backward()    # values.grad, b.grad
"""
from __future__ import annotations

import importlib.util

import numpy as np
import torch

from .build import TORCH_EXT, torch_ext_path
from .solver import BicgstabSolver


def _load_ext():
    path = torch_ext_path()
    if not path.exists():
        raise RuntimeError(f"{path} is not built (python -m fpga_amd.build)")
    spec = importlib.util.spec_from_file_location(TORCH_EXT, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


_ext = _load_ext()


class DifferentiableSolver:
    """Owns one solver handle (pattern uploaded, coloured once)."""

    def __init__(self, bs: int, rowptr, colidx, values0, level_scheduling: bool = False):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        n = rowptr.size - 1
        self.solver = BicgstabSolver(bs, n, int(rowptr[-1]))
        self.solver.upload_bcsr(rowptr, colidx, np.ascontiguousarray(values0, np.float64))
        self.solver.analyse(level_scheduling)
        self.bs, self.n = bs, n

    def __call__(self, values: torch.Tensor, b: torch.Tensor, tol: float = 1e-10, maxit: int = 1000,
                 adj_tol: float = 1e-10, adj_maxit: int = 1000, precision_form: bool = False) -> torch.Tensor:
        """x = A(values)^-1 b.  precision_form: forward solve and adjoint in the
        precision form (FP64 refinement around FP32-value inner solves)."""
        return _ext.solve(values, b, self.solver.handle, tol, maxit, adj_tol, adj_maxit, int(precision_form))

    def close(self):
        self.solver.close()
