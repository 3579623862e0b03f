"""The five synthetic workloads of this solver, by what they are.

grid32x32x16   32x32x16 cells, 3x3 blocks, lognormal(0,1) perm, rel tol 1e-2,
               maxit 200 -- the CPU-runnable parity case; also the matrix of
               the precision-form case.
grid128x128x64 128x128x64 cells, 3x3 blocks, lognormal(0,2) perm with layer
               correlation, 10% cells inactive, 1% non-neighbour connections
               within Chebyshev distance 3, rel tol 1e-6, maxit 1000 -- the
               headline single-GPU solve.
grid64x64x32   131,072 cells, generator of grid32x32x16 -- the adjoint case
               (backward tolerance 1e-10).
grid256x256x64 4x4 blocks, otherwise as grid128x128x64 -- the large combined
               run.

Layer correlation: log K = sigma (sqrt(0.5) z_layer + sqrt(0.5) z_cell), so
half of the log-permeability variance is shared by all cells of a z layer.
"""
from __future__ import annotations

from .solver import GridSpec

GRID_32x32x16 = GridSpec(32, 32, 16, bs=3, seed=0, perm_sigma=1.0)
GRID_128x128x64 = GridSpec(128, 128, 64, bs=3, seed=0, perm_sigma=2.0, layer_corr=0.5,
                           inactive_frac=0.10, nnc_frac=0.01, nnc_dist=3)
GRID_64x64x32 = GridSpec(64, 64, 32, bs=3, seed=0, perm_sigma=1.0)
GRID_256x256x64 = GridSpec(256, 256, 64, bs=4, seed=0, perm_sigma=2.0, layer_corr=0.5,
                           inactive_frac=0.10, nnc_frac=0.01, nnc_dist=3)

WORKLOADS = {
    "grid32x32x16": dict(spec=GRID_32x32x16, tol=1e-2, maxit=200),
    "grid128x128x64": dict(spec=GRID_128x128x64, tol=1e-6, maxit=1000),
    "grid64x64x32": dict(spec=GRID_64x64x32, tol=1e-10, maxit=1000),
    "grid256x256x64": dict(spec=GRID_256x256x64, tol=1e-6, maxit=1000),
}


def scaled(spec: GridSpec, nx: int, ny: int, nz: int) -> GridSpec:
    """Same generator settings on another grid size (for quick tests)."""
    d = dict(spec.__dict__)
    d.update(nx=nx, ny=ny, nz=nz)
    return GridSpec(**d)
