"""Write the golden fixtures of tests/golden/ from numpy / scipy routines.

Nothing here calls the oracle or the solver library's arithmetic: the block
systems come from the synthetic generator (fpga_amd.generate, host code),
and every expected output is computed with numpy/scipy or a closed form:

  sys_*.npz         y = A @ x with scipy.sparse.bsr_matrix; x = numpy.linalg.solve(A, b);
                    lambda = A^{-T} w; the gradient -lambda_i x_j^T on the block pattern
  redblack_*.npz    block ILU0 of a 2-colour 7-point grid in closed form
                    (red rows first): D_r = A_rr, U_rb = A_rb,
                    L_br = A_br A_rr^-1, D_b = A_bb - sum_r A_br A_rr^-1 A_rb,
                    and the right-preconditioned BiCGSTAB iterates of
                    scipy.sparse.linalg.bicgstab with M = (L U)^-1
  colouring.json    SHA-256 of the closed-form colouring (i+j+k) mod 2 and of
                    its stable colour-ordered permutation for the 32x32x16 and
                    64x64x32 grids, and of the 32x32x16 generated matrix

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
OUT = Path(__file__).resolve().parent

import fpga_amd as fa  # noqa: E402  (host generator only)

SMALL = {
    "g4x3x2": fa.GridSpec(4, 3, 2, bs=3, seed=11, perm_sigma=1.0),
    "g5x4x3_nnc": fa.GridSpec(5, 4, 3, bs=3, seed=7, perm_sigma=2.0, layer_corr=0.5, inactive_frac=0.1,
                              nnc_frac=0.2, nnc_dist=3),
    "g3x3x2_bs4": fa.GridSpec(3, 3, 2, bs=4, seed=5, perm_sigma=1.0, nnc_frac=0.3, nnc_dist=2),
}
REDBLACK = {
    "rb6x5x4": fa.GridSpec(6, 5, 4, bs=3, seed=3, perm_sigma=1.0),
    "rb4x4x3_bs4": fa.GridSpec(4, 4, 3, bs=4, seed=9, perm_sigma=1.0),
}


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.int32).tobytes()).hexdigest()


def dense(s: fa.BlockSystem) -> np.ndarray:
    return sp.bsr_matrix((s.values, s.colidx, s.rowptr), shape=(s.N, s.N)).toarray()


def closed_form_colouring(nx, ny, nz):
    i = np.arange(nx * ny * nz)
    colour = ((i % nx) + (i // nx) % ny + i // (nx * ny)) % 2
    perm = np.concatenate([np.nonzero(colour == 0)[0], np.nonzero(colour == 1)[0]])
    return colour.astype(np.int32), perm.astype(np.int32)


def redblack_fixture(s: fa.BlockSystem, grid):
    bs, n = s.bs, s.n
    A = dense(s)
    colour, spec_perm = closed_form_colouring(*grid)
    P = np.concatenate([np.arange(k * bs, (k + 1) * bs) for k in spec_perm])
    Ap = A[np.ix_(P, P)]
    nred = int(np.sum(colour == 0))
    Lf = np.eye(s.N)
    Uf = np.zeros((s.N, s.N))
    blk = lambda M, i, j: M[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs]
    dinv = np.zeros((n, bs, bs))
    for i in range(n):
        for j in range(n):
            a = blk(Ap, i, j)
            if i == j or not np.any(a):
                continue
            if i < nred:                       # red row: U_rb = A_rb
                Uf[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] = a
            else:                              # black row: L_br = A_br A_rr^-1
                Lf[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] = a @ np.linalg.inv(blk(Ap, j, j))
    for i in range(n):
        d = blk(Ap, i, i).copy()
        if i >= nred:
            for j in range(nred):
                a = blk(Ap, i, j)
                if np.any(a):
                    d -= a @ np.linalg.inv(blk(Ap, j, j)) @ blk(Ap, j, i)
        Uf[i * bs:(i + 1) * bs, i * bs:(i + 1) * bs] = d
        dinv[i] = np.linalg.inv(d)
    # right-preconditioned BiCGSTAB with M = L U (scipy's step structure)
    bp = s.b[P]

    def psolve(v):
        # L has identity diagonal blocks (scalar unit-lower); U is block upper
        # triangular with full diagonal blocks, so it takes a general solve
        y = sla.solve_triangular(Lf, v, lower=True, unit_diagonal=True)
        return np.linalg.solve(Uf, y)

    M = spla.LinearOperator(Ap.shape, matvec=psolve)
    iters = []
    x_s, info = spla.bicgstab(sp.csr_matrix(Ap), bp, rtol=1e-8, atol=0.0, maxiter=200, M=M,
                              callback=lambda xk: iters.append(1))
    rng = np.random.default_rng(1)
    pvec = rng.uniform(-1, 1, s.N)
    return dict(perm=spec_perm, L=Lf, U=Uf, dinv=dinv, z=psolve(pvec), p=pvec, x_bicgstab=x_s,
                scipy_iters=len(iters), scipy_info=info, x_dense=np.linalg.solve(Ap, bp))



def main():
    hashes = {}
    for name, spec in SMALL.items():
        s = fa.generate(spec)
        A = dense(s)
        rng = np.random.default_rng(2)
        x = rng.uniform(-1, 1, s.N)
        w = rng.standard_normal(s.N)
        xs = np.linalg.solve(A, s.b)
        lam = np.linalg.solve(A.T, w)
        rows = np.repeat(np.arange(s.n), np.diff(s.rowptr))
        grad = -lam.reshape(s.n, s.bs)[rows][:, :, None] * xs.reshape(s.n, s.bs)[s.colidx][:, None, :]
        np.savez_compressed(OUT / f"sys_{name}.npz", rowptr=s.rowptr, colidx=s.colidx, values=s.values, b=s.b,
                            bs=s.bs, x=x,
                            y=sp.bsr_matrix((s.values, s.colidx, s.rowptr), shape=(s.N, s.N)) @ x,
                            x_solve=xs, w=w, lam=lam, grad=grad)
    for name, spec in REDBLACK.items():
        s = fa.generate(spec)
        fx = redblack_fixture(s, (spec.nx, spec.ny, spec.nz))
        np.savez_compressed(OUT / f"redblack_{name}.npz", rowptr=s.rowptr, colidx=s.colidx, values=s.values,
                            b=s.b, bs=s.bs, grid=np.array([spec.nx, spec.ny, spec.nz]), **fx)
    for nx, ny, nz in [(32, 32, 16), (64, 64, 32)]:
        c, p = closed_form_colouring(nx, ny, nz)
        hashes[f"grid{nx}x{ny}x{nz}"] = {"colour_sha256": sha(c), "perm_sha256": sha(p),
                                          "colour_ptr": [0, int(np.sum(c == 0)), int(c.size)]}
    s = fa.generate(fa.configs.GRID_32x32x16)
    hashes["grid32x32x16"]["matrix_sha256"] = hashlib.sha256(
        s.rowptr.tobytes() + s.colidx.tobytes() + s.values.tobytes() + s.b.tobytes()).hexdigest()
    hashes["grid32x32x16"]["nnzb"] = int(s.nnzb)
    (OUT / "colouring.json").write_text(json.dumps(hashes, indent=2) + "\n")
    print("wrote", sorted(p.name for p in OUT.iterdir() if p.suffix in (".npz", ".json")))


if __name__ == "__main__":
    main()
