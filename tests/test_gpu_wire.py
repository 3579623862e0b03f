"""Wire-format export of a GPU handle (factors read back from the device)
against the same export built from the oracle's factors: identical arrays,
since the device factors match the oracle's bit for bit."""
import numpy as np
import pytest

import fpga_amd as fa
import oracle
from fpga_amd import wire
from fpga_amd.configs import GRID_32x32x16, scaled

from test_wire_format import permuted

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spec", [scaled(GRID_32x32x16, 12, 10, 8),
                                  fa.GridSpec(8, 7, 5, bs=4, seed=4, perm_sigma=2.0, nnc_frac=0.1, nnc_dist=2)],
                         ids=["bs3", "bs4"])
def test_export_matches_oracle_build(spec):
    print(f"case {spec}", flush=True)
    s = fa.generate(spec)
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    g.upload_system(s)
    g.analyse()
    g.factor()
    rng = np.random.default_rng(0)
    x0 = rng.standard_normal(s.N)
    wg = wire.export_handle(g, s.b, x0)
    g.close()

    o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, s.values)
    o.factor()
    _, perm, cptr = o.colouring()
    L, U, dinv = o.factors()
    sp = lambda v: v.reshape(s.n, s.bs)[perm].ravel()
    wo = wire.build(s.bs, cptr, permuted(s, perm), L, U, dinv, sp(s.b), sp(x0))
    assert wg.sizes == wo.sizes and wg.nnz_val_sizes == wo.nnz_val_sizes
    for slot, a in wo.arrays.items():
        assert np.array_equal(wg.arrays[slot].view(np.uint8), a.view(np.uint8)), slot
