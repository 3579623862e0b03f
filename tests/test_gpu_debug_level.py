"""Diagnostic lines of the library on stderr at levels 0, 1 and 2.

The counterpart of the reference's BDA_DEBUG printouts
(src/sda_app/common/bda_utils.hpp:46-57), e.g. its per-colour sizes
(fpga_functions_bicgstab.cpp:146-167): the lines must state what the
handle reports through its own queries (colours, launches, iterations).
"""
import re

import numpy as np
import pytest

import fpga_amd as fa
from fpga_amd.configs import GRID_128x128x64, scaled

pytestmark = pytest.mark.gpu


def test_debug_lines(capfd):
    spec = scaled(GRID_128x128x64, 14, 12, 8)
    print(f"case {spec}", flush=True)
    s = fa.generate(spec)
    prev = fa.set_debug_level(2)
    try:
        capfd.readouterr()
        g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
        g.upload_system(s)
        g.analyse()
        g.factor()
        x, r = g.solve(s.b, tol=1e-8, maxit=500)
        xm, rm = g.solve_mixed(s.b, tol=1e-6)
        err = capfd.readouterr().err
        assert re.search(rf"INFO: bda_create: 3x3 blocks, {s.n} block rows", err), err
        m = re.search(r"INFO: analyse: (\d+) block rows of 3x3, (\d+) blocks, (\d+) colours \(greedy\)", err)
        assert m and int(m.group(1)) == s.n and int(m.group(2)) == s.nnzb
        nc = int(m.group(3))
        _, _, cptr = g.colouring()
        assert nc == len(cptr) - 1
        rows = [int(v) for v in re.findall(r"INFO: analyse: colour \d+: (\d+) rows", err)]
        assert rows == list(np.diff(cptr))
        m = re.search(r"preconditioner launches per application: (\d+), transposed (\d+)", err)
        assert m and int(m.group(1)) == g.precond_launches(False) and int(m.group(2)) == g.precond_launches(True)
        assert "INFO: factor: block ILU0" in err
        m = re.search(r"INFO: solve: (\d+) half-iterations", err)
        assert m and int(m.group(1)) == r.half_iterations
        m = re.search(r"INFO: precision-form solve: (\d+) half-iterations \((\d+) refinement steps\)", err)
        assert m and int(m.group(2)) == rm.outer_iterations
        assert "INFO: solve: ||r0||" in err

        # an error is printed at level 1 (and still reported through the return code)
        fa.set_debug_level(1)
        with pytest.raises(fa.BdaError):
            g.solve(s.b, tol=-1.0)
        assert "ERROR: " in capfd.readouterr().err

        # level 0: silent
        fa.set_debug_level(0)
        g.solve(s.b, tol=1e-8, maxit=500)
        assert "INFO:" not in capfd.readouterr().err
        g.close()
    finally:
        fa.set_debug_level(prev)
