"""The C++ solver-backend adapter (include/bda_backend.hpp): the object a
simulator's linear-solver slot holds, over the C ABI, in the shape of the
reference's caller (the FPGA backend that links fpga_lib_alveo_u280.a,
alveo_u280/vitis_20192/OPM-integration/makefile:32,63-66, and drives
fpga_functions_bicgstab.hpp:30-137).

CPU: the header compiles on its own with -Wall -Werror, and the Newton driver
tests/cabi/backend_newton.cpp is built and linked against libfpga_amd.so.
GPU: the driver runs four solves through one backend object (upload and
colouring, a value update without new analysis, the caller's own colouring
and factors, a new pattern) and its solutions match the oracle's.
"""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import fpga_amd as fa
import oracle
from fpga_amd.configs import GRID_128x128x64, scaled

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "tests" / "cabi"))
import build_backend_test  # noqa: E402

NX, NY, NZ = 14, 12, 8


def test_header_compiles_alone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "bda_backend.hpp"\nint main() { bda_backend::Mi355xBackend be(3, 100, 1e-6); '
                   'return be.handle() != nullptr; }\n')
    res = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I",
                          str(ROOT / "include"), str(src)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_driver_builds_and_links(tmp_path):
    exe = build_backend_test.EXE
    assert exe.exists(), "tests/cabi/backend_newton is built by __graft_entry__.build()"
    res = subprocess.run(build_backend_test.command(tmp_path / "backend_newton"), capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    ldd = subprocess.run(["ldd", str(exe)], capture_output=True, text=True).stdout
    assert "libfpga_amd.so" in ldd and "not found" not in ldd


@pytest.mark.gpu
def test_newton_steps_through_the_adapter(tmp_path):
    print(f"case backend_newton {NX}x{NY}x{NZ} bs3", flush=True)
    res = subprocess.run([str(build_backend_test.EXE), str(NX), str(NY), str(NZ), "3", str(tmp_path)],
                         capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "step3 entries differing from step2: 0" in res.stdout

    spec = scaled(GRID_128x128x64, NX, NY, NZ)
    s = fa.generate(spec)
    o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, s.values)
    o.factor()
    x1 = np.fromfile(tmp_path / "x1.bin")
    xo, ro = o.solve(s.b, tol=1e-10, maxit=1000)
    assert ro["converged"] and np.linalg.norm(x1 - xo) / np.linalg.norm(xo) < 1e-8
    v2 = s.values * 1.25
    for i in range(s.n):
        for q in range(s.rowptr[i], s.rowptr[i + 1]):
            if s.colidx[q] == i:
                v2[q] += 0.5 * np.eye(s.bs)
    o.update_values(v2)
    o.factor()
    x2 = np.fromfile(tmp_path / "x2.bin")
    xo, ro = o.solve(s.b + 0.1, tol=1e-10, maxit=1000)
    assert ro["converged"] and np.linalg.norm(x2 - xo) / np.linalg.norm(xo) < 1e-8
    s4 = fa.generate(scaled(GRID_128x128x64, NX + 2, NY, NZ))
    o4 = oracle.OracleSystem(s4.bs, s4.rowptr, s4.colidx, s4.values)
    o4.factor()
    x4 = np.fromfile(tmp_path / "x4.bin")
    xo, ro = o4.solve(s4.b, tol=1e-10, maxit=1000)
    assert ro["converged"] and np.linalg.norm(x4 - xo) / np.linalg.norm(xo) < 1e-8
