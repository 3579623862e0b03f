"""Build step of the backend test program: tests/cabi/backend_newton, a plain
C++ program compiled with g++ against include/bda_backend.hpp and
fpga_amd/libfpga_amd.so (rpath to the library's directory)."""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
SRC = HERE / "backend_newton.cpp"
EXE = HERE / "backend_newton"


def command(out: Path = EXE):
    return ["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(SRC), "-o", str(out),
            "-L", str(ROOT / "fpga_amd"), "-lfpga_amd", "-Wl,-rpath,$ORIGIN/../../fpga_amd"]


def build(verbose: bool = False) -> Path:
    tmp = EXE.with_suffix(".tmp")
    res = subprocess.run(command(tmp), capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"backend test program build failed:\n{res.stderr}")
    tmp.replace(EXE)
    if verbose:
        print(f"built {EXE}")
    return EXE


if __name__ == "__main__":
    build(verbose=True)
