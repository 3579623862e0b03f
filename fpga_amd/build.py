"""Build the in-tree shared library ``fpga_amd/libfpga_amd.so`` for gfx950.

Every source is compiled by hipcc (host C++ and HIP alike) and linked into
one library exporting the C ABI of ``include/bda_bicgstab.h``.  Objects go to
``fpga_amd/_build/``; the library is written next to this file so that it
travels with the source tree.

Flags:
  --offload-arch=gfx950          MI355X only
  -ffp-contract=off              separately rounded multiply and add, as the
                                 reference's FP units (dot_axpy.vhd:249-275)
  -target-feature -packed-fp32-ops  no packed FP32 instructions
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = PKG / "libfpga_amd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = ["grid_gen.cpp", "analysis.cpp", "host_capi.cpp", "solver.hip"]


def _compile(src: Path) -> Path:
    obj = OUT / (src.name + ".o")
    cmd = [HIPCC] + COMMON
    if src.suffix == ".hip":
        cmd += ["-x", "hip"] + HIP_FLAGS
    cmd += ["-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile all sources and link libfpga_amd.so; returns its path."""
    if not Path(HIPCC).exists() and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    OUT.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", str(tmp)] + [str(o) for o in objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
