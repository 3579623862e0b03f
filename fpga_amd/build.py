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
SOURCES = ["grid_gen.cpp", "analysis.cpp", "host_capi.cpp", "wire_format.cpp", "solver.hip"]
TORCH_SRC = "torch_adjoint.cpp"
TORCH_EXT = "_torch_bda"


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


def _torch_flags():
    import sysconfig

    import torch
    import torch.utils.cpp_extension as ce
    inc = sum((["-I", p] for p in ce.include_paths()), []) + ["-I", sysconfig.get_paths()["include"]]
    libs = sum((["-L", p] for p in ce.library_paths()), []) + ["-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python"]
    defs = [f"-D_GLIBCXX_USE_CXX11_ABI={int(torch._C._GLIBCXX_USE_CXX11_ABI)}",
            f"-DTORCH_EXTENSION_NAME={TORCH_EXT}", "-DTORCH_API_INCLUDE_EXTENSION_H"]
    return inc, libs, defs, sysconfig.get_config_var("EXT_SUFFIX")


def _compile_torch() -> Path:
    inc, _, defs, _ = _torch_flags()
    obj = OUT / (TORCH_SRC + ".o")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-w"] + defs + inc + ["-c", str(CSRC / TORCH_SRC), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def torch_ext_path() -> Path:
    import sysconfig
    return PKG / (TORCH_EXT + sysconfig.get_config_var("EXT_SUFFIX"))


def build(verbose: bool = False) -> Path:
    """Compile all sources, link libfpga_amd.so and the LibTorch extension
    (the autograd Function of torch_adjoint.cpp, linked against the library);
    returns the library path."""
    if not Path(HIPCC).exists() and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    OUT.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs) + 1)) as ex:
        tfut = ex.submit(_compile_torch)
        objs = list(ex.map(_compile, srcs))
        tobj = tfut.result()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", str(tmp)] + [str(o) for o in objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    _, libs, _, _ = _torch_flags()
    ext = torch_ext_path()
    tmp = ext.with_suffix(".tmp")
    cmd = ["g++", "-shared", "-fPIC", "-o", str(tmp), str(tobj), "-L", str(PKG), "-lfpga_amd",
           "-Wl,-rpath,$ORIGIN"] + libs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, ext)
    if verbose:
        print(f"built {LIB}")
        print(f"built {ext}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
