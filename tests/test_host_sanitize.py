"""The host-only C-ABI paths under AddressSanitizer and UBSan (CPU only).

tests/cabi/host_sanitize.cpp is compiled with g++ together with the
library's host sources (grid generator, colouring and analysis, the
per-colour device format, the status/limits encoders; no device code) and
run: grid generation, colouring of 3x3 and 4x4 systems with fault
connections, build -> import and pack -> unpack round trips of the
reference's per-colour format (fpga_functions_bicgstab.cpp:506-629), and
malformed patterns and streams that must be rejected.  Any out-of-bounds
access, use after free, leak or undefined behaviour aborts the program with
a sanitizer report, which fails the test.
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "fpga_amd" / "csrc"
SOURCES = [ROOT / "tests" / "cabi" / "host_sanitize.cpp"] + [
    CSRC / f for f in ("analysis.cpp", "grid_gen.cpp", "wire_format.cpp", "host_capi.cpp")]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_host_paths_under_sanitizers(tmp_path):
    exe = tmp_path / "host_sanitize"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-Wall", "-I", str(ROOT / "include")] + [str(s) for s in SOURCES] + [
           "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-6000:]
    assert "0 failed checks" in r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
