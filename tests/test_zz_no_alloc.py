"""bda_spmv_dev_noalloc, the entry point that allocates nothing, against
bda_spmv_dev, whose reductions use the handle's own buffers.

Both run the same SpMV kernels with the same fixed-order sums, so the
outputs and the fused dots must agree bit for bit, in every kernel form, on
two systems in a row (the second smaller: the shared scratch array keeps the
first call's contents).  The entry point must refuse what would need an
allocation (FP32 or transposed copies not yet made) and systems beyond its
scratch capacity, with a message, and the device's free memory must not
change while it runs.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import fpga_amd as fa
from fpga_amd._lib import lib
from fpga_amd.configs import GRID_128x128x64, GRID_256x256x64, scaled

pytestmark = pytest.mark.gpu

DT = {0: torch.float64, 1: torch.float32, 2: torch.float64}


def _hip():
    """The HIP runtime already in the process (the library's and torch's,
    one soname), to read the device's free memory."""
    try:
        h = C.CDLL("libamdhip64.so.7")
    except OSError:
        return None
    h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return h


def free_bytes(hip):
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def call(fn, g, prec, x, y, w, red, transpose=0):
    dots = (C.c_double * 2)()
    rc = fn(C.c_void_p(g.handle), prec, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), transpose,
            C.c_void_p(w.data_ptr()) if w is not None else None, red, dots)
    return rc, (dots[0], dots[1])


def make(spec):
    print(f"case {spec}", flush=True)
    s = fa.generate(spec)
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    g.upload_system(s)
    g.analyse()
    g.factor()
    return s, g


def test_noalloc_equals_spmv_dev_two_calls():
    dev = torch.device("cuda:0")
    hip = _hip()
    systems = [make(scaled(GRID_128x128x64, 40, 36, 20)), make(scaled(GRID_128x128x64, 14, 12, 8))]
    assert systems[1][0].n < systems[0][0].n
    rng = np.random.default_rng(31)
    try:
        # what the entry point cannot make itself: refused, with a message
        s0, g0 = systems[0]
        x = torch.zeros(s0.N, dtype=torch.float32, device=dev)
        rc, _ = call(lib.bda_spmv_dev_noalloc, g0, 1, x, torch.empty_like(x), None, 0)
        assert rc != 0 and b"not current" in lib.bda_last_error(C.c_void_p(g0.handle))
        # inputs and the regular entry point's runs (these may allocate: the
        # FP32 and transposed copies of the matrix, torch's tensors)
        cases = []
        for s, g in systems:
            for prec in (0, 1, 2):
                x = torch.tensor(rng.uniform(-1, 1, s.N), dtype=DT[prec], device=dev)
                w = torch.tensor(rng.uniform(-1, 1, s.N), dtype=DT[prec], device=dev)
                for red, tr in ((0, 0), (1, 0), (3, 0), (0, 1)):
                    y_ref = torch.empty_like(x)
                    torch.cuda.synchronize()
                    rc, d_ref = call(lib.bda_spmv_dev, g, prec, x, y_ref, w, red, tr)
                    assert rc == 0, lib.bda_last_error(C.c_void_p(g.handle))
                    cases.append((s, g, prec, red, tr, x, w, y_ref, d_ref, torch.full_like(x, float("nan"))))
        torch.cuda.synchronize()
        before = free_bytes(hip) if hip else None
        # the no-allocation runs: the larger system first, then the smaller
        for s, g, prec, red, tr, x, w, y_ref, d_ref, y in cases:
            rc, d = call(lib.bda_spmv_dev_noalloc, g, prec, x, y, w, red, tr)
            assert rc == 0, lib.bda_last_error(C.c_void_p(g.handle))
        torch.cuda.synchronize()
        if hip:
            assert free_bytes(hip) == before
        for s, g, prec, red, tr, x, w, y_ref, d_ref, y in cases:
            assert torch.equal(y, y_ref), (s.n, prec, red, tr)
            rc, d = call(lib.bda_spmv_dev_noalloc, g, prec, x, y, w, red, tr)
            assert rc == 0 and d == d_ref, (s.n, prec, red, tr, d, d_ref)
    finally:
        for _, g in systems:
            g.close()


def test_noalloc_rejects_beyond_scratch():
    """256x256x64 with 4x4 blocks: its SpMV has more tiles than the scratch
    array holds partial sums; a reduction call is refused, a plain product
    (no reductions, no scratch) runs and equals bda_spmv_dev."""
    dev = torch.device("cuda:0")
    s, g = make(GRID_256x256x64)
    try:
        x = torch.tensor(np.random.default_rng(32).uniform(-1, 1, s.N), device=dev)
        y = torch.empty_like(x)
        rc, _ = call(lib.bda_spmv_dev_noalloc, g, 0, x, y, x, 1)
        assert rc != 0 and b"at most 65504" in lib.bda_last_error(C.c_void_p(g.handle))
        y_ref = torch.empty_like(x)
        assert call(lib.bda_spmv_dev, g, 0, x, y_ref, None, 0)[0] == 0
        assert call(lib.bda_spmv_dev_noalloc, g, 0, x, y, None, 0)[0] == 0
        assert torch.equal(y, y_ref)
    finally:
        g.close()
