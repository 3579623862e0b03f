"""The reference's per-colour device format ("wire format"), host side.

Mirrors the arrays the reference's host library hands to the FPGA kernel
(``vectorPointers[0..20]`` / ``vectorSizes[0..17]``, fpga_functions_bicgstab.
cpp:506-629), the five data buffers with their setup block (:84-496) and the
``dump_input_data_<b>_seq_<s>.bin`` buffer dumps (:649-665).  All work is done
by the C ABI in libfpga_amd.so (bda_wire_*); this module only allocates the
arrays and keeps them alive.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from ._lib import BdaWireSizes, lib

HEADER_WORDS = 8
LIMITS = {"rows_per_colour": 2048, "entries_per_colour": 8192, "values_per_colour": 32768,
          "colours": 256, "matrix_values": 1 << 21, "vector_elems": 163840}


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


@dataclass
class WireArrays:
    """The 21 vector pointers of the reference, as numpy arrays."""
    sizes: list            # vectorSizes[18]
    nnz_val_sizes: list    # nnzValArrays_sizes of A, L, U
    device_fits: bool
    arrays: dict = field(default_factory=dict)   # slot -> ndarray
    _keep: list = field(default_factory=list)

    def pointers(self):
        vp = (C.c_void_p * 21)()
        for slot, a in self.arrays.items():
            vp[slot] = a.ctypes.data
        # the value slots hold double** (one array each, FFB:105)
        for k in range(3):
            pp = (C.c_void_p * 1)(self.arrays[6 * k + 2].ctypes.data)
            self._keep.append(pp)
            vp[6 * k + 2] = C.cast(pp, C.c_void_p).value
        return vp

    def colour_sizes(self, k):
        """{row, arow, col, val} per partition of matrix k (0 A, 1 L, 2 U)."""
        a = self.arrays[6 * k]
        return a[HEADER_WORDS:].reshape(-1, 4)


def allocate(sizes: BdaWireSizes) -> WireArrays:
    s = list(sizes.vector_sizes)
    nv = list(sizes.nnz_val_sizes)
    w = WireArrays(s, nv, bool(sizes.device_fits))
    for k in range(3):
        b = 6 * k
        w.arrays[b + 0] = np.zeros(HEADER_WORDS + 4 * s[b + 2], np.int32)
        w.arrays[b + 1] = np.zeros(max(s[b + 3], 1), np.uint32)
        w.arrays[b + 2] = np.zeros(max(nv[k], 1), np.float64)
        w.arrays[b + 3] = np.zeros(max(s[b + 1], 1), np.uint16)
        w.arrays[b + 4] = np.zeros(max(s[b + 4], 1), np.uint8)
    w.arrays[18] = np.zeros(max(s[5], 1), np.float64)
    w.arrays[19] = np.zeros(s[0], np.float64)
    w.arrays[20] = np.zeros(s[0], np.float64)
    return w


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed")


def export_handle(solver, b=None, x0=None) -> WireArrays:
    """Arrays of a factored fpga_amd.BicgstabSolver (b, x0 in the caller's ordering)."""
    ws = BdaWireSizes()
    _check(lib.bda_wire_export_sizes(solver._h, C.byref(ws)), "bda_wire_export_sizes")
    w = allocate(ws)
    bb = None if b is None else np.ascontiguousarray(b, np.float64)
    xx = None if x0 is None else np.ascontiguousarray(x0, np.float64)
    _check(lib.bda_wire_export(solver._h, _p(bb), _p(xx), w.pointers()), "bda_wire_export")
    return w


def build(bs, colour_ptr, A, L, U, dinv, b_perm=None, x0_perm=None) -> WireArrays:
    """Arrays from colour-ordered host data: A, L, U = (rowptr, colidx, values[nnzb, bs, bs])."""
    n = len(A[0]) - 1
    cp = np.ascontiguousarray(colour_ptr, np.int32)
    A = [np.ascontiguousarray(A[0], np.int32), np.ascontiguousarray(A[1], np.int32),
         np.ascontiguousarray(A[2], np.float64)]
    L = [np.ascontiguousarray(L[0], np.int32), np.ascontiguousarray(L[1], np.int32),
         np.ascontiguousarray(L[2], np.float64)]
    U = [np.ascontiguousarray(U[0], np.int32), np.ascontiguousarray(U[1], np.int32),
         np.ascontiguousarray(U[2], np.float64)]
    dinv = np.ascontiguousarray(dinv, np.float64)
    ws = BdaWireSizes()
    _check(lib.bda_wire_build_sizes(bs, n, _p(cp), cp.size - 1, _p(A[0]), _p(A[1]), _p(L[0]), _p(L[1]),
                                    _p(U[0]), _p(U[1]), C.byref(ws)), "bda_wire_build_sizes")
    w = allocate(ws)
    bp = None if b_perm is None else np.ascontiguousarray(b_perm, np.float64)
    xp = None if x0_perm is None else np.ascontiguousarray(x0_perm, np.float64)
    _check(lib.bda_wire_build(bs, n, _p(cp), cp.size - 1, _p(A[0]), _p(A[1]), _p(A[2]), _p(L[0]), _p(L[1]),
                              _p(L[2]), _p(U[0]), _p(U[1]), _p(U[2]), _p(dinv), _p(bp), _p(xp), w.pointers()),
           "bda_wire_build")
    return w


def layout(sizes, nnz_val_sizes, config_bits=0):
    """(buffer sizes in bytes [5], setup words [40]) of fpga_setup_host_datamem."""
    s = np.ascontiguousarray(sizes, np.int32)
    nv = np.ascontiguousarray(nnz_val_sizes, np.int32)
    tot = np.zeros(5, np.uint32)
    setup = np.zeros(40, np.uint64)
    _check(lib.bda_wire_layout(_p(s), _p(nv), config_bits, _p(tot), _p(setup)), "bda_wire_layout")
    return tot, setup


def pack(w: WireArrays, config_bits=0):
    """The five data buffers of fpga_copy_host_datamem, as uint8 arrays."""
    tot, _ = layout(w.sizes, w.nnz_val_sizes, config_bits)
    bufs = [np.zeros(int(t), np.uint8) for t in tot]
    ptrs = (C.c_void_p * 5)(*[b.ctypes.data for b in bufs])
    s = np.ascontiguousarray(w.sizes, np.int32)
    nv = np.ascontiguousarray(w.nnz_val_sizes, np.int32)
    _check(lib.bda_wire_pack(_p(s), _p(nv), config_bits, w.pointers(), ptrs), "bda_wire_pack")
    return bufs


def unpack(bufs) -> WireArrays:
    """Arrays viewed inside packed buffers (sizes recovered from the setup block)."""
    bufs = [np.ascontiguousarray(b, np.uint8) for b in bufs]
    ptrs = (C.c_void_p * 5)(*[b.ctypes.data for b in bufs])
    s = np.zeros(18, np.int32)
    nv = np.zeros(3, np.int32)
    vp = (C.c_void_p * 21)()
    va = (C.c_void_p * 3)()
    _check(lib.bda_wire_unpack(ptrs, _p(s), _p(nv), vp, va), "bda_wire_unpack")
    w = WireArrays(list(s), list(nv), False)
    w._keep.extend([bufs, va])

    def view(ptr, n, dt):
        if n <= 0:
            return np.zeros(1, dt)
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), (n,)).copy()
    for k in range(3):
        b = 6 * k
        w.arrays[b + 0] = view(vp[b + 0], HEADER_WORDS + 4 * s[b + 2], np.int32)
        w.arrays[b + 0][:HEADER_WORDS] = 0
        w.arrays[b + 1] = view(vp[b + 1], s[b + 3], np.uint32)
        w.arrays[b + 2] = view(va[k], nv[k], np.float64)
        w.arrays[b + 3] = view(vp[b + 3], s[b + 1], np.uint16)
        w.arrays[b + 4] = view(vp[b + 4], s[b + 4], np.uint8)
    w.arrays[18] = view(vp[18], s[5], np.float64)
    w.arrays[19] = view(vp[19], s[0], np.float64)
    w.arrays[20] = view(vp[20], s[0], np.float64)
    return w


def import_arrays(bs, w: WireArrays):
    """Colour-ordered block CSR back from the arrays:
    (rowptr, colidx, values, lu_values, dinv, part_rows, b, x0)."""
    s = np.ascontiguousarray(w.sizes, np.int32)
    vp = w.pointers()
    n, nnzb, parts = C.c_int32(), C.c_int64(), C.c_int32()
    _check(lib.bda_wire_import_sizes(bs, _p(s), vp, C.byref(n), C.byref(nnzb), C.byref(parts)),
           "bda_wire_import_sizes")
    n, nnzb, parts = n.value, nnzb.value, parts.value
    rowptr = np.zeros(n + 1, np.int32)
    colidx = np.zeros(max(nnzb, 1), np.int32)
    vals = np.zeros((max(nnzb, 1), bs, bs))
    lu = np.zeros((max(nnzb, 1), bs, bs))
    dinv = np.zeros((n, bs, bs))
    part_rows = np.zeros(parts + 1, np.int32)
    b = np.zeros(n * bs)
    x0 = np.zeros(n * bs)
    _check(lib.bda_wire_import(bs, _p(s), vp, _p(rowptr), _p(colidx), _p(vals), _p(lu), _p(dinv), _p(part_rows),
                               _p(b), _p(x0)), "bda_wire_import")
    return rowptr, colidx[:nnzb], vals[:nnzb], lu[:nnzb], dinv, part_rows, b, x0


def write_dump(bufs, directory, sequence=0):
    """dump_input_data_<b>_seq_<sequence>.bin files, one per data buffer."""
    d = Path(directory)
    paths = []
    for i, buf in enumerate(bufs):
        p = d / f"dump_input_data_{i}_seq_{sequence}.bin"
        p.write_bytes(np.asarray(buf, np.uint8).tobytes())
        paths.append(p)
    return paths


def read_dump(directory, sequence=0):
    d = Path(directory)
    return [np.frombuffer((d / f"dump_input_data_{i}_seq_{sequence}.bin").read_bytes(), np.uint8).copy()
            for i in range(5)]
