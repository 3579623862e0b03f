"""Status and limits lines decoded by the reference's own host decoder.

bda_encode_status / bda_encode_query write the layout of the reference
kernel's debug port; decode_debuginfo_bicgstab and
decode_debuginfo_bicgstab_query (sda_app/common/bicgstab_utils.cpp, built from
the reference sources into oracle/_ref by oracle/ref.mk) must read back the
same flags, half-iteration count, norms, newest-norm index and limits.
"""
import ctypes as C

import numpy as np
import pytest

import fpga_amd as fa
import oracle
from fpga_amd._lib import BdaLimits, BdaResult, lib

dec = oracle.ref_status_decoder()
pytestmark = pytest.mark.skipif(dec is None, reason="reference status decoder not built (needs the reference sources)")


def decode_status(line):
    buf = (C.c_ulong * 16)(*[int(v) for v in line])
    cycles, iters = C.c_uint(), C.c_uint()
    norms = (C.c_double * 4)()
    last = C.c_ubyte()
    flags = [C.c_bool() for _ in range(6)]
    rc = dec[0](C.c_bool(True), C.c_bool(False), buf, C.c_uint(2), C.c_uint(8), C.c_uint(0), C.byref(cycles),
                C.byref(iters), norms, C.byref(last), *[C.byref(f) for f in flags])
    aborted, signature_bad, overflow, noresults, wrafter, fifofull = [f.value for f in flags]
    return dict(rc=rc, cycles=cycles.value, iters=iters.value, norms=list(norms), last=last.value,
                aborted=aborted, signature_bad=signature_bad, overflow=overflow, noresults=noresults)


def encode(res: BdaResult):
    line = (C.c_uint64 * 16)()
    assert lib.bda_encode_status(C.byref(res), line) == 0
    return list(line)


@pytest.mark.parametrize("h", [0, 1, 2, 5, 66, 1000])
def test_status_roundtrip(h):
    r = BdaResult()
    r.half_iterations = h
    r.iterations = h / 2 + 0.5
    for i, v in enumerate([128.5, 2.25, 1e-3, 7.5e-7]):
        r.norms[i] = v
    r.last_norm_idx = (h % 3) + 1
    r.solve_ms = 12.5
    d = decode_status(encode(r))
    assert d["rc"] == 0 and not d["signature_bad"] and not d["overflow"] and not d["aborted"]
    assert d["iters"] == h and d["norms"] == [128.5, 2.25, 1e-3, 7.5e-7]
    assert d["last"] == r.last_norm_idx
    assert d["cycles"] == round(12.5e-3 * 250e6)


def test_status_flags():
    r = BdaResult()
    r.no_results = 1
    d = decode_status(encode(r))
    assert d["noresults"] and not d["aborted"] and d["rc"] == 0
    r.aborted = 1
    d = decode_status(encode(r))
    assert d["aborted"] and d["rc"] == 1         # the reference treats an abort as failure


@pytest.mark.parametrize("tol,maxit", [(1e-2, 200), (1e-8, 200), (0.0, 3), (0.0, 4)])
def test_oracle_solve_status_decodes(tol, maxit):
    """The norm ring of a real solve: the reference decoder's newest-norm
    index (h % 3) + 1 points at the last norm the solver produced."""
    s = fa.generate(fa.configs.scaled(fa.configs.GRID_32x32x16, 10, 8, 6))
    o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, s.values)
    o.factor()
    x, res = o.solve(s.b, tol=tol, maxit=maxit)
    r = BdaResult()
    r.half_iterations = res["half_iterations"]
    for i in range(4):
        r.norms[i] = res["norms"][i]
    d = decode_status(encode(r))
    assert d["iters"] == res["half_iterations"]
    assert d["last"] == res["last_norm_idx"]
    assert d["norms"][d["last"]] == res["final_norm"]


def test_query_roundtrip():
    lim = BdaLimits(4, 128, 32767, 64, 288 << 30, 256, 64)
    line = (C.c_uint64 * 8)()
    assert lib.bda_encode_query(C.byref(lim), 943550, 6003904, 3, line) == 0
    buf = (C.c_ulong * 8)(*list(line))
    u = [C.c_uint() for _ in range(4)]
    s16 = [C.c_ushort() for _ in range(4)]
    b = [C.c_bool() for _ in range(2)]
    u8 = [C.c_ubyte() for _ in range(6)]
    # argument order: x_vector_elem, max_row_size, max_column_size, max_colors_size, max_nnzs_per_row,
    # max_matrix_size, use_uram, write_ilu0_results, dma_data_width, mult_num, x_vector_latency,
    # add_latency, mult_latency, num_read_ports, num_write_ports, reset_cycles, reset_settle
    max_matrix = C.c_uint()
    rc = dec[1](C.c_bool(True), buf, C.c_uint(1), C.c_uint(8), C.byref(u[0]), C.byref(u[1]), C.byref(u[2]),
                C.byref(u[3]), C.byref(s16[0]), C.byref(max_matrix), C.byref(b[0]), C.byref(b[1]),
                C.byref(s16[1]), C.byref(u8[0]), C.byref(u8[1]), C.byref(u8[2]), C.byref(u8[3]), C.byref(u8[4]),
                C.byref(u8[5]), C.byref(s16[2]), C.byref(s16[3]))
    assert rc == 0
    assert [v.value for v in u] == [943550 * 3, 943550 * 3, 943550 * 3, 943550]
    assert s16[0].value == 128 * 3 and max_matrix.value == 6003904 * 9
    assert (b[0].value, b[1].value, s16[1].value, u8[0].value) == (False, True, 512, 64)
    assert (u8[4].value, u8[5].value) == (1, 1)


def decode_query(line):
    buf = (C.c_ulong * 8)(*[int(v) for v in line])
    u = [C.c_uint() for _ in range(4)]
    s16 = [C.c_ushort() for _ in range(4)]
    b = [C.c_bool() for _ in range(2)]
    u8 = [C.c_ubyte() for _ in range(6)]
    max_matrix = C.c_uint()
    rc = dec[1](C.c_bool(False), buf, C.c_uint(1), C.c_uint(8), C.byref(u[0]), C.byref(u[1]), C.byref(u[2]),
                C.byref(u[3]), C.byref(s16[0]), C.byref(max_matrix), C.byref(b[0]), C.byref(b[1]),
                C.byref(s16[1]), C.byref(u8[0]), C.byref(u8[1]), C.byref(u8[2]), C.byref(u8[3]), C.byref(u8[4]),
                C.byref(u8[5]), C.byref(s16[2]), C.byref(s16[3]))
    return dict(rc=rc, x_elems=u[0].value, max_rows=u[1].value, max_cols=u[2].value, max_colours=u[3].value,
                nnz_per_row=s16[0].value, max_matrix=max_matrix.value, uram=b[0].value, ilu0_out=b[1].value,
                width=s16[1].value, mults=u8[0].value, rd_ports=u8[4].value, wr_ports=u8[5].value)


@pytest.mark.gpu
def test_live_query_decodes():
    """bda_query on a live handle (fpga_kernel_query, fpga_functions_bicgstab.cpp:1332-1452), its limits
    line decoded by the reference's decode_debuginfo_bicgstab_query (bicgstab_utils.cpp:282-369)."""
    s = fa.generate(fa.configs.scaled(fa.configs.GRID_32x32x16, 10, 8, 6))
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    try:
        lim = g.query()
        print("live limits", {k: getattr(lim, k) for k, _ in lim._fields_}, flush=True)
        assert lim.max_block_size == 4 and lim.max_iterations == 32767
        assert lim.wavefront_size == 64 and lim.compute_units >= 200
        assert lim.device_memory_bytes > (200 << 30)
        assert lim.max_blocks_per_row == 168 and lim.rows_per_tile == 64
        line = (C.c_uint64 * 8)()
        assert lib.bda_encode_query(C.byref(lim), s.n, s.nnzb, s.bs, line) == 0
        d = decode_query(list(line))
        assert d["rc"] == 0
        assert (d["x_elems"], d["max_rows"], d["max_cols"], d["max_colours"]) == (s.N, s.N, s.N, s.n)
        assert d["nnz_per_row"] == 168 * 3 and d["max_matrix"] == s.nnzb * 9
        assert (d["uram"], d["ilu0_out"], d["width"], d["mults"]) == (False, True, 512, 64)
        assert (d["rd_ports"], d["wr_ports"]) == (1, 1)
    finally:
        g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("tol,maxit", [(1e-2, 200), (1e-8, 200), (0.0, 3), (0.0, 4)])
def test_gpu_solve_status_decodes(tol, maxit):
    """A real GPU solve's result encoded as the kernel's status line and read
    back by decode_debuginfo_bicgstab (bicgstab_utils.cpp:90-280): half
    iterations, the four norms, the newest-norm index and the flags."""
    s = fa.generate(fa.configs.scaled(fa.configs.GRID_32x32x16, 10, 8, 6))
    g = fa.BicgstabSolver(s.bs, s.n, s.nnzb)
    try:
        g.upload_system(s)
        g.analyse()
        g.factor()
        r = BdaResult()
        x = np.zeros(s.N)
        assert lib.bda_solve(C.c_void_p(g.handle), s.b.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p),
                             tol, maxit, C.byref(r)) == 0
        d = decode_status(encode(r))
        assert d["rc"] == 0 and not d["signature_bad"] and not d["aborted"] and not d["noresults"]
        assert d["iters"] == r.half_iterations and d["last"] == r.last_norm_idx
        assert d["norms"] == list(r.norms) and d["norms"][d["last"]] == r.final_norm
        assert abs(d["cycles"] - r.solve_ms * 1e-3 * 250e6) <= 0.5  # cycles of the 250 MHz kernel clock
        # no results: a zero right-hand side and start vector
        z = np.zeros(s.N)
        x = np.zeros(s.N)
        assert lib.bda_solve(C.c_void_p(g.handle), z.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p),
                             tol, maxit, C.byref(r)) == 0
        assert r.no_results
        d = decode_status(encode(r))
        assert d["noresults"] and d["iters"] == 0
    finally:
        g.close()
