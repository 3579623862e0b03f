"""The reference's per-colour device format (host side, CPU only).

The arrays are checked three ways:
- decoding them with the consuming RTL's semantics (row of entry k =
  sum NR - 1, NR2offset.vhd:32-41; column = P[slot]; L swept over partitions
  in order, U in reverse with rows descending and D^-1 four per scalar row,
  write_unit.vhd:200-203, ext_read_unit.vhd:114-116) reproduces the oracle's
  SpMV and ILU0 application on the same colour-ordered system;
- import(export) gives back the colour-ordered matrix, factors, D^-1, b, x0
  bit for bit;
- the device limits and line paddings the RTL read sizes imply
  (constants.vhd:80-98; ext_read_unit.vhd:97,114-116,147-152,246,282-287)
  hold, and the packed buffers decode back to the same arrays.
"""
import numpy as np
import pytest

import fpga_amd as fa
import oracle
from fpga_amd import wire
from fpga_amd.configs import GRID_32x32x16, scaled

CASES = {
    "grid32x32x16": GRID_32x32x16,
    "grid12x10x8_nnc": scaled(GRID_32x32x16, 12, 10, 8),
    "grid9x8x6_bs4_nnc": fa.GridSpec(9, 8, 6, bs=4, seed=3, perm_sigma=2.0, layer_corr=0.5, inactive_frac=0.1,
                                     nnc_frac=0.1, nnc_dist=3),
}


def permuted(s, perm):
    """A in colour order: block row k = original row perm[k], columns renumbered and sorted."""
    n = s.n
    iperm = np.empty(n, np.int64)
    iperm[perm] = np.arange(n)
    rowptr = np.zeros(n + 1, np.int32)
    cols, vals = [], []
    for k in range(n):
        i = perm[k]
        q = np.arange(s.rowptr[i], s.rowptr[i + 1])
        c = iperm[s.colidx[q]]
        o = np.argsort(c)
        cols.append(c[o])
        vals.append(s.values[q[o]])
        rowptr[k + 1] = rowptr[k] + len(q)
    return rowptr, np.concatenate(cols).astype(np.int32), np.concatenate(vals)


def setup_case(spec):
    s = fa.generate(spec)
    o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, s.values)
    o.factor()
    _, perm, cptr = o.colouring()
    A = permuted(s, perm)
    L, U, dinv = o.factors()
    rng = np.random.default_rng(1)
    bp, xp = rng.standard_normal(s.N), rng.standard_normal(s.N)
    w = wire.build(s.bs, cptr, A, L, U, dinv, bp, xp)
    return s, o, perm, cptr, A, (L, U, dinv), bp, xp, w


@pytest.fixture(scope="module", params=list(CASES))
def case(request):
    return setup_case(CASES[request.param])


def partitions(w, k):
    """(row0, arow, P, val, col, NR) per partition of matrix k, in processing order."""
    cs = w.colour_sizes(k)
    P, V, Ci, NR = (w.arrays[6 * k + j] for j in (1, 2, 3, 4))
    op = ov = onr = 0
    out = []
    for row, arow, col, val in cs:
        out.append((int(arow), P[op:op + col], V[ov:ov + val], Ci[ov:ov + val], NR[onr:onr + val]))
        op += -(-col // 16) * 16
        ov += val
        onr += -(-val // 64) * 64
    return out


def stream_rows(nr):
    return np.cumsum(nr.astype(np.int64)) - 1


def decode_spmv(w, x):
    y = np.zeros_like(x)
    row0 = 0
    for arow, P, V, Ci, NR in partitions(w, 0):
        r = stream_rows(NR)
        np.add.at(y, row0 + r, V * x[P[Ci].astype(np.int64)])
        row0 += arow
    return y


def decode_precond(w, bs, p):
    N = p.size
    y = np.zeros(N)
    row0 = 0
    for arow, P, V, Ci, NR in partitions(w, 1):        # L: y = p - L y, partitions in order
        t = np.zeros(arow)
        np.add.at(t, stream_rows(NR), V * y[P[Ci].astype(np.int64)])
        y[row0:row0 + arow] = p[row0:row0 + arow] - t
        row0 += arow
    z = np.zeros(N)
    D = w.arrays[18]
    up, od = N, 0
    for arow, P, V, Ci, NR in partitions(w, 2):        # U: z = D^-1 (y - U z), rows descending
        rows = up - 1 - np.arange(arow)
        t = np.zeros(arow)
        np.add.at(t, stream_rows(NR), V * z[P[Ci].astype(np.int64)])
        tv = y[rows] - t                               # processing order (descending rows)
        blk = D[od:od + 4 * arow].reshape(arow, 4)[:, :bs]
        tr = np.zeros(N)
        tr[rows] = tv
        for q, r in enumerate(rows):                   # z_r = sum_b D^-1[r, b] t[block(r), b]
            base = (r // bs) * bs
            z[r] = blk[q] @ tr[base:base + bs]
        up -= arow
        od += -(-4 * arow // 8) * 8
    return z


def test_limits_and_padding(case):
    s, o, perm, cptr, A, F, bp, xp, w = case
    N = s.N
    for k in range(3):
        cs = w.colour_sizes(k)
        assert cs.shape[0] == w.sizes[6 * k + 2]
        assert np.all(w.arrays[6 * k][:wire.HEADER_WORDS] == 0)
        assert np.all(cs[:, 0] == cs[:, 1]) and cs[:, 1].sum() == N
        assert np.all(cs[:, 1] % s.bs == 0) and np.all(cs[:, 1] <= 2048)
        assert np.all(cs[:, 2] <= 8192) and np.all(cs[:, 3] <= 32768) and np.all(cs[:, 3] % 32 == 0)
        assert w.sizes[6 * k + 0] == N and w.sizes[6 * k + 1] == cs[:, 3].sum() == w.nnz_val_sizes[k]
        assert w.sizes[6 * k + 3] == sum(-(-c // 16) * 16 for c in cs[:, 2])
        assert w.sizes[6 * k + 4] == sum(-(-v // 64) * 64 for v in cs[:, 3])
        # partitions never straddle a colour; L and U gather their own rows last
        bounds = np.cumsum(cs[:, 1]) // s.bs
        if k == 2:
            bounds = s.n - bounds
        assert set(cptr[1:-1]) <= set(bounds.tolist()) | {0}
        for arow, P, V, Ci, NR in partitions(w, k):
            assert np.all(Ci < len(P)) and np.all(NR <= 255)
            if len(NR):
                assert NR[0] >= 1 and stream_rows(NR)[-1] < arow
            if k > 0:
                own = np.sort(P[len(P) - arow:])
                assert np.all(np.diff(own) == 1) and np.all(P[len(P) - arow:] == own)
    cu = w.colour_sizes(2)
    assert w.sizes[5] == w.sizes[11] == w.sizes[17] == sum(-(-4 * a // 8) * 8 for a in cu[:, 1])


def test_decoded_spmv_matches_oracle(case):
    s, o, perm, cptr, A, F, bp, xp, w = case
    x = np.random.default_rng(2).standard_normal(s.N)
    ref = o.spmv_perm(x)
    np.testing.assert_allclose(decode_spmv(w, x), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_decoded_precond_matches_oracle(case):
    s, o, perm, cptr, A, F, bp, xp, w = case
    p = np.random.default_rng(3).standard_normal(s.N)
    ref = o.precond_perm(p)
    np.testing.assert_allclose(decode_precond(w, s.bs, p), ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_import_round_trip(case):
    s, o, perm, cptr, A, (L, U, dinv), bp, xp, w = case
    rowptr, colidx, vals, lu, dv, part_rows, b, x0 = wire.import_arrays(s.bs, w)
    assert np.array_equal(rowptr, A[0]) and np.array_equal(colidx, A[1])
    assert np.array_equal(vals.view(np.uint64), A[2].view(np.uint64))
    # L and U land on their blocks of A's pattern, the diagonal stays empty
    want = np.zeros_like(A[2])
    for (rp, ci, v) in (L, U):
        for i in range(s.n):
            for q in range(rp[i], rp[i + 1]):
                j = ci[q]
                blk = A[0][i] + np.searchsorted(A[1][A[0][i]:A[0][i + 1]], j)
                want[blk] = v[q]
    assert np.array_equal(lu.view(np.uint64), want.view(np.uint64))
    assert np.array_equal(dv, dinv) and np.array_equal(b, bp) and np.array_equal(x0, xp)
    assert part_rows[0] == 0 and part_rows[-1] == s.N and np.all(np.diff(part_rows) > 0)


def test_pack_unpack_and_dump(case, tmp_path):
    s, o, perm, cptr, A, F, bp, xp, w = case
    bufs = wire.pack(w, config_bits=5)
    tot, setup = wire.layout(w.sizes, w.nnz_val_sizes, 5)
    assert [b.size for b in bufs] == list(tot)
    words = bufs[0][:320].view(np.uint64)
    assert np.array_equal(words, setup)
    assert int(words[0]) == (w.sizes[1] << 32) + w.sizes[0] and int(words[1]) == (5 << 32) + w.sizes[2]
    assert int(words[8]) == (w.sizes[7] << 32) + w.sizes[6] and int(words[9]) == w.sizes[8]
    assert int(words[16]) == (w.sizes[13] << 32) + w.sizes[12] and int(words[17]) == w.sizes[14]
    assert np.all(words[31:] == np.uint64(0xDEADC0DEDEADC0DE))
    # every array pointer is a line index inside its buffer, arrays of one buffer do not overlap
    buf_of = {2: 2, 3: 3, 4: 3, 5: 2, 6: 3, 7: 3, 10: 0, 11: 1, 12: 0, 13: 1, 14: 1, 15: 3, 18: 0, 19: 1, 20: 0,
              21: 1, 22: 1, 23: 0, 24: 0, 25: 1, 26: 0, 27: 1, 28: 1, 29: 4, 30: 4}
    starts = {}
    for wd, b in buf_of.items():
        st = int(words[wd]) * 64
        assert (0 < st if b == 0 else 0 <= st) and st < tot[b]   # buffer 0 starts with the setup lines
        starts.setdefault(b, []).append(st)
    for b, st in starts.items():
        assert len(set(st)) == len(st)
    back = wire.unpack(bufs)
    assert back.sizes == list(w.sizes) and back.nnz_val_sizes == list(w.nnz_val_sizes)
    for slot, a in w.arrays.items():
        assert np.array_equal(back.arrays[slot][:a.size], a[:back.arrays[slot].size]), slot
    wire.write_dump(bufs, tmp_path, sequence=7)
    again = wire.read_dump(tmp_path, sequence=7)
    assert all(np.array_equal(a, b) for a, b in zip(again, bufs))
    r2 = wire.import_arrays(s.bs, wire.unpack(again))
    assert np.array_equal(r2[2].view(np.uint64), A[2].view(np.uint64))


def test_long_empty_gaps_use_fillers():
    """U rows of one colour with more than 255 empty scalar rows between entries."""
    n, bs = 400, 3
    rng = np.random.default_rng(5)
    pairs = {(0, 1), (300, 301)}
    rowptr, colidx = [0], []
    for i in range(n):
        cols = {i} | {j for a, j in pairs if a == i} | {a for a, j in pairs if j == i}
        colidx += sorted(cols)
        rowptr.append(len(colidx))
    rowptr, colidx = np.array(rowptr, np.int32), np.array(colidx, np.int32)
    vals = rng.uniform(-0.1, 0.1, (len(colidx), bs, bs))
    for i in range(n):
        vals[rowptr[i] + list(colidx[rowptr[i]:rowptr[i + 1]]).index(i)] += 4 * np.eye(bs)
    o = oracle.OracleSystem(bs, rowptr, colidx, vals)
    o.factor()
    _, perm, cptr = o.colouring()

    class S:
        pass
    s = S()
    s.n, s.bs, s.N, s.rowptr, s.colidx, s.values = n, bs, n * bs, rowptr, colidx, vals
    A = permuted(s, perm)
    L, U, dinv = o.factors()
    w = wire.build(bs, cptr, A, L, U, dinv)
    fillers = [(NR == 255).sum() for k in (1, 2) for (_, _, _, _, NR) in partitions(w, k)]
    assert sum(fillers) > 0
    p = rng.standard_normal(n * bs)
    ref = o.precond_perm(p)
    np.testing.assert_allclose(decode_precond(w, bs, p), ref, rtol=1e-12, atol=1e-14)
    x = rng.standard_normal(n * bs)
    np.testing.assert_allclose(decode_spmv(w, x), o.spmv_perm(x), rtol=1e-13, atol=1e-14)
    rp, ci, v, lu, dv, *_ = wire.import_arrays(bs, w)
    assert np.array_equal(rp, A[0]) and np.array_equal(v, A[2]) and np.array_equal(dv, dinv)


def test_device_fit_flag():
    small = setup_case(scaled(GRID_32x32x16, 6, 5, 4))[-1]
    assert small.device_fits
    big = setup_case(fa.GridSpec(40, 40, 40, bs=3, seed=0))[-1]   # N = 192,000 > 163,840 vector entries
    assert not big.device_fits


def _copy(w):
    c = wire.WireArrays(list(w.sizes), list(w.nnz_val_sizes), w.device_fits)
    c.arrays = {k: v.copy() for k, v in w.arrays.items()}
    return c


@pytest.mark.parametrize("damage", ["L_slot", "U_slot", "L_p", "U_row", "U_parts", "dinv", "bs"])
def test_import_rejects_corrupted_streams(case, damage):
    """A foreign or damaged buffer image is refused before anything is
    decoded: every stream's rows, slots and gathered indices, the partition
    counts against the arrays and D^-1's extent are checked, not only A's."""
    s, o, perm, cptr, A, (L, U, dinv), bp, xp, w = case
    c = _copy(w)
    bs = s.bs
    if damage in ("L_slot", "U_slot"):
        k = 1 if damage == "L_slot" else 2
        col = c.colour_sizes(k)[0][2]
        c.arrays[6 * k + 3][0] = col           # slot one past the partition's gathered list
    elif damage == "L_p":
        c.arrays[6 * 1 + 1][0] = s.N           # gathered index outside the vector
    elif damage == "U_row":
        c.arrays[6 * 2 + 4][0] = 250           # row advance beyond the partition
    elif damage == "U_parts":
        c.sizes[6 * 2 + 2] += 1                # one partition more than the header holds
        c.arrays[6 * 2 + 0] = np.concatenate([c.arrays[12], np.array([0, 1 << 20, 0, 0], np.int32)])
    elif damage == "dinv":
        c.sizes[5] = 8                          # D^-1 array shorter than its rows
    elif damage == "bs":
        bs = 5
    with pytest.raises(RuntimeError):
        wire.import_arrays(bs, c)
