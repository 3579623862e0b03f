"""Host logic on the CPU: the C-ABI library loads and exports every symbol of
include/bda_bicgstab.h; the synthetic generator; the library's host
colouring against the oracle's (bit-exact) and against the first-fit and
validity properties."""
import re
from pathlib import Path

import numpy as np
import pytest

import fpga_amd as fa
import oracle
from fpga_amd import _lib
from fpga_amd.configs import GRID_128x128x64, GRID_32x32x16, scaled

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "bda_bicgstab.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(bda_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    lib = _lib.load()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_library_missing_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="not built"):
        _lib.load(tmp_path / "libfpga_amd.so")


def test_generator_sizes_and_determinism():
    s = fa.generate(GRID_32x32x16)
    assert (s.n, s.nnzb, s.bs) == (16384, 110592, 3)
    s2 = fa.generate(GRID_32x32x16)
    assert np.array_equal(s.values, s2.values) and np.array_equal(s.b, s2.b)
    # structurally symmetric, sorted, diagonal present
    rows = np.repeat(np.arange(s.n), np.diff(s.rowptr))
    pairs = set(zip(rows.tolist(), s.colidx.tolist()))
    assert all((j, i) in pairs for i, j in pairs)
    assert all((i, i) in pairs for i in range(s.n))
    for i in range(0, s.n, 997):
        seg = s.colidx[s.rowptr[i]:s.rowptr[i + 1]]
        assert np.all(np.diff(seg) > 0)
    # diagonal dominance by row: diag = sum |offdiag| + 1 (+ N(0, 0.05) noise)
    d = s.values[s.rowptr[:-1] + np.array([np.searchsorted(s.colidx[s.rowptr[i]:s.rowptr[i + 1]], i)
                                           for i in range(s.n)])]
    assert np.all(np.diagonal(d, axis1=1, axis2=2) > 0)


def test_generator_inactive_and_nnc():
    spec = scaled(GRID_128x128x64, 24, 20, 10)
    s = fa.generate(spec)
    ncell = 24 * 20 * 10
    assert 0.85 * ncell < s.n < 0.95 * ncell          # ~10% removed
    deg = np.diff(s.rowptr)
    assert deg.max() >= 8                              # NNCs raise some rows above 7 blocks
    assert deg.min() >= 1


@pytest.mark.parametrize("spec", [GRID_32x32x16, scaled(GRID_128x128x64, 20, 18, 10),
                                  fa.GridSpec(9, 7, 5, bs=4, seed=4, inactive_frac=0.2, nnc_frac=0.1)],
                         ids=["grid32", "grid20_nnc", "grid9_bs4"])
@pytest.mark.parametrize("level", [False, True], ids=["greedy", "levels"])
def test_host_colouring_matches_oracle(spec, level):
    s = fa.generate(spec)
    c, p, cp = fa.colour_pattern(s.rowptr, s.colidx, level)
    co = oracle.colour(s.rowptr, s.colidx, level)
    np.testing.assert_array_equal(c, co)
    o = oracle.OracleSystem(s.bs, s.rowptr, s.colidx, s.values, level)
    oc, op, ocp = o.colouring()
    np.testing.assert_array_equal(p, op)
    np.testing.assert_array_equal(cp, ocp)
    # validity: no two coupled rows share a colour; first fit / level property
    rows = np.repeat(np.arange(s.n), np.diff(s.rowptr))
    off = rows != s.colidx
    assert not np.any(c[rows[off]] == c[s.colidx[off]])
    nb = [[] for _ in range(s.n)]
    for i, j in zip(rows[off], s.colidx[off]):
        nb[i].append(j)
    for i in range(0, s.n, max(1, s.n // 500)):
        earlier = {int(c[j]) for j in nb[i] if j < i}
        if level:
            assert c[i] == (max(earlier) + 1 if earlier else 0)
        else:
            assert c[i] == min(set(range(len(earlier) + 1)) - earlier)
    # stable permutation by (colour, index)
    assert np.all(np.diff(c[p]) >= 0)
    for k in range(len(cp) - 1):
        assert np.all(np.diff(p[cp[k]:cp[k + 1]]) > 0)


def test_config2_colour_count_is_stable():
    s = fa.generate(GRID_128x128x64)
    _, _, cp = fa.colour_pattern(s.rowptr, s.colidx)
    assert (s.n, s.nnzb) == (943550, 6003904)
    assert list(np.diff(cp)) == [449499, 448480, 36644, 8905, 22]


def test_one_hip_runtime_per_process():
    """Importing the package before torch must not leave two HIP runtimes in
    the process (torch would then see no GPU): the package binds torch's."""
    import subprocess
    import sys
    code = ("import fpga_amd, torch\n"
            "print(len(set(l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l)))")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=str(ROOT))
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "1", out.stdout
