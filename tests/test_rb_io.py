"""Rutherford-Boeing input and the spldlt_test driver (the reference's
drivers/spldlt_test.F90, which reads matrices with SPRAL's rb_read).

CPU: the reader against an independent writer (scipy.io.hb_write, full
unsymmetric storage), hand-written symmetric files in other fixed-width
formats (D exponents, 1P scale factors, a missing exponent letter, an entry
stored above the diagonal), pattern-only files with made-up values, and the
error codes; the driver builds and, without a device, stops at factorize with
the no-device flag. GPU: the driver runs the reference's test sequence (b =
A * ones, analyse, factorize, solve) to x = 1 on a posdef and an indefinite
matrix.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

from sylver_amd import matrices as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "sylver_amd", "spldlt_test")


def rb_read(S, path, values=2, seed=0):
    L = S.lib()
    n, m = C.c_int(), C.c_int()
    ptr, row, val = C.POINTER(C.c_long)(), C.POINTER(C.c_int)(), C.POINTER(C.c_double)()
    typ = C.create_string_buffer(4)
    flag = L.sylver_amd_rb_read(path.encode(), values, C.c_ulonglong(seed), C.byref(m), C.byref(n),
                                C.byref(ptr), C.byref(row), C.byref(val), typ)
    if flag < 0:
        return flag, None, typ.value.decode()
    nn = n.value
    p = np.ctypeslib.as_array(ptr, (nn + 1,)).copy()
    nz = int(p[-1] - 1)
    r = np.ctypeslib.as_array(row, (max(nz, 1),))[:nz].copy()
    v = np.ctypeslib.as_array(val, (max(nz, 1),))[:nz].copy()
    for q in (ptr, row, val):
        L.sylver_amd_rb_free(q)
    lower = sp.csc_matrix((v, r - 1, p - 1), shape=(nn, nn))
    return flag, lower, typ.value.decode()


def _sym_random(n, seed, density=0.3):
    a = sp.random(n, n, density=density, random_state=seed, format="csc")
    return (a + a.T + sp.eye(n) * 4.0).tocsc()


def test_reader_against_scipy_writer(S, tmp_path):
    for seed, n in ((1, 6), (2, 40), (3, 211)):
        A = _sym_random(n, seed)
        f = str(tmp_path / f"a{seed}.rb")
        sio.hb_write(f, A)
        flag, lower, typ = rb_read(S, f)
        assert flag == 0 and typ == "rua"
        ref = sp.tril(A).tocsc()
        ref.sort_indices()
        assert lower.shape == ref.shape
        assert np.array_equal(lower.indptr, ref.indptr) and np.array_equal(lower.indices, ref.indices)
        assert np.array_equal(lower.data, ref.data)  # 17 significant digits: exact round trip


RSA = """\
hand-written symmetric test matrix                                      RSA1
             6             1             2             3
rsa                        4             4             6             0
(5I4)           (8I3)           (1P,3D15.7)
   1   4   5   6
   7
  1  2  4  3  3  4
  1.0000000D+01  2.0000000D+00  3.0000000-01
  1.0000000D+01  1.0000000D+01  1.0000000D+01
"""


def test_reader_symmetric_formats(S, tmp_path):
    """1P scale factor, D exponents, a value without exponent letter
    (3.0-01 = 0.3), pointers continued on a second line."""
    f = tmp_path / "s.rsa"
    f.write_text(RSA)
    flag, lower, typ = rb_read(S, str(f))
    assert flag == 0 and typ == "rsa"
    d = lower.toarray()
    assert lower.nnz == 6
    # pointers 1 4 5 6 7: column 1 rows {1, 2, 4}, columns 2 and 3 row 3,
    # column 4 row 4
    exp = np.zeros((4, 4))
    exp[0, 0], exp[1, 0], exp[3, 0] = 10.0, 2.0, 0.3
    exp[2, 1] = 10.0
    exp[2, 2] = 10.0
    exp[3, 3] = 10.0
    assert np.allclose(d, exp, rtol=1e-15, atol=0), d


UPPER = """\
upper-stored symmetric matrix                                           UP1
             4             1             1             2
rsa                        3             3             5             0
(4I4)           (5I4)           (3E20.12)
   1   2   4   6
   1   1   2   2   3
  4.000000000000E+00 -1.000000000000E+00  4.000000000000E+00
 -1.000000000000E+00  4.000000000000E+00
"""


def test_reader_mirrors_upper_storage(S, tmp_path):
    f = tmp_path / "u.rsa"
    f.write_text(UPPER)
    flag, lower, typ = rb_read(S, str(f))
    assert flag == 0
    exp = np.array([[4.0, 0, 0], [-1.0, 4.0, 0], [0, -1.0, 4.0]])
    assert np.array_equal(lower.toarray(), exp)
    assert np.array_equal(lower.indices, [0, 1, 1, 2, 2])  # sorted within columns


def _pattern_file(path, A):
    low = sp.tril(A).tocsc()
    low.sort_indices()
    n, nz = low.shape[0], low.nnz
    ptr = low.indptr + 1
    idx = low.indices + 1
    lines = ["pattern-only test matrix".ljust(72) + "PAT1".ljust(8)]
    pl, il = (len(ptr) + 9) // 10, (nz + 9) // 10
    lines.append(f"{pl + il:14d}{pl:14d}{il:14d}{0:14d}")
    lines.append(f"psa           {n:14d}{n:14d}{nz:14d}{0:14d}")
    lines.append("(10I8)          (10I8)")
    for a in (ptr, idx):
        for k in range(0, len(a), 10):
            lines.append("".join(f"{int(v):8d}" for v in a[k:k + 10]))
    path.write_text("\n".join(lines) + "\n")


def test_reader_pattern_made_up_values(S, tmp_path):
    A = M.laplacian_2d(7)
    full = sp.csc_matrix((A.val, A.row - 1, A.ptr - 1), shape=(A.n, A.n))
    f = tmp_path / "p.psa"
    _pattern_file(f, full)
    flag, lower, typ = rb_read(S, str(f), values=2, seed=3)
    assert flag == 1 and typ == "psa"  # warning: values made up
    ref = sp.tril(full).tocsc()
    assert np.array_equal(lower.indptr, ref.indptr) and np.array_equal(lower.indices, ref.indices)
    sym = (lower + sp.tril(lower, -1).T).toarray()
    off = sym - np.diag(np.diag(sym))
    assert np.abs(off).max() <= 1.0 and np.abs(off).max() > 0
    assert np.allclose(np.diag(sym), 1.0 + np.abs(off).sum(axis=1), rtol=1e-14, atol=0)  # dominant
    assert np.linalg.eigvalsh(sym).min() > 0
    flag2, lower2, _ = rb_read(S, str(f), values=2, seed=3)
    assert np.array_equal(lower2.data, lower.data)  # seeded
    flag0, lower0, _ = rb_read(S, str(f), values=0)
    assert flag0 == 1 and np.all(lower0.data == 0)
    assert rb_read(S, str(f), values=1)[0] == -5  # SYLVER_AMD_RB_ERROR_NO_VALUES


def test_reader_errors(S, tmp_path):
    assert rb_read(S, str(tmp_path / "missing.rb"))[0] == -2
    bad = tmp_path / "c.rb"
    bad.write_text(RSA.replace("rsa ", "csa "))
    assert rb_read(S, str(bad))[0] == -4
    junk = tmp_path / "j.rb"
    junk.write_text("title\n1 1\n")
    assert rb_read(S, str(junk))[0] == -3


def test_driver_builds_and_reads_without_device(S, tmp_path):
    assert os.path.exists(DRIVER)
    if S.device_count() > 0:
        pytest.skip("a HIP device is visible: the GPU tests run the driver")
    f = str(tmp_path / "lap.rb")
    sio.hb_write(f, _full(M.laplacian_3d(6)))
    r = subprocess.run([DRIVER, "--mat", f, "--posdef"], capture_output=True, text=True, timeout=120)
    assert "Reading...\nok" in r.stdout
    assert r.returncode == 3 and "spldlt_factorize flag -51" in r.stdout, r.stdout
    r = subprocess.run([DRIVER, "--bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Unrecognised command line argument" in r.stdout


def _full(A):
    low = sp.csc_matrix((A.val, A.row - 1, A.ptr - 1), shape=(A.n, A.n))
    return (low + sp.tril(low, -1).T).tocsc()


def _run_driver(args):
    r = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {}
    for line in r.stdout.splitlines():
        if "bwd error scaled" in line:
            out["bwd"] = [float(v) for v in line.split("=")[1].split()]
        elif "number bad cmp" in line:
            out["bad"] = int(line.split("=")[1])
        elif line.startswith("num_neg:"):
            out["neg"] = int(line.split(":")[1])
        elif line.startswith("delay:"):
            out["delay"] = int(line.split(":")[1])
    return out, r.stdout


@pytest.mark.gpu
def test_driver_posdef_known_answer(S, gpu, tmp_path):
    f = str(tmp_path / "lap.rb")
    sio.hb_write(f, _full(M.laplacian_3d(14)))
    for extra in ([], ["--precision=fp32"]):
        out, txt = _run_driver(["--mat", f, "--posdef", "--nrhs", "3", *extra])
        assert out["bad"] == 0, txt
        assert len(out["bwd"]) == 3 and max(out["bwd"]) <= 1e-15, txt


@pytest.mark.gpu
def test_driver_indefinite_kkt(S, gpu, tmp_path):
    side, mrows = 8, 128
    f = str(tmp_path / "kkt.rb")
    sio.hb_write(f, _full(M.kkt(side, mrows, seed=4)))
    for extra in ([], ["--ordering=mc64-metis"]):
        out, txt = _run_driver(["--mat", f, "--indef", *extra])
        assert out["neg"] == mrows, txt
        assert max(out["bwd"]) <= 1e-14, txt
