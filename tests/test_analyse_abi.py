"""CPU tests of the library boundary and its symbolic analysis (no GPU):
- the shared library loads and exports every function include/*.h declares;
- sylver_inform_t / sylver_options_t have the reference's layout (checked
  against offsetof() from the C compiler on include/sylver/sylver.h, the
  reference's include path);
- default options are the reference's (src/sylver_datatypes_mod.F90:97-198);
- the supernodal structure contains the exact Cholesky structure (brute
  force symbolic elimination on small graphs) for several orderings;
- input-check flags of spldlt_analyse for the reference's warning/error
  cases (tests/drivers/sylver_test.F90:90-300, 420-530);
- without a HIP device the numeric phase refuses with -51 (no CPU fallback).
"""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from sylver_amd import matrices as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = []
    for h in ("sylver/sylver.h", "sylver_amd.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:extern\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", txt, re.M):
            if m.group(1) not in ("if", "sizeof"):
                names.append(m.group(1))
    return sorted(set(names))


def test_exports_every_declared_symbol(S):
    L = S.lib()
    names = _declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(S.EXPORTED)


def test_struct_layout_matches_header(S):
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "sylver/sylver.h"
#define P(T, f) printf(#T " " #f " %zu\n", offsetof(T, f));
int main(void) {
  P(sylver_inform_t, flag) P(sylver_inform_t, num_factor) P(sylver_inform_t, num_flops)
  P(sylver_inform_t, num_neg) P(sylver_inform_t, cublas_error) P(sylver_inform_t, unused)
  P(sylver_options_t, ordering) P(sylver_options_t, prune_tree) P(sylver_options_t, min_gpu_work)
  P(sylver_options_t, small) P(sylver_options_t, u) P(sylver_options_t, small_subtree_threshold)
  P(sylver_options_t, nb) P(sylver_options_t, action) P(sylver_options_t, use_gpu)
  P(sylver_options_t, gpu_perf_coeff) P(sylver_options_t, scheduler)
  printf("sizeof inform %zu\nsizeof options %zu\n", sizeof(sylver_inform_t), sizeof(sylver_options_t));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "l")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    want = {}
    for line in out:
        parts = line.split()
        if len(parts) == 3 and not line.startswith("sizeof"):
            want[(parts[0], parts[1])] = int(parts[2])
    for (T, f), off in want.items():
        cls = S.Inform if T == "sylver_inform_t" else S.Options
        assert getattr(cls, f).offset == off, (T, f)
    sizes = [l for l in out if l.startswith("sizeof")]
    assert C.sizeof(S.Inform) == int(sizes[0].split()[-1])
    assert C.sizeof(S.Options) == int(sizes[1].split()[-1])


def test_default_options(S):
    o = S.default_options()
    assert (o.ordering, o.nemin, o.nb, o.pivot_method, o.failed_pivot_method) == (1, 32, 256, 2, 1)
    assert o.small == 1e-20 and o.u == 0.01 and o.action and o.prune_tree
    assert o.min_gpu_work == 5000000000 and o.small_subtree_threshold == 4000000


def _structure_ok(A, sym):
    n = A.n
    inv = sym["invp"] - 1
    P = (A.to_dense() != 0)[np.ix_(inv, inv)]
    L = np.tril(P)
    for k in range(n):
        rows = np.nonzero(L[k + 1:, k])[0] + k + 1
        if len(rows):
            L[np.ix_(rows, rows)] |= np.tril(np.ones((len(rows), len(rows)), bool))
    for s in range(sym["nnodes"]):
        rl = set((sym["rlist"][sym["rptr"][s] - 1:sym["rptr"][s + 1] - 1] - 1).tolist())
        for c in range(sym["sptr"][s] - 1, sym["sptr"][s + 1] - 1):
            need = set(np.nonzero(L[c:, c])[0] + c)
            if not need <= rl:
                return False
    return True


@pytest.mark.parametrize("kind", ["2d8", "3d4", "kkt", "rand"])
def test_supernodes_contain_cholesky_structure(S, kind):
    if kind == "2d8":
        A = M.laplacian_2d(8)
    elif kind == "3d4":
        A = M.laplacian_3d(4)
    elif kind == "kkt":
        A = M.kkt(4, 12, seed=1)
    else:
        rng = np.random.default_rng(3)
        n = 60
        d = (rng.random((n, n)) < 0.06)
        d = np.tril(d | d.T | np.eye(n, dtype=bool))
        rows, cols = np.nonzero(d)
        A = M._from_coo_lower(n, rows, cols, np.where(rows == cols, n, -0.5))
    for ordering, order in ((1, None), (0, np.arange(1, A.n + 1, dtype=np.int32))):
        opts = S.default_options()
        opts.ordering = ordering
        akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, order=order, options=opts)
        assert inf.flag >= 0
        sym = S.akeep_symbolic(akeep)
        assert sorted(sym["order"].tolist()) == list(range(1, A.n + 1))
        assert _structure_ok(A, sym)
        # rlist: own pivots first then ascending rows; sparent after child
        for s in range(sym["nnodes"]):
            rl = sym["rlist"][sym["rptr"][s] - 1:sym["rptr"][s + 1] - 1]
            assert np.all(np.diff(rl) > 0)
            assert rl[0] == sym["sptr"][s]
            assert sym["sparent"][s] > s + 1
        S.spldlt_free_akeep(akeep)


def _flag(S, n, ptr, row, order=None, check=True):
    opts = S.default_options()
    opts.ordering = 0 if order is not None else 1
    akeep, inf, _ = S.spldlt_analyse(n, np.array(ptr, np.int64), np.array(row, np.int32),
                                     order=order, check=check, options=opts)
    return inf


SIMPLE_PTR = [1, 4, 5, 7, 8]
SIMPLE_ROW = [1, 2, 4, 2, 3, 4, 4]


def test_analyse_warning_and_error_flags(S):
    nat = np.arange(1, 5, dtype=np.int32)
    assert _flag(S, 4, SIMPLE_PTR, SIMPLE_ROW, nat).flag == 0
    # out of range (row -1 and row n+1)
    assert _flag(S, 4, [1, 4, 5, 7, 9], SIMPLE_ROW + [-1], nat).flag == S.WARNING_IDX_OOR
    assert _flag(S, 4, [1, 4, 5, 7, 9], SIMPLE_ROW + [5], nat).flag == S.WARNING_IDX_OOR
    # duplicate
    inf = _flag(S, 4, [1, 4, 5, 7, 9], SIMPLE_ROW + [4], nat)
    assert inf.flag == S.WARNING_DUP_IDX and inf.matrix_dup == 1
    assert _flag(S, 4, [1, 4, 5, 7, 10], SIMPLE_ROW + [5, 4], nat).flag == S.WARNING_DUP_AND_OOR
    # missing diagonal (column 3 has no diagonal)
    assert _flag(S, 4, [1, 4, 5, 6, 7], [1, 2, 4, 2, 4, 4], nat).flag == S.WARNING_MISSING_DIAGONAL
    assert _flag(S, 4, [1, 4, 5, 6, 8], [1, 2, 4, 2, 4, 4, -1], nat).flag == S.WARNING_MISS_DIAG_OORDUP
    # structurally singular: a variable with no entry at all
    inf = _flag(S, 3, [1, 3, 4, 4], [1, 2, 2], np.arange(1, 4, dtype=np.int32))
    assert inf.flag == S.WARNING_ANAL_SINGULAR
    # errors
    assert _flag(S, -1, [1], []).flag == S.ERROR_A_N_OOR
    assert _flag(S, 4, [0, 4, 5, 7, 8], SIMPLE_ROW).flag == S.ERROR_A_PTR
    assert _flag(S, 4, [1, 4, 3, 7, 8], SIMPLE_ROW).flag == S.ERROR_A_PTR
    assert _flag(S, 2, [1, 2, 3], [3, 4]).flag == S.ERROR_A_ALL_OOR
    for bad in ([1, 2, 3, 3], [0, 1, 2, 3], [1, 2, 3, 5]):
        assert _flag(S, 4, SIMPLE_PTR, SIMPLE_ROW, np.array(bad, np.int32)).flag == S.ERROR_ORDER


def test_duplicates_are_summed(S):
    """Duplicated entries are summed through the conversion map (the
    analysed pattern has one entry per position)."""
    ptr = [1, 5, 6, 8, 9]
    row = [1, 2, 2, 4, 2, 3, 4, 4]
    akeep, inf, _ = S.spldlt_analyse(4, np.array(ptr, np.int64), np.array(row, np.int32),
                                     options=S.default_options())
    assert inf.flag == S.WARNING_DUP_IDX
    sym = S.akeep_symbolic(akeep)
    assert len(sym["nlist"]) // 2 == 7


def test_numeric_phase_refuses_without_device(S):
    if S.device_count() > 0:
        pytest.skip("a HIP device is visible")
    A = M.laplacian_2d(5)
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=S.default_options())
    fkeep, inf = S.spldlt_factorize(True, A.val, akeep)
    assert inf.flag == S.ERROR_CUDA_UNKNOWN
    x = np.ones(A.n)
    assert S.spldlt_solve(x, akeep, fkeep).flag == S.ERROR_CALL_SEQUENCE


def test_call_sequence_errors(S):
    x = np.ones(4)
    assert S.spldlt_solve(x, C.c_void_p(), C.c_void_p()).flag == S.ERROR_CALL_SEQUENCE
    fk, inf = S.spldlt_factorize(True, np.ones(7), C.c_void_p())
    assert inf.flag == S.ERROR_CALL_SEQUENCE


def test_seam_symbolic_tree_roundtrip(S):
    """spldlt_create_symbolic_tree (the reference's internal C ABI) accepts
    the analyse arrays and can be destroyed."""
    A = M.laplacian_2d(6)
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=S.default_options())
    sym = S.akeep_symbolic(akeep)
    L = S.lib()
    L.spldlt_create_symbolic_tree.restype = C.c_void_p
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    lp = lambda a: a.ctypes.data_as(C.POINTER(C.c_long))  # noqa: E731
    t = L.spldlt_create_symbolic_tree(None, A.n, sym["nnodes"], ip(sym["sptr"]), ip(sym["sparent"]),
                                      lp(sym["rptr"]), ip(sym["rlist"]), lp(sym["nptr"]),
                                      lp(sym["nlist"]), 0, None, None, None, None)
    assert t
    L.spldlt_destroy_symbolic_tree.argtypes = [C.c_void_p]
    L.spldlt_destroy_symbolic_tree(t)


def test_seam_struct_layouts(S):
    """sylver_options_c / sylver_inform_c as ctypes mirror the C header."""
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include <stdbool.h>
#include "sylver_amd.h"
#define P(T, f) printf(#T " " #f " %zu\n", offsetof(T, f));
int main(void) {
  P(sylver_options_c, print_level) P(sylver_options_c, action) P(sylver_options_c, small)
  P(sylver_options_c, u) P(sylver_options_c, multiplier) P(sylver_options_c, small_subtree_threshold)
  P(sylver_options_c, nb) P(sylver_options_c, pivot_method) P(sylver_options_c, failed_pivot_method)
  P(sylver_options_c, cpu_topology) P(sylver_inform_c, flag) P(sylver_inform_c, not_second_pass)
  printf("sizeof %zu %zu\n", sizeof(sylver_options_c), sizeof(sylver_inform_c));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "l")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    for line in out:
        parts = line.split()
        if len(parts) == 3 and parts[0] != "sizeof":
            cls = S.OptionsC if parts[0] == "sylver_options_c" else S.InformC
            assert getattr(cls, parts[1]).offset == int(parts[2]), line
        elif len(parts) == 3:
            assert C.sizeof(S.OptionsC) == int(parts[1]) and C.sizeof(S.InformC) == int(parts[2])
