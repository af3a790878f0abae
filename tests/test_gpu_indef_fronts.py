"""The reference's single-front APP test list on the device path
(tests/testing_factor_node_indef.cxx:8-80, checks of
tests/testing_factor_node_indef.hxx:426,440): an m x m random symmetric
indefinite matrix (entries U[-1, 1], tests/common.hxx:753-771), optionally
with delays caused as the reference does (tests/common.hxx:163-180: n/8
random rows and columns and n/8 random entries scaled by 1000), factored as
a front of m rows and n fully-summed columns. The reference then finishes
the remaining (m - nelim) columns with a second TPP factorization; here the
trailing block is the contribution block of a parent front holding columns
n+1..m, so the whole matrix is factored by the device's APP + TPP path.

Checks: every column eliminated (nelim == m: the root delays nothing and
inform.flag == 0), the reference's u * bwderr <= 5e-14, inertia equal to
the eigenvalue count (numpy) where that is affordable, and the front under
test (node 0) bit for bit equal to the oracle where the one-thread oracle
is affordable (its panel arithmetic is written with the same roundings).
The reference's block-size argument (its CPU tile) has no counterpart: the
device panel has a fixed 64-column block.
"""
import ctypes as C

import numpy as np
import pytest

from sylver_amd import matrices as M

pytestmark = pytest.mark.gpu

U, SMALL = 0.01, 1e-20

# (m, n, cause delays), tests/testing_factor_node_indef.cxx:13-80
CASES = [
    # square, no delays
    (32, 32, False), (128, 128, False), (1024, 1024, False), (2048, 2048, False), (3000, 3000, False),
    (5179, 5179, False), (4085, 4085, False),
    # square, delays
    (32, 32, True), (64, 64, True), (1024, 1024, True), (2048, 2048, True), (3000, 3000, True),
    # rectangular, no delays
    (2, 1, False), (8, 4, False), (128, 64, False), (1000, 200, False), (512, 256, False), (64, 32, False),
    (1000, 500, False), (1092, 451, False), (1061, 419, False), (2000, 500, False), (3000, 500, False),
    # rectangular, delays
    (2, 1, True), (4, 2, True), (8, 4, True), (16, 8, True), (64, 32, True), (128, 64, True),
    (2048, 1024, True), (2000, 1000, True), (1000, 500, True), (1092, 451, True), (2000, 500, True),
    (3000, 500, True),
]


def _matrix(m, delays, seed):
    """Dense symmetric U[-1,1]; delays as tests/common.hxx:163-180."""
    rng = np.random.default_rng(seed)
    a = np.tril(rng.uniform(-1.0, 1.0, (m, m)))
    if delays:
        for _ in range(max(1, m // 8)):
            idx = min(int(m * rng.uniform()), m - 1)
            a[idx, :idx] *= 1000.0   # row idx left of the diagonal
            a[idx:, idx] *= 1000.0   # column idx from the diagonal down
            r, c = min(int(m * rng.uniform()), m - 1), min(int(m * rng.uniform()), m - 1)
            a[max(r, c), min(r, c)] *= 1000.0
    return a + np.tril(a, -1).T


def _front_tree(S, O, a, m, n):
    """The front (rows 1..m, columns 1..n) and, when m > n, its parent
    (columns n+1..m): the 1-based symbolic arrays of the reference's
    SymbolicTree constructor and the A-entry map of build_map."""
    rows, cols = np.nonzero(np.tril(np.ones((m, m), bool)))
    order = np.argsort(cols * m + rows, kind="stable")
    rows, cols = rows[order], cols[order]
    A = M._from_coo_lower(m, rows, cols, a[rows, cols])
    nn = 1 if m == n else 2
    sym = {"n": m, "nnodes": nn}
    sym["sptr"] = np.array([1, n + 1] + ([m + 1] if nn == 2 else []), np.int32)
    sym["sparent"] = np.array([2, 3][:nn] if nn == 2 else [2], np.int32)
    rl = list(range(1, m + 1)) + (list(range(n + 1, m + 1)) if nn == 2 else [])
    sym["rlist"] = np.array(rl, np.int32)
    sym["rptr"] = np.array([1, m + 1] + ([m + 1 + (m - n)] if nn == 2 else []), np.int64)
    ident = np.arange(1, m + 1, dtype=np.int32)
    sym["nptr"], sym["nlist"] = O.build_map(A, ident, ident, sym)
    return A, sym


def _seam_factor(S, A, sym):
    L = S.lib()
    ip = lambda v: v.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    lp = lambda v: v.ctypes.data_as(C.POINTER(C.c_long))  # noqa: E731
    L.spldlt_create_symbolic_tree.restype = C.c_void_p
    tree = L.spldlt_create_symbolic_tree(None, A.n, sym["nnodes"], ip(sym["sptr"]), ip(sym["sparent"]),
                                         lp(sym["rptr"]), ip(sym["rlist"]), lp(sym["nptr"]),
                                         lp(sym["nlist"]), 0, None, None, None, None)
    assert tree
    oc = S.OptionsC()
    oc.print_level, oc.action, oc.small, oc.u, oc.nb = 0, True, SMALL, U, 256
    oc.pivot_method, oc.failed_pivot_method = 2, 1
    ic = S.InformC()
    L.spldlt_create_numeric_tree_dbl.restype = C.c_void_p
    L.spldlt_create_numeric_tree_dbl.argtypes = [C.c_bool, C.c_void_p, C.c_void_p, C.POINTER(C.c_double),
                                                 C.c_void_p, C.c_void_p, C.POINTER(S.OptionsC),
                                                 C.POINTER(S.InformC)]
    val = np.ascontiguousarray(A.val)
    nt = L.spldlt_create_numeric_tree_dbl(False, None, tree, val.ctypes.data_as(C.POINTER(C.c_double)),
                                          None, None, C.byref(oc), C.byref(ic))
    return L, tree, nt, ic


def _seam_free(S, L, tree, nt):
    L.spldlt_destroy_numeric_tree_dbl.argtypes = [C.c_bool, C.c_void_p]
    if nt:
        L.spldlt_destroy_numeric_tree_dbl(False, nt)
    L.spldlt_destroy_symbolic_tree.argtypes = [C.c_void_p]
    L.spldlt_destroy_symbolic_tree(tree)


@pytest.mark.parametrize("m,n,delays", CASES)
def test_reference_front_case(S, O, gpu, m, n, delays):
    seed = 1000 * m + n + (7 if delays else 0)
    a = _matrix(m, delays, seed)
    A, sym = _front_tree(S, O, a, m, n)
    S.set_ext_options(precision=S.PREC_FP64, indef_panel_min=0)
    L, tree, nt, ic = _seam_factor(S, A, sym)
    try:
        assert nt, ic.flag
        assert ic.flag == 0, ic.flag  # nelim == m: nothing left uneliminated at the root
        # solve through the tree solves (identity order, no scaling)
        b = M.random_rhs(m, 1, seed=seed)[:, 0]
        x = np.ascontiguousarray(b.copy())
        dp = x.ctypes.data_as(C.POINTER(C.c_double))
        for fn in ("spldlt_tree_solve_fwd_dbl", "spldlt_tree_solve_diag_dbl", "spldlt_tree_solve_bwd_dbl"):
            f = getattr(L, fn)
            f.argtypes = [C.c_bool, C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int]
            assert f(False, nt, 1, dp, m) >= 0
        err = O.bwd_err(A, x, b)
        assert U * err <= 5e-14, err  # the reference's bar
        assert err <= 1e-13, err
        if m <= 3000:
            ev = np.linalg.eigvalsh(a)
            assert ic.num_neg == int((ev < 0).sum())
        if m * n * n <= 2e9:
            ref = O.IndefFactor(sym, A.val, nb=64, u=U, small=SMALL)
            assert {k: ref.stats[k] for k in ("num_delay", "num_two", "num_neg")} == \
                {"num_delay": ic.num_delay, "num_two": ic.num_two, "num_neg": ic.num_neg}
            got, want = S.fkeep_indef_node(nt, 0), ref.node(0)
            assert got["nelim"] == want["nelim"]
            assert np.array_equal(got["rows"], want["rows"])
            assert np.array_equal(got["dinv"], want["dinv"])
            # strictly below the diagonal: above it the device front keeps the
            # symmetric copy, the oracle zeros
            assert np.array_equal(np.tril(got["L"], -1), np.tril(want["L"], -1))
    finally:
        _seam_free(S, L, tree, nt)
