"""Parity of the device APP LDL^T path (indefinite, delayed pivots) against
the CPU oracle (MI355X only).

Bars:
- a single front (dense matrix, one supernode): the device makes the same
  pivoting decisions as the oracle and produces the same numbers, bit for
  bit (row order, D^{-1}, L), because the panel arithmetic is written with
  the same explicit roundings on both sides;
- whole trees: inertia is a Sylvester invariant and must equal the exact
  eigenvalue count; delays/2x2 counts equal the oracle's (the contribution
  blocks are summed by MFMA on the device, so a pivot choice between
  near-equal candidates could in principle differ; the cases below are
  checked to agree); scaled backward error <= 1e-14 in FP64;
- the reference's indefinite known-answer cases
  (tests/sylver_test_mod.F90:319-436, tests/drivers/sylver_test.F90:215-700).
"""
import ctypes as C

import numpy as np
import pytest

from sylver_amd import matrices as M

pytestmark = pytest.mark.gpu

BWD_TOL = 1e-14


@pytest.fixture(params=["wg", "multi"])
def panel(request, S):
    """Run a test with every front on the one-workgroup panel (k_app_front)
    or on the multi-workgroup panel (k_lp_*)."""
    S.set_ext_options(indef_panel_min=(1 << 30) if request.param == "wg" else 1)
    yield request.param
    S.set_ext_options(indef_panel_min=0)


def _analyse(S, A, natural=False):
    opts = S.default_options()
    order = None
    if natural:
        opts.ordering = 0
        order = np.arange(1, A.n + 1, dtype=np.int32)
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, order=order, options=opts)
    assert inf.flag >= 0
    return akeep, opts, S.akeep_symbolic(akeep)


def _dense_indef(n, seed, zero_block=0, scale_fs=1.0):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, (n, n))
    a = a + a.T
    if zero_block:
        a[:zero_block, :zero_block] = 0.0
    a *= scale_fs
    # explicit zeros stay in the pattern: one dense supernode
    rows, cols = np.nonzero(np.tril(np.ones((n, n), bool)))
    return M._from_coo_lower(n, rows, cols, a[rows, cols]), a


def _solve_check(S, A, akeep, fkeep, nrhs=2, seed=3):
    b = M.random_rhs(A.n, nrhs, seed=seed)
    x = np.asfortranarray(b.copy())
    inf = S.spldlt_solve(x, akeep, fkeep)
    assert inf.flag >= 0
    return M.backward_error(A, x, b)


@pytest.mark.parametrize("n,zero_block,seed", [(40, 0, 1), (150, 0, 2), (130, 60, 3), (200, 100, 4),
                                                (333, 150, 5)])
def test_single_front_bit_exact(S, O, gpu, panel, n, zero_block, seed):
    A, dense = _dense_indef(n, seed, zero_block)
    akeep, opts, sym = _analyse(S, A, natural=True)
    assert sym["nnodes"] == 1
    fkeep, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
    assert inf.flag >= 0, inf.flag
    ref = O.IndefFactor(sym, A.val, nb=64, u=opts.u, small=opts.small)
    want = ref.node(0)
    got = S.fkeep_indef_node(fkeep, 0)
    assert (got["m"], got["n"], got["nelim"]) == (want["m"], want["n"], want["nelim"])
    assert np.array_equal(got["rows"], want["rows"])
    assert np.array_equal(got["dinv"], want["dinv"])
    assert np.array_equal(np.tril(got["L"], -1), np.tril(want["L"], -1))
    ev = np.linalg.eigvalsh(dense)
    assert inf.num_neg == int((ev < 0).sum())
    assert inf.num_two == ref.stats["num_two"]
    assert _solve_check(S, A, akeep, fkeep) <= BWD_TOL


@pytest.mark.parametrize("side,mrows,seed", [(4, 12, 1), (6, 40, 2), (8, 100, 3), (10, 250, 4)])
def test_kkt_tree_vs_oracle(S, O, gpu, panel, side, mrows, seed):
    A = M.kkt(side, mrows, seed=seed)
    akeep, opts, sym = _analyse(S, A)
    fkeep, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
    assert inf.flag >= 0, inf.flag
    nh = side ** 3
    # Sylvester: [H A^T; A 0] with H SPD and A of full row rank has inertia (nh, m, 0)
    assert inf.num_neg == mrows
    assert inf.matrix_rank == A.n
    ref = O.IndefFactor(sym, A.val, nb=64, u=opts.u, small=opts.small)
    st = S.fkeep_indef_stats(fkeep)
    for k in ("num_delay", "num_neg", "num_two", "num_zero", "maxfront", "not_first_pass",
              "not_second_pass"):
        assert st[k] == ref.stats[k], (k, st[k], ref.stats[k])
    assert inf.num_delay == ref.stats["num_delay"]
    assert _solve_check(S, A, akeep, fkeep) <= BWD_TOL
    assert nh + mrows == A.n


def test_panels_identical_on_tree(S, gpu):
    """The multi-workgroup panel reproduces the one-workgroup panel bit for
    bit on every front of a tree with many delays (12^3 KKT, 432 rows; the
    root front has 478 fully-summed columns)."""
    A = M.kkt(12, 432, seed=0)
    akeep, opts, sym = _analyse(S, A)
    got = {}
    try:
        for name, pmin in (("wg", 1 << 30), ("multi", 1), ("auto", 0)):
            S.set_ext_options(indef_panel_min=pmin)
            fkeep, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
            assert inf.flag >= 0, (name, inf.flag)
            assert inf.num_neg == 432
            got[name] = [S.fkeep_indef_node(fkeep, i) for i in range(sym["nnodes"])]
            assert _solve_check(S, A, akeep, fkeep) <= BWD_TOL
    finally:
        S.set_ext_options(indef_panel_min=0)
    for name in ("multi", "auto"):
        for i, (a, b) in enumerate(zip(got["wg"], got[name])):
            assert (a["m"], a["n"], a["nelim"]) == (b["m"], b["n"], b["nelim"]), (name, i)
            assert np.array_equal(a["rows"], b["rows"]), (name, i)
            assert np.array_equal(a["dinv"], b["dinv"]), (name, i)
            assert np.array_equal(np.tril(a["L"], -1), np.tril(b["L"], -1)), (name, i)


def test_kkt_16_large_panels_vs_oracle(S, O, gpu):
    """16^3 KKT with 1024 constraint rows: fronts up to 1813 rows with 1105
    delayed pivots in all, through the multi-workgroup panel, the parallel
    second pass and the blocked solves (auto mode). Inertia is the closed
    form (4096, 1024, 0); delay statistics equal the oracle's."""
    A = M.kkt(16, 1024, seed=0)
    akeep, opts, sym = _analyse(S, A)
    fkeep, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
    assert inf.flag >= 0, inf.flag
    assert inf.num_neg == 1024 and inf.matrix_rank == A.n
    ref = O.IndefFactor(sym, A.val, nb=64, u=opts.u, small=opts.small)
    st = S.fkeep_indef_stats(fkeep)
    for k in ("num_delay", "num_neg", "num_two", "num_zero", "maxfront", "not_first_pass", "not_second_pass"):
        assert st[k] == ref.stats[k], (k, st[k], ref.stats[k])
    assert _solve_check(S, A, akeep, fkeep, nrhs=3) <= BWD_TOL


def test_partial_solves_compose(S, gpu):
    """job 1 (L), 2 (D), 3 (L^T) and 4 (D then L^T) compose to job 0
    (spldlt_solve job semantics, src/spldlt_factorize_mod.F90:1169-1286)."""
    A = M.kkt(6, 40, seed=7)
    akeep, opts, _ = _analyse(S, A)
    fkeep, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
    assert inf.flag >= 0
    b = M.random_rhs(A.n, 3, seed=5)
    full = np.asfortranarray(b.copy())
    assert S.spldlt_solve(full, akeep, fkeep, job=0).flag >= 0
    x = np.asfortranarray(b.copy())
    for job in (1, 2, 3):
        assert S.spldlt_solve(x, akeep, fkeep, job=job).flag >= 0
    np.testing.assert_allclose(x, full, rtol=0, atol=1e-12 * np.abs(full).max())
    y = np.asfortranarray(b.copy())
    assert S.spldlt_solve(y, akeep, fkeep, job=1).flag >= 0
    assert S.spldlt_solve(y, akeep, fkeep, job=4).flag >= 0
    np.testing.assert_allclose(y, full, rtol=0, atol=1e-12 * np.abs(full).max())


def test_simple_mat_indef(S, gpu):
    """tests/sylver_test_mod.F90:319-354 (4x4 indefinite)."""
    A = M.CSCLower(4, [1, 4, 5, 7, 8], [1, 2, 4, 2, 3, 4, 4], [1.0, 2.0, 3.0, 1.0, 1.0, 4.0, 1.0])
    akeep, opts, _ = _analyse(S, A)
    fkeep, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
    assert inf.flag == 0
    ev = np.linalg.eigvalsh(A.to_dense())
    assert inf.num_neg == int((ev < 0).sum())
    x1 = np.ones(4)
    b = A.to_dense() @ x1
    x = b.copy()
    assert S.spldlt_solve(x, akeep, fkeep).flag >= 0
    np.testing.assert_allclose(x, x1, rtol=0, atol=1e-14)
    # the same matrix as posdef is refused (sylver_test.F90:700-720)
    fk2, inf2 = S.spldlt_factorize(True, A.val, akeep, options=opts)
    assert inf2.flag == S.ERROR_NOT_POS_DEF


def test_zero_diagonal(S, gpu):
    """tests/sylver_test_mod.F90:412-436: [0 1 2; 1 0 1; 2 1 0] needs 2x2
    pivots; analyse warns about the missing diagonal."""
    A = M.CSCLower(3, [1, 3, 4, 4], [2, 3, 3], [1.0, 2.0, 1.0])
    opts = S.default_options()
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=opts)
    assert inf.flag == S.WARNING_MISSING_DIAGONAL
    fkeep, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
    assert inf.flag >= 0
    D = A.to_dense()
    assert inf.num_neg == int((np.linalg.eigvalsh(D) < 0).sum())
    x1 = np.array([1.0, -2.0, 3.0])
    x = D @ x1
    assert S.spldlt_solve(x, akeep, fkeep).flag >= 0
    np.testing.assert_allclose(x, x1, rtol=0, atol=1e-13)


def test_singular(S, gpu):
    """tests/drivers/sylver_test.F90:640-700: a singular matrix is an error
    with action = false and a warning (zero pivot, rank n-1) with action."""
    A = M.CSCLower(3, [1, 3, 4, 4], [1, 2, 2], [-1.0, 2.0, 1.0])  # simple_sing_mat2
    opts = S.default_options()
    opts.ordering = 0
    order = np.arange(1, 4, dtype=np.int32)
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, order=order, options=opts)
    assert inf.flag == S.WARNING_ANAL_SINGULAR
    opts.action = False
    fkeep, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
    assert inf.flag == S.ERROR_SINGULAR
    opts.action = True
    fkeep, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
    assert inf.flag == S.WARNING_FACT_SINGULAR
    assert inf.matrix_rank == 2
    x1 = np.array([1.0, 1.0, 0.0])
    x = A.to_dense() @ x1
    assert S.spldlt_solve(x, akeep, fkeep).flag >= 0
    np.testing.assert_allclose(x, x1, rtol=0, atol=1e-14)


def test_seam_numeric_tree_indef(S, gpu):
    """The internal C ABI: spldlt_create_numeric_tree_dbl(posdef = false) and
    the four tree solves (src/NumericTree.cxx:27-136)."""
    A = M.kkt(5, 30, seed=9)
    akeep, opts, sym = _analyse(S, A)
    L = S.lib()
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    lp = lambda a: a.ctypes.data_as(C.POINTER(C.c_long))  # noqa: E731
    L.spldlt_create_symbolic_tree.restype = C.c_void_p
    tree = L.spldlt_create_symbolic_tree(None, A.n, sym["nnodes"], ip(sym["sptr"]), ip(sym["sparent"]),
                                         lp(sym["rptr"]), ip(sym["rlist"]), lp(sym["nptr"]),
                                         lp(sym["nlist"]), 0, None, None, None, None)
    assert tree
    oc = S.OptionsC()
    oc.print_level, oc.action, oc.small, oc.u, oc.nb = 0, True, 1e-20, 0.01, 256
    oc.pivot_method, oc.failed_pivot_method = 2, 1
    ic = S.InformC()
    L.spldlt_create_numeric_tree_dbl.restype = C.c_void_p
    L.spldlt_create_numeric_tree_dbl.argtypes = [C.c_bool, C.c_void_p, C.c_void_p,
                                                 C.POINTER(C.c_double), C.c_void_p, C.c_void_p,
                                                 C.POINTER(S.OptionsC), C.POINTER(S.InformC)]
    # aval in the analysed (cleaned) pattern order, as the Fortran driver passes it
    val = np.ascontiguousarray(A.val)
    nt = L.spldlt_create_numeric_tree_dbl(False, None, tree, val.ctypes.data_as(C.POINTER(C.c_double)),
                                          None, None, C.byref(oc), C.byref(ic))
    assert nt and ic.flag >= 0, ic.flag
    assert ic.num_neg == 30
    inv = sym["invp"] - 1
    b = M.random_rhs(A.n, 1, seed=2)[:, 0]
    xp = np.ascontiguousarray(b[inv])
    dp = xp.ctypes.data_as(C.POINTER(C.c_double))
    for fn in ("spldlt_tree_solve_fwd_dbl", "spldlt_tree_solve_diag_dbl", "spldlt_tree_solve_bwd_dbl"):
        f = getattr(L, fn)
        f.argtypes = [C.c_bool, C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int]
        assert f(False, nt, 1, dp, A.n) >= 0
    x = np.empty(A.n)
    x[inv] = xp
    assert M.backward_error(A, x, b) <= BWD_TOL
    L.spldlt_destroy_numeric_tree_dbl.argtypes = [C.c_bool, C.c_void_p]
    L.spldlt_destroy_numeric_tree_dbl(False, nt)
    L.spldlt_destroy_symbolic_tree.argtypes = [C.c_void_p]
    L.spldlt_destroy_symbolic_tree(tree)


def test_seam_numeric_tree_posdef(S, gpu):
    """spldlt_create_numeric_tree_posdef_dbl and the posdef tree solves
    (src/NumericTreePosdef.cxx:19-83) on an SPD matrix."""
    A = M.laplacian_3d(9)
    akeep, opts, sym = _analyse(S, A)
    L = S.lib()
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    lp = lambda a: a.ctypes.data_as(C.POINTER(C.c_long))  # noqa: E731
    L.spldlt_create_symbolic_tree.restype = C.c_void_p
    tree = L.spldlt_create_symbolic_tree(None, A.n, sym["nnodes"], ip(sym["sptr"]), ip(sym["sparent"]),
                                         lp(sym["rptr"]), ip(sym["rlist"]), lp(sym["nptr"]),
                                         lp(sym["nlist"]), 0, None, None, None, None)
    oc, ic = S.OptionsC(), S.InformC()
    oc.action, oc.small, oc.u, oc.nb, oc.pivot_method, oc.failed_pivot_method = True, 1e-20, 0.01, 256, 2, 1
    f = L.spldlt_create_numeric_tree_posdef_dbl
    f.restype = C.c_void_p
    f.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_void_p, C.c_void_p,
                  C.POINTER(S.OptionsC), C.POINTER(S.InformC)]
    val = np.ascontiguousarray(A.val)
    nt = f(None, tree, val.ctypes.data_as(C.POINTER(C.c_double)), None, None, C.byref(oc), C.byref(ic))
    assert nt and ic.flag == 0
    inv = sym["invp"] - 1
    b = M.random_rhs(A.n, 1, seed=4)[:, 0]
    xp = np.ascontiguousarray(b[inv])
    dp = xp.ctypes.data_as(C.POINTER(C.c_double))
    for fn in ("spldlt_tree_solve_fwd_posdef_dbl", "spldlt_tree_solve_bwd_posdef_dbl"):
        g = getattr(L, fn)
        g.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int]
        assert g(nt, 1, dp, A.n) >= 0
    x = np.empty(A.n)
    x[inv] = xp
    assert M.backward_error(A, x, b) <= 1e-15
    L.spldlt_destroy_numeric_tree_posdef_dbl.argtypes = [C.c_void_p]
    L.spldlt_destroy_numeric_tree_posdef_dbl(nt)
    L.spldlt_destroy_symbolic_tree.argtypes = [C.c_void_p]
    L.spldlt_destroy_symbolic_tree(tree)


def test_kkt_48_matching_order_full_size(S, O, gpu):
    """The KKT workload at full size (H = 48^3 Laplacian, 27,648 constraint
    rows) with the matching-based ordering and its saved scaling: one tree
    level then holds more than 2^32 front entries (its zero fill is a
    grid-stride launch), fronts reach 56,263 rows; inertia exactly
    (110,592, 27,648, 0) and the scaled backward error <= 1e-14."""
    A = M.kkt(48, 27648, seed=0)
    opts = S.default_options()
    opts.ordering = 2
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=opts)
    assert inf.flag >= 0
    opts.scaling = 3
    fk, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
    try:
        assert inf.flag == 0, inf.flag
        assert inf.num_neg == 27648 and inf.matrix_rank == A.n
        assert inf.maxfront > 50000
        b = M.random_rhs(A.n, 1, seed=0)
        x = b.copy(order="F")
        assert S.spldlt_solve(x, akeep, fk, options=opts).flag == 0
        assert O.bwd_err(A, x, b) <= BWD_TOL
    finally:
        S.spldlt_free_fkeep(fk)
        S.spldlt_free_akeep(akeep)


@pytest.mark.parametrize("refine", ["gmres", "auto"])
def test_fp32_tier_refined_kkt16(S, O, gpu, refine):
    """Indefinite FP32 tier (contribution-block Schur updates on FP32 MFMA;
    APP pivot tests, D and its 2x2 inverses in FP64) with FP64 GMRES-IR
    (forced, or AUTO: IR falling back to GMRES-IR) against the FP64 matrix:
    KKT with H = 16^3, 1,024 constraint rows. Inertia exactly (4096, 1024,
    0) and scaled backward error <= 1e-15 on every right-hand side (oracle)."""
    A = M.kkt(16, 1024, seed=0)
    akeep, opts, sym = _analyse(S, A)
    mode = S.REFINE_GMRES if refine == "gmres" else S.REFINE_AUTO
    S.set_ext_options(precision=S.PREC_FP32, refine=mode, refine_tol=1e-15, refine_max_iter=20)
    try:
        fk, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
        assert inf.flag == 0, inf.flag
        assert inf.tail().precision == S.PREC_FP32
        assert inf.num_neg == 1024 and inf.matrix_rank == A.n
        b = M.random_rhs(A.n, 3, seed=7)
        x = b.copy(order="F")
        sinf = S.spldlt_solve(x, akeep, fk, options=opts)
        assert sinf.flag == 0, sinf.flag
        t = sinf.tail()
        assert t.refine_iters >= 1
        if refine == "gmres":
            assert t.refine_mode == S.REFINE_GMRES
        for c in range(b.shape[1]):
            assert O.bwd_err(A, x[:, c], b[:, c]) <= 1e-15, c
        S.spldlt_free_fkeep(fk)
    finally:
        S.set_ext_options(precision=S.PREC_FP64, refine=S.REFINE_AUTO)
        S.spldlt_free_akeep(akeep)


def test_fp32_tier_matching_order_kkt(S, O, gpu):
    """The FP32 tier with the matching-based ordering and its saved scaling
    (scaling = 3): the refinement runs on the unscaled system."""
    A = M.kkt(12, 432, seed=2)
    opts = S.default_options()
    opts.ordering = 2
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=opts)
    opts.scaling = 3
    S.set_ext_options(precision=S.PREC_FP32, refine=S.REFINE_GMRES, refine_tol=1e-15, refine_max_iter=20)
    try:
        fk, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
        assert inf.flag == 0 and inf.num_neg == 432
        b = M.random_rhs(A.n, 2, seed=3)
        x = b.copy(order="F")
        assert S.spldlt_solve(x, akeep, fk, options=opts).flag == 0
        assert O.bwd_err(A, x, b) <= 1e-15
        S.spldlt_free_fkeep(fk)
    finally:
        S.set_ext_options(precision=S.PREC_FP64, refine=S.REFINE_AUTO)
        S.spldlt_free_akeep(akeep)
