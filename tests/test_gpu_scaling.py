"""User scaling, refinement modes and factor/plan lifetime through the public
API (MI355X only).

Scaling follows the reference's driver (src/spldlt_factorize_mod.F90):
options.scaling <= 0 applies the caller's scale array when one is passed
(:728-733), the factorization scatters S A S (init_a_block,
src/kernels/assemble.hxx:161-214) and the solve scales the right-hand side
before the forward sweep and the solution after the backward one (:951-966,
:1039-1053). The result must therefore be the solution of the UNSCALED
system: checked against the unscaled factorization's solution, the oracle's
factors of S A S and the scaled backward error of the original matrix.
"""
import numpy as np
import pytest

from sylver_amd import matrices as M

pytestmark = pytest.mark.gpu

BWD_TOL = 1e-15


@pytest.fixture(autouse=True)
def _defaults(S):
    S.set_ext_options(precision=S.PREC_FP64, small_front_max=0, use_graph=1, refine=S.REFINE_AUTO,
                      refine_tol=1e-15, refine_max_iter=30, indef_panel_min=0)
    yield
    S.set_ext_options(precision=S.PREC_FP64, small_front_max=0, use_graph=1, refine=S.REFINE_AUTO,
                      refine_tol=1e-15, refine_max_iter=30, indef_panel_min=0)


def _analyse(S, A):
    opts = S.default_options()
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=opts)
    assert inf.flag >= 0
    return akeep, opts, S.akeep_symbolic(akeep)


def _scale(n, seed):
    return np.random.default_rng(seed).uniform(0.25, 4.0, n)


def _solve(S, akeep, fkeep, opts, b, job=0):
    x = np.asfortranarray(b.copy())
    inf = S.spldlt_solve(x, akeep, fkeep, job=job, options=opts)
    return x, inf


@pytest.mark.parametrize("scaling_opt", [0, -1])
def test_posdef_user_scaling_fp64(S, O, gpu, scaling_opt):
    """FP64 posdef with a random positive scale (options.scaling 0 and -1
    both apply it): factors equal the oracle's factors of S A S, the solution
    is the unscaled system's, backward error <= 1e-15."""
    A = M.laplacian_3d(14)
    akeep, opts, sym = _analyse(S, A)
    opts.scaling = scaling_opt
    s = _scale(A.n, 1)
    fk, inf = S.spldlt_factorize(True, A.val, akeep, scale=s, options=opts)
    assert inf.flag == 0
    sc_piv = s[sym["invp"] - 1]
    st, off, lcol = O.factor_posdef(sym, A.val, scaling=sc_piv)
    assert st == 0
    for node in range(sym["nnodes"]):
        m = int(sym["rptr"][node + 1] - sym["rptr"][node])
        n = int(sym["sptr"][node + 1] - sym["sptr"][node])
        ref = np.tril(O.node_factor(sym, off, lcol, node))
        got = np.tril(S.fkeep_node_factor(fk, node, m, n))
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), node
    b = M.random_rhs(A.n, 3, seed=2)
    x, inf = _solve(S, akeep, fk, opts, b)
    assert inf.flag == 0
    assert O.bwd_err(A, x, b) <= BWD_TOL
    # the same system without the scaling
    opts.scaling = 0
    fk0, inf = S.spldlt_factorize(True, A.val, akeep, options=opts)
    assert inf.flag == 0
    x0, inf = _solve(S, akeep, fk0, opts, b)
    assert np.abs(x - x0).max() <= 1e-12 * np.abs(x0).max()
    # partial solves compose: job 3 (bwd, scaled out) after job 1 (fwd, scaled in)
    opts.scaling = scaling_opt
    y, inf = _solve(S, akeep, fk, opts, b, job=1)
    assert inf.flag == 0
    z, inf = _solve(S, akeep, fk, opts, y, job=3)
    assert inf.flag == 0
    assert np.abs(z - x).max() <= 1e-13 * np.abs(x).max()
    S.spldlt_free_fkeep(fk0)
    S.spldlt_free_fkeep(fk)
    S.spldlt_free_akeep(akeep)


def test_posdef_user_scaling_fp32_refined(S, O, gpu):
    """FP32 factors of S A S refined in FP64 against the unscaled A: the
    preconditioner is S (L L^T)^{-1} S, backward error <= 1e-15."""
    S.set_ext_options(precision=S.PREC_FP32, refine=S.REFINE_IR)
    A = M.laplacian_3d(16)
    akeep, opts, sym = _analyse(S, A)
    s = _scale(A.n, 3)
    fk, inf = S.spldlt_factorize(True, A.val, akeep, scale=s, options=opts)
    assert inf.flag == 0
    b = M.random_rhs(A.n, 4, seed=4)
    x, inf = _solve(S, akeep, fk, opts, b)
    assert inf.flag == 0, inf.flag
    t = inf.tail()
    assert t.refine_mode == S.REFINE_IR and t.refine_iters >= 2
    for c in range(b.shape[1]):
        assert O.bwd_err(A, x[:, c], b[:, c]) <= BWD_TOL, c
    S.spldlt_free_fkeep(fk)
    S.spldlt_free_akeep(akeep)


@pytest.mark.parametrize("panel_min", [1 << 30, 1])
def test_indef_user_scaling_fp64(S, O, gpu, panel_min):
    """Indefinite (KKT) with a random positive scale on both panels: the
    delay statistics equal the oracle's on S A S, the inertia is A's
    (congruence), and the solution solves the unscaled system."""
    S.set_ext_options(indef_panel_min=panel_min)
    A = M.kkt(8, 100, seed=3)
    akeep, opts, sym = _analyse(S, A)
    s = _scale(A.n, 5)
    fk, inf = S.spldlt_factorize(False, A.val, akeep, scale=s, options=opts)
    assert inf.flag >= 0, inf.flag
    assert inf.num_neg == 100
    ref = O.IndefFactor(sym, A.val, nb=64, u=opts.u, small=opts.small, scaling=s[sym["invp"] - 1])
    st = S.fkeep_indef_stats(fk)
    for k in ("num_delay", "num_neg", "num_two", "num_zero", "maxfront"):
        assert st[k] == ref.stats[k], (k, st[k], ref.stats[k])
    b = M.random_rhs(A.n, 2, seed=6)
    x, inf = _solve(S, akeep, fk, opts, b)
    assert inf.flag >= 0
    assert O.bwd_err(A, x, b) <= 1e-14
    S.spldlt_free_fkeep(fk)
    S.spldlt_free_akeep(akeep)


def test_refine_gmres_forced(S, O, gpu):
    """REFINE_GMRES: every correction is a GMRES-IR step; converges to
    <= 1e-15 and reports the mode."""
    S.set_ext_options(precision=S.PREC_FP32, refine=S.REFINE_GMRES)
    A = M.laplacian_3d(16)
    akeep, opts, sym = _analyse(S, A)
    fk, inf = S.spldlt_factorize(True, A.val, akeep, options=opts)
    assert inf.flag == 0
    b = M.random_rhs(A.n, 3, seed=7)
    x, inf = _solve(S, akeep, fk, opts, b)
    assert inf.flag == 0
    t = inf.tail()
    assert t.refine_mode == S.REFINE_GMRES
    assert t.bwd_err <= BWD_TOL
    assert O.bwd_err(A, x, b) <= BWD_TOL
    S.spldlt_free_fkeep(fk)
    S.spldlt_free_akeep(akeep)


def test_refine_auto_falls_back_and_warns(S, O, gpu):
    """AUTO with an unreachable target: classical IR stalls at the FP64
    floor, the next step is GMRES-IR (mode reported), and after
    refine_max_iter steps the solve returns the non-convergence warning
    while the iterate keeps its backward error <= 1e-15."""
    S.set_ext_options(precision=S.PREC_FP32, refine=S.REFINE_AUTO, refine_tol=1e-30, refine_max_iter=8)
    A = M.laplacian_3d(16)
    akeep, opts, sym = _analyse(S, A)
    fk, inf = S.spldlt_factorize(True, A.val, akeep, options=opts)
    assert inf.flag == 0
    b = M.random_rhs(A.n, 2, seed=8)
    x, inf = _solve(S, akeep, fk, opts, b)
    assert inf.flag == 9  # SYLVER_AMD_WARNING_REFINE
    t = inf.tail()
    assert t.refine_mode == S.REFINE_GMRES
    assert t.refine_iters == 8
    assert O.bwd_err(A, x, b) <= BWD_TOL
    S.spldlt_free_fkeep(fk)
    S.spldlt_free_akeep(akeep)


def test_plan_rebuild_keeps_older_factors(S, O, gpu):
    """Two factorizations on one analysis; new front variants rebuild the
    analysis's plan for the second; the first keeps solving with its own
    plan and captured graphs."""
    A = M.laplacian_3d(16)
    akeep, opts, sym = _analyse(S, A)
    b = M.random_rhs(A.n, 2, seed=9)
    fk1, inf = S.spldlt_factorize(True, A.val, akeep, options=opts)
    assert inf.flag == 0
    x1, inf = _solve(S, akeep, fk1, opts, b)  # captures fk1's solve graph
    assert inf.flag == 0
    S.set_front_variants(akeep, np.zeros(sym["nnodes"], np.int8))  # every front tiled
    fk2, inf = S.spldlt_factorize(True, A.val, akeep, options=opts)
    assert inf.flag == 0
    x2, inf = _solve(S, akeep, fk2, opts, b)
    assert inf.flag == 0
    x1b, inf = _solve(S, akeep, fk1, opts, b)
    assert inf.flag == 0
    assert np.array_equal(x1, x1b)
    assert O.bwd_err(A, x1b, b) <= BWD_TOL and O.bwd_err(A, x2, b) <= BWD_TOL
    S.set_front_variants(akeep, None)
    S.spldlt_free_fkeep(fk1)
    S.spldlt_free_fkeep(fk2)
    S.spldlt_free_akeep(akeep)


def test_equilibration_on_device_matches_oracle(S, O, gpu):
    """options.scaling >= 4: the norm equilibration runs on the device
    (sweeps over the pivot-order CSR); the scale returned to the caller
    equals the oracle's restatement bit for bit, and the scaled posdef
    factorization solves the unscaled system."""
    from oracle import scaling as OS
    A = M.laplacian_3d(12)
    rng = np.random.default_rng(11)
    f = 10.0 ** rng.uniform(-3, 3, A.n)  # a badly scaled SPD matrix
    col = np.repeat(np.arange(A.n), np.diff(A.ptr))
    row = A.row - 1
    B = M.CSCLower(A.n, A.ptr, A.row, A.val * f[row] * f[col])
    akeep, opts, sym = _analyse(S, B)
    opts.scaling = 4
    s = np.zeros(B.n)
    fk, inf = S.spldlt_factorize(True, B.val, akeep, scale=s, options=opts)
    assert inf.flag == 0
    want, sweeps = OS.equilib_scale_sym(B)
    assert np.array_equal(s, want), np.abs(s - want).max()
    b = M.random_rhs(B.n, 2, seed=12)
    x, inf = _solve(S, akeep, fk, opts, b)
    assert inf.flag == 0
    assert O.bwd_err(B, x, b) <= BWD_TOL
    S.spldlt_free_fkeep(fk)
    S.spldlt_free_akeep(akeep)


@pytest.mark.parametrize("mode", [1, 2, 4])
def test_computed_scalings_indefinite(S, O, gpu, mode):
    """KKT with badly scaled constraint rows, scaled by the matching-based
    scalings (1 shortest augmenting paths, 2 auction) or equilibration (4):
    exact inertia, backward error of the original system <= 1e-14, and the
    returned scale equals the host routine's (modes 1, 2)."""
    A = M.kkt(8, 120, seed=7)
    rng = np.random.default_rng(13)
    f = np.ones(A.n)
    f[512:] = 10.0 ** rng.uniform(-4, 4, 120)
    col = np.repeat(np.arange(A.n), np.diff(A.ptr))
    row = A.row - 1
    B = M.CSCLower(A.n, A.ptr, A.row, A.val * f[row] * f[col])
    akeep, opts, sym = _analyse(S, B)
    opts.scaling = mode
    s = np.zeros(B.n)
    fk, inf = S.spldlt_factorize(False, B.val, akeep, scale=s, options=opts)
    assert inf.flag >= 0, inf.flag
    assert inf.num_neg == 120
    if mode == 1:
        assert np.array_equal(s, S.hungarian_scale_sym(B)[1])
    elif mode == 2:
        assert np.array_equal(s, S.auction_scale_sym(B)[1])
    assert np.all(s > 0)
    b = M.random_rhs(B.n, 2, seed=14)
    x, inf = _solve(S, akeep, fk, opts, b)
    assert inf.flag >= 0
    assert O.bwd_err(B, x, b) <= 1e-14
    S.spldlt_free_fkeep(fk)
    S.spldlt_free_akeep(akeep)


def test_matching_order_and_saved_scaling(S, O, gpu):
    """ordering = 2 (matching-based) keeps its scaling for scaling = 3;
    scaling = 3 without it is SYLVER_ERROR_NO_SAVED_SCALING; another scaling
    after a matching order returns the warning
    SYLVER_WARNING_MATCH_ORD_NO_SCALE (src/spldlt_factorize_mod.F90:722-725)."""
    A = M.kkt(8, 150, seed=9)
    opts = S.default_options()
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=opts)
    opts.scaling = 3
    fk, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
    assert inf.flag == -15
    S.spldlt_free_akeep(akeep)
    opts = S.default_options()
    opts.ordering = 2
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=opts)
    assert inf.flag >= 0
    opts.scaling = 3
    fk, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
    assert inf.flag == 0, inf.flag
    assert inf.num_neg == 150
    b = M.random_rhs(A.n, 2, seed=15)
    x, inf = _solve(S, akeep, fk, opts, b)
    assert O.bwd_err(A, x, b) <= 1e-14
    opts.scaling = 0
    fk2, inf = S.spldlt_factorize(False, A.val, akeep, options=opts)
    assert inf.flag == 8
    S.spldlt_free_fkeep(fk)
    S.spldlt_free_fkeep(fk2)
    S.spldlt_free_akeep(akeep)
