"""Parity of the device posdef path against the CPU oracle (MI355X only).

Bars: integer structures (device copies of the symbolic arrays, child maps)
bit-exact; FP64 factors within 1e-12*max|L| of the oracle (the LL^T factor
of an SPD front is unique, differences come only from summation order);
scaled backward error <= 1e-15 in FP64; FP32 factors within 2e-4 relative;
mixed precision reaches <= 1e-15 through FP64 refinement.
"""
import numpy as np
import pytest

from sylver_amd import matrices as M

pytestmark = pytest.mark.gpu

FP64_L_TOL = 1e-12
BWD_TOL = 1e-15


def _analyse(S, A, order=None, ordering=1):
    opts = S.default_options()
    opts.ordering = 0 if order is not None else ordering
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, order=order, options=opts)
    assert inf.flag >= 0
    return akeep, opts, S.akeep_symbolic(akeep)


def _factor(S, A, akeep, opts, posdef=True):
    fkeep, inf = S.spldlt_factorize(posdef, A.val, akeep, options=opts)
    return fkeep, inf


def _compare_factors(S, O, sym, fkeep, A, tol):
    st, off, lcol = O.factor_posdef(sym, A.val)
    assert st == 0
    worst = 0.0
    for s in range(sym["nnodes"]):
        m = int(sym["rptr"][s + 1] - sym["rptr"][s])
        n = int(sym["sptr"][s + 1] - sym["sptr"][s])
        ref = O.node_factor(sym, off, lcol, s)
        got = np.tril(S.fkeep_node_factor(fkeep, s, m, n))
        scale = max(1.0, np.abs(ref).max())
        worst = max(worst, np.abs(got - np.tril(ref)).max() / scale)
    assert worst <= tol, worst
    return off, lcol


@pytest.fixture(autouse=True)
def _fp64_defaults(S):
    S.set_ext_options(precision=S.PREC_FP64, small_front_max=0, use_graph=1, refine=S.REFINE_AUTO)
    yield
    S.set_ext_options(precision=S.PREC_FP64, small_front_max=0, use_graph=1, refine=S.REFINE_AUTO)


def test_c_example_tridiagonal(S, gpu):
    # examples/C/spldlt_simple_example_c.c:30-40 (solved as posdef here)
    A = M.CSCLower(3, [1, 3, 5, 6], [1, 2, 2, 3, 3], [2.0, -1.0, 2.0, -1.0, 2.0])
    akeep, opts, _ = _analyse(S, A)
    fkeep, inf = _factor(S, A, akeep, opts)
    assert inf.flag == 0
    x = np.ones(3)
    inf = S.spldlt_solve(x, akeep, fkeep, options=opts)
    assert inf.flag == 0
    np.testing.assert_allclose(x, [1.5, 2.0, 1.5], rtol=0, atol=1e-15)


def test_simple_mat_known_answer(S, gpu):
    # tests/sylver_test_mod.F90:120-155 (simple_mat), x = 1
    A = M.CSCLower(4, [1, 4, 5, 7, 8], [1, 2, 4, 2, 3, 4, 4],
                   [10.0, 2.0, 3.0, 10.0, 10.0, 4.0, 10.0])
    b = A.to_dense() @ np.ones(4)
    for order in (None, np.array([1, 2, 3, 4], np.int32), np.array([4, 3, 2, 1], np.int32)):
        akeep, opts, _ = _analyse(S, A, order=order)
        fkeep, inf = _factor(S, A, akeep, opts)
        assert inf.flag == 0
        x = b.copy()
        assert S.spldlt_solve(x, akeep, fkeep, options=opts).flag == 0
        np.testing.assert_allclose(x, np.ones(4), rtol=0, atol=1e-14)


def test_laplacian_2d_100_vs_oracle(S, O, gpu):
    """2D 5-point Laplacian on 100x100 (n=10,000), nested dissection, one RHS
    uniform in [-1,1] from seed 0."""
    A = M.laplacian_2d(100)
    akeep, opts, sym = _analyse(S, A)
    fkeep, inf = _factor(S, A, akeep, opts)
    assert inf.flag == 0
    _compare_factors(S, O, sym, fkeep, A, FP64_L_TOL)
    b = M.random_rhs(A.n, 1, seed=0)
    x = b.copy(order="F")
    assert S.spldlt_solve(x, akeep, fkeep, options=opts).flag == 0
    err = O.bwd_err(A, x, b)
    assert err <= BWD_TOL, err


def test_device_symbolic_bit_exact(S, O, gpu):
    A = M.laplacian_3d(12)
    akeep, opts, sym = _analyse(S, A)
    fkeep, inf = _factor(S, A, akeep, opts)
    assert inf.flag == 0
    dev = S.fkeep_device_symbolic(fkeep, sym)
    for k in ("sptr", "sparent", "rptr", "rlist", "nptr", "nlist"):
        assert np.array_equal(dev[k], sym[k]), k
    nptr, nlist = O.build_map(A, sym["order"], sym["invp"], sym)
    assert np.array_equal(dev["nptr"], nptr) and np.array_equal(dev["nlist"], nlist)
    off, cmap = O.child_maps(A.n, sym)
    assert np.array_equal(dev["cmap_off"], off)
    assert np.array_equal(dev["cmap"][:len(cmap)], cmap)


@pytest.mark.parametrize("small_max", [0, 48])
def test_laplacian_3d_large_front_path(S, O, gpu, small_max):
    """20^3 Laplacian; small_front_max=48 routes most fronts through the
    tiled large-front path (diagonal POTRF/TRTRI, TRSM-by-GEMM, MFMA update)."""
    S.set_ext_options(small_front_max=small_max)
    A = M.laplacian_3d(20)
    akeep, opts, sym = _analyse(S, A)
    fkeep, inf = _factor(S, A, akeep, opts)
    assert inf.flag == 0
    _compare_factors(S, O, sym, fkeep, A, FP64_L_TOL)
    b = M.random_rhs(A.n, 3, seed=1)
    x = b.copy(order="F")
    assert S.spldlt_solve(x, akeep, fkeep, options=opts).flag == 0
    assert O.bwd_err(A, x, b) <= BWD_TOL


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("grid", ["3d24", "3d48"])
def test_chained_solve_matches_steps(S, O, gpu, monkeypatch, prec, grid):
    """The level-wide chained solve (one launch per level, workgroups of a
    front linked through the exchange buffer) performs the per-step solve's
    arithmetic in the same order: solutions bitwise equal, for two
    right-hand-side chunks (20 columns) and the forward-only /
    backward-only jobs. 3D 24^3: every level's contribution blocks fit in
    its chained launch; 3D 48^3: the wider levels (over 256 workgroups)
    take the separate contribution-block launch."""
    precision = S.PREC_FP64 if prec == "fp64" else S.PREC_FP32
    S.set_ext_options(precision=precision, small_front_max=48, refine=S.REFINE_NONE)
    A = M.laplacian_3d(24) if grid == "3d24" else M.laplacian_3d(48)
    b = M.random_rhs(A.n, 20, seed=3)
    out = {}
    for mode in ("steps", "chain"):
        if mode == "steps":
            monkeypatch.setenv("SYLVER_AMD_SOLVE_STEPS", "1")
        else:
            monkeypatch.delenv("SYLVER_AMD_SOLVE_STEPS", raising=False)
        akeep, opts, sym = _analyse(S, A)
        fkeep, inf = _factor(S, A, akeep, opts)
        assert inf.flag == 0
        res = []
        for job in (0, 1, 3):
            x = b.copy(order="F")
            assert S.spldlt_solve(x, akeep, fkeep, job=job, options=opts).flag == 0
            res.append(x)
        out[mode] = res
        S.spldlt_free_fkeep(fkeep)
        S.spldlt_free_akeep(akeep)
    for a, c in zip(out["steps"], out["chain"]):
        assert np.array_equal(a, c), np.abs(a - c).max()
    if prec == "fp64":
        assert O.bwd_err(A, out["chain"][0], b) <= BWD_TOL


def test_geometric_nd_order(S, O, gpu):
    A = M.laplacian_3d(16)
    akeep, opts, sym = _analyse(S, A, order=M.grid_nd_order(16))
    fkeep, inf = _factor(S, A, akeep, opts)
    assert inf.flag == 0
    _compare_factors(S, O, sym, fkeep, A, FP64_L_TOL)


def test_fp32_factor_and_refinement(S, O, gpu):
    S.set_ext_options(precision=S.PREC_FP32, refine=S.REFINE_IR, refine_tol=1e-15)
    A = M.laplacian_3d(16)
    akeep, opts, sym = _analyse(S, A)
    fkeep, inf = _factor(S, A, akeep, opts)
    assert inf.flag == 0
    _compare_factors(S, O, sym, fkeep, A, 2e-4)
    b = M.random_rhs(A.n, 4, seed=2)
    x = b.copy(order="F")
    inf = S.spldlt_solve(x, akeep, fkeep, options=opts)
    assert inf.flag == 0
    t = inf.tail()
    assert t.refine_iters >= 2
    assert O.bwd_err(A, x, b) <= BWD_TOL


def test_fp32_refinement_two_rhs_chunks(S, O, gpu):
    """FP32 factor + FP64 refinement with 20 right-hand sides: two 16-column
    chunks through the chained solves, the row-major copy of x, the residual
    and the per-workgroup column maxima; every column reaches the bound and
    the refined solution matches the FP64 path's within it."""
    A = M.laplacian_3d(20)
    b = M.random_rhs(A.n, 20, seed=4)
    S.set_ext_options(precision=S.PREC_FP32, refine=S.REFINE_IR, refine_tol=1e-15)
    akeep, opts, sym = _analyse(S, A)
    fkeep, inf = _factor(S, A, akeep, opts)
    assert inf.flag == 0
    x = b.copy(order="F")
    inf = S.spldlt_solve(x, akeep, fkeep, options=opts)
    assert inf.flag == 0
    assert inf.tail().refine_iters >= 2
    for c in range(b.shape[1]):
        assert O.bwd_err(A, x[:, c], b[:, c]) <= BWD_TOL, c
    S.set_ext_options(precision=S.PREC_FP64, refine=S.REFINE_AUTO)
    fk64, inf = _factor(S, A, akeep, opts)
    assert inf.flag == 0
    x64 = b.copy(order="F")
    assert S.spldlt_solve(x64, akeep, fk64, options=opts).flag == 0
    scale = np.abs(x64).max()
    assert np.abs(x - x64).max() <= 1e-12 * scale


def test_not_posdef_flag(S, gpu):
    # tests/sylver_test_mod.F90 simple_mat_indef factorized as posdef -> -6
    A = M.CSCLower(4, [1, 4, 5, 7, 8], [1, 2, 4, 2, 3, 4, 4], [1.0, 2.0, 3.0, 1.0, 1.0, 4.0, 1.0])
    akeep, opts, _ = _analyse(S, A)
    _, inf = _factor(S, A, akeep, opts)
    assert inf.flag == S.ERROR_NOT_POS_DEF


def test_refactorize_deterministic(S, gpu):
    """Same values twice -> bitwise identical factors (no atomics anywhere in
    assembly or factorization)."""
    A = M.laplacian_3d(14)
    akeep, opts, sym = _analyse(S, A)
    fkeep, inf = _factor(S, A, akeep, opts)
    first = [S.fkeep_node_factor(fkeep, s, int(sym["rptr"][s + 1] - sym["rptr"][s]),
                                 int(sym["sptr"][s + 1] - sym["sptr"][s])).copy()
             for s in range(sym["nnodes"])]
    fkeep, inf = S.spldlt_factorize(True, A.val, akeep, fkeep, options=opts)
    assert inf.flag == 0
    for s in range(sym["nnodes"]):
        got = S.fkeep_node_factor(fkeep, s, *first[s].shape)
        assert np.array_equal(np.tril(got), np.tril(first[s]))


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("path", ["one_launch", "plan_default", "plan_tiled", "plan_one_wg"])
def test_batched_dense_fronts(S, O, gpu, dtype, path):
    """Slice of the batched small-front workload: fronts with ncol
    log-uniform in [8,256], nrow = ncol*(1+r), r in [0,2]; one
    variable-size launch, or the launch plan with per-front variants (the
    default rule, every front tiled, every front on one workgroup); every
    front (pivot columns and contribution block) checked against the oracle."""
    nf = 48
    nrow, ncol = M.dense_front_sizes(nf, seed=3)
    offs = np.zeros(nf, np.int64)
    tot = 0
    fronts = []
    for f in range(nf):
        offs[f] = tot
        a = M.dense_front_values(int(nrow[f]), seed=100 + f)
        fronts.append(a)
        tot += int(nrow[f]) ** 2
    host = np.concatenate([a.ravel(order="F") for a in fronts]).astype(dtype)
    dev = S.DeviceArray(host)
    if path == "one_launch":
        rc, info, _ = S.batched_front_potrf(dev, nrow, ncol, offs, dtype)
    else:
        v = {"plan_default": None, "plan_tiled": np.zeros(nf, np.int8), "plan_one_wg": np.ones(nf, np.int8)}[path]
        rc, info, _ = S.batched_front_potrf_v(dev, nrow, ncol, offs, dtype, v)
    assert rc == 0
    assert int(np.abs(info).sum()) == 0
    out = dev.numpy().astype(np.float64)
    dev.free()
    tol = 1e-12 if dtype == "float64" else 2e-5
    for f in range(nf):
        m, n = int(nrow[f]), int(ncol[f])
        got = out[offs[f]:offs[f] + m * m].reshape((m, m), order="F")
        st, ref = O.front_potrf(fronts[f], n)
        assert st == 0
        scale = np.abs(ref).max()
        assert np.abs(np.tril(got) - np.tril(ref)).max() <= tol * scale * max(1, m / 64), f


def test_mixed_bf16_wide_fronts_refined(S, O, gpu):
    """Mixed tier with the BF16 update forced on every large front
    (bf16_min_ncol=64): factors stay close to the FP64 oracle (BF16 unit
    roundoff 2^-8 on the update operands) and FP64 refinement brings the
    scaled backward error to <= 1e-15."""
    S.set_ext_options(precision=S.PREC_MIXED, bf16_min_ncol=64, small_front_max=48,
                      refine=S.REFINE_AUTO, refine_tol=1e-15, refine_max_iter=60)
    try:
        A = M.laplacian_3d(16)
        akeep, opts, sym = _analyse(S, A)
        fkeep, inf = _factor(S, A, akeep, opts)
        assert inf.flag == 0
        _compare_factors(S, O, sym, fkeep, A, 3e-2)
        b = M.random_rhs(A.n, 2, seed=5)
        x = b.copy(order="F")
        inf = S.spldlt_solve(x, akeep, fkeep, options=opts)
        assert inf.flag == 0
        assert O.bwd_err(A, x, b) <= BWD_TOL
    finally:
        S.set_ext_options(bf16_min_ncol=1024, refine_max_iter=30)


@pytest.fixture(scope="module")
def lap80(S):
    """The 80^3 3D Laplacian (n = 512,000) analysed once, 16 RHS U[-1,1]
    from seed 0: the full-size mixed-precision workload."""
    A = M.laplacian_3d(80)
    akeep, opts, sym = _analyse(S, A)
    b = M.random_rhs(A.n, 16, seed=0)
    yield A, akeep, opts, sym, b
    S.spldlt_free_akeep(akeep)


@pytest.mark.parametrize("tier", ["fp32", "mixed_bf16"])
def test_laplacian_80_full_size_refined(S, O, gpu, lap80, tier):
    """80^3 at full size (7,580-row root, 60-block chains): FP32 fronts, and
    the mixed tier with BF16 Schur updates on fronts of more than 1024 pivots
    (the default threshold), each refined in FP64 to scaled backward error
    <= 1e-15 on all 16 right-hand sides, checked by the oracle."""
    A, akeep, opts, sym, b = lap80
    ncol = np.diff(sym["sptr"])
    assert int(ncol.max()) > 1024  # the BF16 tier has fronts to work on
    S.set_ext_options(precision=S.PREC_FP32 if tier == "fp32" else S.PREC_MIXED, bf16_min_ncol=1024,
                      refine=S.REFINE_AUTO, refine_tol=1e-15, refine_max_iter=30)
    fkeep, inf = _factor(S, A, akeep, opts)
    try:
        assert inf.flag == 0
        x = b.copy(order="F")
        sinf = S.spldlt_solve(x, akeep, fkeep, options=opts)
        assert sinf.flag == 0
        t = sinf.tail()
        assert 1 <= t.refine_iters <= 30
        for c in range(b.shape[1]):
            assert O.bwd_err(A, x[:, c], b[:, c]) <= BWD_TOL, c
    finally:
        S.spldlt_free_fkeep(fkeep)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_small_front_assembly_fused_bitwise(S, gpu, prec):
    """One-workgroup fronts assembled inside their factor launch (zero, A
    entries, children in list order) give bitwise the factors of the
    separate zero / scatter / extend-add launches, with and without a user
    scaling."""
    A = M.laplacian_3d(22)
    akeep, opts, sym = _analyse(S, A)
    rng = np.random.default_rng(5)
    scale = np.exp(rng.uniform(-1.0, 1.0, A.n))
    pr = S.PREC_FP64 if prec == "fp64" else S.PREC_FP32
    try:
        for sc in (None, scale):
            got = []
            for fuse in (0, 1):
                S.set_ext_options(precision=pr, small_assembly=fuse)
                o = S.default_options()
                o.scaling = 0
                fk, inf = S.spldlt_factorize(True, A.val, akeep, options=o,
                                             scale=None if sc is None else sc.copy())
                assert inf.flag == 0
                got.append([np.tril(S.fkeep_node_factor(fk, s, int(sym["rptr"][s + 1] - sym["rptr"][s]),
                                                        int(sym["sptr"][s + 1] - sym["sptr"][s])))
                            for s in range(sym["nnodes"])])
                S.spldlt_free_fkeep(fk)
            for a, b in zip(*got):
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    finally:
        S.set_ext_options(precision=S.PREC_FP64, small_assembly=0)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_strip_trsm_vs_separate_launches(S, O, gpu, prec):
    """The panel TRSM by substitution in 64-row strips (after a factor-only
    diagonal launch; inverses formed at the end) against the TRSM by the
    inverse: both within the oracle's bar, and within rounding of each other
    (the TRSM's summation differs)."""
    A = M.laplacian_3d(20)
    pr = S.PREC_FP64 if prec == "fp64" else S.PREC_FP32
    akeep, opts, sym = _analyse(S, A)
    st, why = S.plan_check(akeep, small_front_max=48)
    assert st["violations"] == 0 and st["strip_tasks"] > 0, why
    b = M.random_rhs(A.n, 3, seed=4)
    got = {}
    try:
        for strip in (1, 0):
            S.set_ext_options(precision=pr, small_front_max=48, strip_trsm=strip, refine=S.REFINE_NONE)
            fk, inf = _factor(S, A, akeep, opts)
            assert inf.flag == 0
            if prec == "fp64":
                _compare_factors(S, O, sym, fk, A, FP64_L_TOL)
            got[strip] = [np.tril(S.fkeep_node_factor(fk, s, int(sym["rptr"][s + 1] - sym["rptr"][s]),
                                                      int(sym["sptr"][s + 1] - sym["sptr"][s])))
                          for s in range(sym["nnodes"])]
            x = b.copy(order="F")
            assert S.spldlt_solve(x, akeep, fk, options=opts).flag == 0
            err = max(O.bwd_err(A, x[:, c], b[:, c]) for c in range(3))
            assert err <= (BWD_TOL if prec == "fp64" else 1e-6), err
            S.spldlt_free_fkeep(fk)
        tol = 1e-13 if prec == "fp64" else 2e-5
        for a, c in zip(got[1], got[0]):
            scale = max(1.0, np.abs(c).max())
            assert np.abs(a - c).max() <= tol * scale
    finally:
        S.set_ext_options(precision=S.PREC_FP64, small_front_max=0, strip_trsm=1, refine=S.REFINE_AUTO)
        S.spldlt_free_akeep(akeep)


@pytest.mark.parametrize("where", ["first_block", "mid_root"])
def test_strip_trsm_not_posdef_then_recovers(S, O, gpu, where):
    """A non-positive pivot in a large front on the strip-TRSM path: the
    factorization reports NOT_POS_DEF, and the next factorization of an SPD
    matrix on the same analysis is exact."""
    A = M.laplacian_3d(20)
    akeep, opts, sym = _analyse(S, A)
    root = int(np.argmax(np.diff(sym["sptr"])))  # the widest front (the root separator)
    p0, p1 = int(sym["sptr"][root]) - 1, int(sym["sptr"][root + 1]) - 1
    assert p1 - p0 > 256
    # a failure in block column 0 or 2 of the 480-column root
    pivot = p0 + 5 if where == "first_block" else p0 + 300
    var = int(sym["invp"][pivot])  # 1-based variable at that pivot
    ks = [k for k in range(A.ptr[var - 1] - 1, A.ptr[var] - 1) if A.row[k] == var]
    assert len(ks) == 1
    bad = A.val.copy()
    bad[ks[0]] = -100.0
    S.set_ext_options(small_front_max=48, strip_trsm=1)
    try:
        fk, inf = S.spldlt_factorize(True, bad, akeep, options=opts)
        assert inf.flag == S.ERROR_NOT_POS_DEF
        fk, inf = S.spldlt_factorize(True, A.val, akeep, fk, options=opts)
        assert inf.flag == 0
        b = M.random_rhs(A.n, 1, seed=2)
        x = b.copy(order="F")
        assert S.spldlt_solve(x, akeep, fk, options=opts).flag == 0
        assert O.bwd_err(A, x[:, 0], b[:, 0]) <= BWD_TOL
        S.spldlt_free_fkeep(fk)
    finally:
        S.set_ext_options(small_front_max=0, strip_trsm=1)
        S.spldlt_free_akeep(akeep)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("small_max", [0, 48])
def test_level_assembly_bitwise(S, gpu, prec, small_max):
    """A level assembled in one launch by receiving columns (zero, original
    entries, every child's columns, children in list order) gives bitwise
    the factors of the separate zero / scatter / per-child-rank extend-add
    launches, with and without a user scaling."""
    A = M.laplacian_3d(22)
    akeep, opts, sym = _analyse(S, A)
    rng = np.random.default_rng(6)
    scale = np.exp(rng.uniform(-1.0, 1.0, A.n))
    pr = S.PREC_FP64 if prec == "fp64" else S.PREC_FP32
    try:
        for sc in (None, scale):
            got = []
            for lasm in (0, 1):
                S.set_ext_options(precision=pr, small_front_max=small_max, level_assembly=lasm)
                o = S.default_options()
                o.scaling = 0
                fk, inf = S.spldlt_factorize(True, A.val, akeep, options=o,
                                             scale=None if sc is None else sc.copy())
                assert inf.flag == 0
                got.append([np.tril(S.fkeep_node_factor(fk, s, int(sym["rptr"][s + 1] - sym["rptr"][s]),
                                                        int(sym["sptr"][s + 1] - sym["sptr"][s])))
                            for s in range(sym["nnodes"])])
                S.spldlt_free_fkeep(fk)
            for a, b in zip(*got):
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    finally:
        S.set_ext_options(precision=S.PREC_FP64, small_front_max=0, level_assembly=1)
        S.spldlt_free_akeep(akeep)
