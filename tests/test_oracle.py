"""The CPU oracle pinned against closed forms and the golden fixtures
(no GPU). Each check is chosen so that a plausible mistake in the oracle (a
dropped term, a transposed operand, a wrong index) fails it:
- front factorization against numpy's Cholesky and the Schur complement
  formula S = A22 - A21 A11^{-1} A12, for several block sizes so that the
  tiled code paths are all exercised;
- the multifrontal factor, node by node, against the dense Cholesky factor
  of the permuted matrix (the LL^T factor of an SPD matrix is unique);
- the solve against numpy.linalg.solve, the backward error against its
  definition evaluated densely;
- build_map / child maps / symbolic tree against structural invariants.
"""
import json
import os

import numpy as np
import pytest

from sylver_amd import matrices as M

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "small_cases.json")


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)


def _case(name):
    for c in _golden()["cases"]:
        if c["name"] == name:
            return c
    raise KeyError(name)


def _csc(c):
    return M.CSCLower(c["n"], c["ptr"], c["row"], c["val"])


@pytest.mark.parametrize("nb", [1, 7, 16, 256])
@pytest.mark.parametrize("m,n", [(51, 51), (51, 1), (51, 23), (90, 40), (33, 33)])
def test_front_potrf_closed_form(O, m, n, nb):
    rng = np.random.default_rng(m * 100 + n)
    a = rng.uniform(-1, 1, (m, m))
    a = 0.5 * (a + a.T) + m * np.eye(m)
    info, f = O.front_potrf(a, n, nb=nb)
    assert info == 0
    L11 = np.linalg.cholesky(a[:n, :n])
    L21 = np.linalg.solve(L11, a[n:, :n].T).T
    S = a[n:, n:] - L21 @ L21.T
    amax = lambda z: float(np.abs(z).max()) if z.size else 0.0  # noqa: E731
    assert amax(np.tril(f[:n, :n]) - L11) <= 1e-14 * m
    assert amax(f[n:, :n] - L21) <= 1e-14 * m
    assert amax(np.tril(f[n:, n:]) - np.tril(S)) <= 1e-13 * m


def test_front_potrf_reference_unit_block(O):
    """src/kernels/test/chol_kernels.cxx: 51x51 diagonally dominant block,
    factorize_diag_block against host_potrf at 1e-14 (FP64)."""
    g = _golden()["chol_dd"]
    m = g["m"]
    a = np.array(g["a"]).reshape((m, m), order="F")
    L = np.array(g["l"]).reshape((m, m), order="F")
    for n in (m, 1, 23):
        info, f = O.front_potrf(a, n)
        assert info == 0
        assert np.abs(np.tril(f[:, :n]) - L[:, :n]).max() < 1e-14


def test_front_potrf_not_posdef(O):
    a = np.diag([4.0, 1.0, -1.0, 2.0])
    info, _ = O.front_potrf(a, 4)
    assert info == 3


def _analyse(S, A, order=None):
    opts = S.default_options()
    opts.ordering = 0 if order is not None else 1
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, order=order, options=opts)
    assert inf.flag >= 0
    return S.akeep_symbolic(akeep)


@pytest.mark.parametrize("nb", [2, 256])
@pytest.mark.parametrize("which", ["lap2d_4", "lap3d_3"])
def test_multifrontal_equals_dense_cholesky(S, O, which, nb):
    c = _case(which)
    A = _csc(c)
    natural = np.arange(1, A.n + 1, dtype=np.int32)
    sym = _analyse(S, A, order=natural)
    st, off, lcol = O.factor_posdef(sym, A.val, nb=nb)
    assert st == 0
    D = A.to_dense()
    inv = sym["invp"] - 1
    L = np.linalg.cholesky(D[np.ix_(inv, inv)])
    if np.array_equal(inv, np.arange(A.n)):
        Lg = np.array(c["chol_natural"]).reshape((A.n, A.n), order="F")
        assert np.abs(L - Lg).max() < 1e-14
    for s in range(sym["nnodes"]):
        rl = sym["rlist"][sym["rptr"][s] - 1:sym["rptr"][s + 1] - 1] - 1
        cols = np.arange(sym["sptr"][s] - 1, sym["sptr"][s + 1] - 1)
        got = O.node_factor(sym, off, lcol, s)
        assert np.abs(np.tril(got) - L[np.ix_(rl, cols)]).max() < 1e-14


@pytest.mark.parametrize("grid", [(6, 6, 1), (5, 4, 3), (7, 7, 1)])
def test_multifrontal_nd_orders(S, O, grid):
    A = M.laplacian_3d(*grid) if grid[2] > 1 else M.laplacian_2d(grid[0], grid[1])
    for order in (None, M.grid_nd_order(*grid)):
        sym = _analyse(S, A, order=order)
        st, off, lcol = O.factor_posdef(sym, A.val, nb=4)
        assert st == 0
        D = A.to_dense()
        inv = sym["invp"] - 1
        L = np.linalg.cholesky(D[np.ix_(inv, inv)])
        for s in range(sym["nnodes"]):
            rl = sym["rlist"][sym["rptr"][s] - 1:sym["rptr"][s + 1] - 1] - 1
            cols = np.arange(sym["sptr"][s] - 1, sym["sptr"][s + 1] - 1)
            assert np.abs(np.tril(O.node_factor(sym, off, lcol, s)) - L[np.ix_(rl, cols)]).max() < 1e-13


def test_multifrontal_scaled_equals_dense_cholesky_of_sas(S, O):
    """With a scaling s (pivot order in the oracle, the reference's
    fkeep%scaling), the factors are the Cholesky factor of P S A S P^T: a
    wrong row/column index into s fails this."""
    A = M.laplacian_3d(5, 4, 3)
    sym = _analyse(S, A, order=M.grid_nd_order(5, 4, 3))
    s = np.random.default_rng(7).uniform(0.25, 4.0, A.n)
    inv = sym["invp"] - 1
    st, off, lcol = O.factor_posdef(sym, A.val, scaling=s[inv], nb=4)
    assert st == 0
    D = A.to_dense() * np.outer(s, s)
    L = np.linalg.cholesky(D[np.ix_(inv, inv)])
    for k in range(sym["nnodes"]):
        rl = sym["rlist"][sym["rptr"][k] - 1:sym["rptr"][k + 1] - 1] - 1
        cols = np.arange(sym["sptr"][k] - 1, sym["sptr"][k + 1] - 1)
        assert np.abs(np.tril(O.node_factor(sym, off, lcol, k)) - L[np.ix_(rl, cols)]).max() < 1e-13


def test_indef_scaled_solves_unscaled_system(S, O):
    """Indefinite oracle with a scaling: inertia unchanged (S A S is congruent
    to A) and x = S (S A S)^{-1} S b solves A x = b; with s applied to the
    wrong index the solve fails this."""
    A = M.kkt(6, 40, seed=4)
    sym = _analyse(S, A)
    s = np.random.default_rng(8).uniform(0.25, 4.0, A.n)
    inv = sym["invp"] - 1
    F = O.IndefFactor(sym, A.val, scaling=s[inv])
    assert F.stats["flag"] == 0 and F.stats["num_neg"] == 40
    b = np.random.default_rng(9).uniform(-1, 1, (A.n, 2))
    xp = F.solve(b[inv] * s[inv][:, None]) * s[inv][:, None]
    x = np.empty_like(xp)
    x[inv] = xp
    assert O.bwd_err(A, x, b) < 1e-15


@pytest.mark.parametrize("name", ["simple_mat", "c_example", "lap2d_4", "lap3d_3"])
def test_solve_known_answers(S, O, name):
    c = _case(name)
    A = _csc(c)
    sym = _analyse(S, A)
    st, off, lcol = O.factor_posdef(sym, A.val)
    assert st == 0
    b = np.array(c["b"])
    xp = O.solve_posdef(sym, off, lcol, b[sym["invp"] - 1])[:, 0]
    x = np.empty(A.n)
    x[sym["invp"] - 1] = xp
    np.testing.assert_allclose(x, c["x"], rtol=0, atol=1e-14)
    assert O.bwd_err(A, x, b) < 1e-16 * 10


def test_bwd_err_matches_definition(O):
    A = M.laplacian_3d(4)
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, (A.n, 3))
    b = rng.uniform(-1, 1, (A.n, 3))
    D = A.to_dense()
    r = b - D @ x
    anorm = np.abs(D).sum(axis=1).max()
    want = max(np.abs(r[:, k]).max() / (anorm * np.abs(x[:, k]).max() + np.abs(b[:, k]).max())
               for k in range(3))
    assert abs(O.bwd_err(A, x, b) - want) <= 1e-15 * want
    assert abs(M.backward_error(A, x, b) - want) <= 1e-14 * want


def test_build_map_invariants(S, O):
    """Every stored entry of A appears exactly once; its position decodes to
    the pivot pair of the entry; the library's map equals the oracle's."""
    A = M.laplacian_3d(5)
    sym = _analyse(S, A)
    nptr, nlist = O.build_map(A, sym["order"], sym["invp"], sym)
    assert np.array_equal(nptr, sym["nptr"]) and np.array_equal(nlist, sym["nlist"])
    pairs = nlist.reshape(-1, 2)
    assert sorted(pairs[:, 0].tolist()) == list(range(1, A.nnz + 1))
    cols = np.repeat(np.arange(A.n), np.diff(A.ptr))
    rows = A.row - 1
    order = sym["order"]
    for node in range(sym["nnodes"]):
        m = sym["rptr"][node + 1] - sym["rptr"][node]
        rl = sym["rlist"][sym["rptr"][node] - 1:sym["rptr"][node + 1] - 1]
        for p in range(sym["nptr"][node] - 1, sym["nptr"][node + 1] - 1):
            src, pos = pairs[p]
            c, r = divmod(int(pos) - 1, int(m))
            assert c < sym["sptr"][node + 1] - sym["sptr"][node]
            got = {int(rl[r]), int(sym["sptr"][node] + c)}
            want = {int(order[rows[src - 1]]), int(order[cols[src - 1]])}
            assert got == want


def test_child_maps_and_tree_invariants(S, O):
    A = M.laplacian_3d(6)
    sym = _analyse(S, A)
    tree = O.symbolic_tree(sym)
    nn = sym["nnodes"]
    par = sym["sparent"] - 1
    assert np.array_equal(tree["parent"][:nn], par)
    # children listed in decreasing index order (prepend order)
    for p in range(nn + 1):
        ch, c = [], tree["first_child"][p]
        while c != -1:
            ch.append(c)
            c = tree["next_child"][c]
        assert ch == sorted([i for i in range(nn) if min(par[i], nn) == p], reverse=True)
    # least_desc is the smallest index in the subtree
    for i in range(nn):
        sub = [i]
        k = 0
        while k < len(sub):
            sub += [c for c in range(nn) if par[c] == sub[k]]
            k += 1
        assert tree["least_desc"][i] == min(sub)
    m = np.diff(sym["rptr"])
    n = np.diff(sym["sptr"])
    assert tree["nfactor"] == int((m * n).sum())
    assert tree["maxfront"] == int(m.max())
    off, cmap = O.child_maps(A.n, sym)
    for c in range(nn):
        p = par[c]
        if p >= nn:
            continue
        rlc = sym["rlist"][sym["rptr"][c] - 1:sym["rptr"][c + 1] - 1]
        rlp = sym["rlist"][sym["rptr"][p] - 1:sym["rptr"][p + 1] - 1]
        mc = cmap[off[c]:off[c + 1]]
        assert np.array_equal(rlp[mc], rlc[n[c]:])
        assert np.all(np.diff(mc) > 0)


# ---- indefinite path (APP LDL^T with delays) ------------------------------
def _dinv_to_D(O, dinv, ne):
    D = np.zeros((ne, ne))
    for i, w, blk in O.dinv_blocks(dinv, ne):
        D[i:i + w, i:i + w] = np.linalg.inv(blk) if blk.any() else 0.0
    return D


@pytest.mark.parametrize("m,n,nb,scale,zero", [
    (40, 40, 16, 1.0, 0), (60, 25, 8, 1.0, 0), (120, 50, 16, 1.0, 0), (90, 90, 128, 1.0, 0),
    (80, 40, 16, 1e-2, 0), (50, 50, 16, 1e-3, 25), (130, 130, 64, 1.0, 60)])
def test_front_ldlt_closed_form(O, m, n, nb, scale, zero):
    """P A P^T restricted to the eliminated pivots equals L D L^T; the failed
    columns and the contribution block hold the Schur complement; every
    entry of L respects the threshold |l| <= 1/u; the eliminated pivots'
    inertia agrees with the Schur complement reduction (Sylvester)."""
    u = 0.01
    rng = np.random.default_rng(m * 7 + n)
    a = rng.uniform(-1, 1, (m, m))
    a = a + a.T
    a[:n, :n] *= scale
    if zero:
        a[:zero, :zero] = 0.0
    ne, nf, perm, dinv, f = O.front_ldlt(a, n, nb=nb, u=u)
    assert 0 <= nf <= ne <= n
    assert sorted(perm.tolist()) == list(range(n))
    P = np.concatenate([perm, np.arange(n, m)])
    Ap = a[np.ix_(P, P)]
    L = np.tril(f[:, :ne], -1)
    L[np.arange(ne), np.arange(ne)] = 1.0
    D = _dinv_to_D(O, dinv, ne)
    amax = max(1.0, np.abs(a).max())
    assert np.abs(Ap[:, :ne] - L @ D @ L[:ne].T).max() <= 1e-12 * amax * max(1.0, np.abs(L).max()) ** 2
    S = Ap[ne:, ne:] - L[ne:] @ D @ L[ne:].T
    if n > ne:
        assert np.abs(f[ne:, ne:n] - S[:, :n - ne]).max() <= 1e-11 * amax * max(1.0, np.abs(L).max()) ** 2
    if m > n:
        assert np.abs(np.tril(f[n:, n:]) - np.tril(S[n - ne:, n - ne:])).max() <= 1e-11 * amax * max(1.0, np.abs(L).max()) ** 2
    assert np.abs(L).max() <= 1.0 / u * (1 + 1e-12)
    if m == n and ne == n:
        ev = np.linalg.eigvalsh(a)
        assert (np.linalg.eigvalsh(D) < 0).sum() == (ev < 0).sum()


def test_front_ldlt_failures_and_second_pass(O):
    """A weak fully-summed block fails the a-posteriori test (nelim_first <
    n); threshold partial pivoting recovers part of it, the rest is delayed;
    with m == n (a root) everything is eliminated."""
    rng = np.random.default_rng(5)
    a = rng.uniform(-1, 1, (80, 80))
    a = a + a.T
    a[:40, :40] *= 1e-2
    ne, nf, perm, dinv, f = O.front_ldlt(a, 40, nb=16)
    assert nf < ne < 40
    ne_nt, nf_nt, *_ = O.front_ldlt(a, 40, nb=16, failed_tpp=False)
    assert ne_nt == nf_nt == nf
    b = a[:40, :40] + np.diag(np.zeros(40))
    ne2, _, _, _, _ = O.front_ldlt(b, 40, nb=16)
    assert ne2 == 40


def test_front_ldlt_singular_action(O):
    a = np.array([[-1.0, 2.0, 0.0], [2.0, 1.0, 0.0], [0.0, 0.0, 0.0]])
    assert O.front_ldlt(a, 3, action=False)[0] == -5
    ne, _, perm, dinv, _ = O.front_ldlt(a, 3, action=True)
    assert ne == 3
    zero = [i for i, w, blk in O.dinv_blocks(dinv, ne) if w == 1 and blk[0, 0] == 0.0]
    assert len(zero) == 1


@pytest.mark.parametrize("side,mrows,nb", [(4, 12, 64), (6, 40, 64), (8, 100, 16), (10, 250, 32)])
def test_indef_tree_inertia_and_solve(S, O, side, mrows, nb):
    """KKT [H A^T; A 0] (H SPD, A full row rank): inertia (nh, m, 0) by
    Sylvester's law; the solve matches numpy.linalg.solve; delays occur."""
    A = M.kkt(side, mrows, seed=side)
    sym = _analyse(S, A)
    F = O.IndefFactor(sym, A.val, nb=nb)
    st = F.stats
    assert st["flag"] == 0 and st["num_neg"] == mrows and st["num_zero"] == 0
    D = A.to_dense()
    rng = np.random.default_rng(1)
    b = rng.uniform(-1, 1, (A.n, 2))
    inv = sym["invp"] - 1
    xp = F.solve(b[inv])
    x = np.empty_like(xp)
    x[inv] = xp
    np.testing.assert_allclose(x, np.linalg.solve(D, b), rtol=0, atol=1e-10 * np.abs(x).max())
    assert O.bwd_err(A, x, b) < 1e-15
    if side >= 8:
        assert st["num_delay"] > 0


def test_indef_tree_partial_solves(S, O):
    A = M.kkt(6, 40, seed=3)
    sym = _analyse(S, A)
    F = O.IndefFactor(sym, A.val)
    b = np.random.default_rng(2).uniform(-1, 1, (A.n, 1))
    full = F.solve(b, job=0)
    y = F.solve(F.solve(F.solve(b, job=1), job=2), job=3)
    np.testing.assert_allclose(y, full, rtol=0, atol=1e-13)
    z = F.solve(F.solve(b, job=1), job=4)
    np.testing.assert_allclose(z, full, rtol=0, atol=1e-13)


def test_indef_tree_node_structure(S, O):
    """Every front: rows are distinct variables; eliminated pivots over all
    nodes cover every variable exactly once; delayed columns reappear as
    fully-summed rows of the parent."""
    A = M.kkt(8, 100, seed=3)
    sym = _analyse(S, A)
    F = O.IndefFactor(sym, A.val, nb=16)
    seen = []
    par = sym["sparent"] - 1
    nn = sym["nnodes"]
    for i in range(nn):
        nd = F.node(i)
        assert len(set(nd["rows"].tolist())) == nd["m"]
        seen += nd["rows"][:nd["nelim"]].tolist()
        fails = set(nd["rows"][nd["nelim"]:nd["n"]].tolist())
        if fails:
            assert par[i] < nn
            pn = F.node(par[i])
            assert fails <= set(pn["rows"][:pn["n"]].tolist())
    assert sorted(seen) == list(range(1, A.n + 1))


def test_mlp_oracle_gradient_finite_differences():
    """The cost model's MLP oracle: analytic gradient equals central finite
    differences (FP64) on a random batch."""
    from oracle import mlp
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, mlp.NPARAMS)
    x = rng.normal(0, 1, (13, 8))
    y = rng.normal(0, 1, (13, 2))
    g = mlp.grad(p, x, y)
    idx = rng.choice(mlp.NPARAMS, 60, replace=False)
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (mlp.loss(p + e, x, y) - mlp.loss(p - e, x, y)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])
