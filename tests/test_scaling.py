"""Matching-based scalings and ordering (host code of the library; CPU only)
and the equilibration oracle.

The scalings solve the assignment problem max prod |a_i,sigma(i)| over
permutations (MC64's objective) and scale symmetrically from its dual
variables. Pinned here by:
- scipy.optimize.linear_sum_assignment on the dense cost -log|a_ij| (the
  same assignment problem): the library's matching reaches the optimum
  (shortest augmenting paths) or lies within n * eps of it (auction);
- dual feasibility, which bounds every |s_i a_ij s_j| by 1;
- structural singularity flags (SPRAL's semantics as the reference uses
  them: src/spldlt_factorize_mod.F90:734-760);
- matching-based ordering: a permutation in which every 2x2 candidate pair
  is adjacent and joined by a nonzero entry.
The equilibration oracle (oracle/scaling.py) is pinned by its fixed point:
every row of |S A S| has largest entry 1 within the tolerance.
"""
import numpy as np
import pytest
from scipy.optimize import linear_sum_assignment

from sylver_amd import matrices as M
from oracle import scaling as OS


def _random_sym(n, density, seed, zero_diag_frac=0.0):
    rng = np.random.default_rng(seed)
    d = rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < density)
    d = np.tril(d) + np.tril(d, -1).T
    f = np.exp(rng.uniform(-6, 6, n))  # badly scaled, still symmetric
    d *= np.outer(f, f)
    diag = rng.uniform(-2, 2, n) * f * f
    diag[rng.random(n) < zero_diag_frac] = 0.0
    np.fill_diagonal(d, diag)
    # a permutation-like backbone keeps it structurally nonsingular
    p = rng.permutation(n)
    for i in range(n):
        j = p[i]
        if d[i, j] == 0.0:
            d[i, j] = d[j, i] = rng.uniform(0.1, 1.0) * f[i] * f[j]
    rows, cols = np.nonzero(np.tril(d))
    return M._from_coo_lower(n, rows, cols, d[rows, cols]), d


def _opt_logprod(d):
    with np.errstate(divide="ignore"):
        c = -np.log(np.abs(d))
    c[~np.isfinite(c)] = 1e6
    r, k = linear_sum_assignment(c)
    return -c[r, k].sum()


def _logprod(d, match):
    return float(np.sum(np.log(np.abs(d[np.arange(len(match)), match]))))


@pytest.mark.parametrize("n,density,seed,zd", [(30, 0.2, 1, 0.0), (60, 0.08, 2, 0.5), (80, 0.05, 3, 1.0),
                                               (120, 0.04, 4, 0.3)])
def test_hungarian_optimal_and_bounded(S, n, density, seed, zd):
    A, d = _random_sym(n, density, seed, zd)
    flag, s, m = S.hungarian_scale_sym(A)
    assert flag == 0
    assert sorted(m.tolist()) == list(range(n))
    assert abs(_logprod(d, m) - _opt_logprod(d)) <= 1e-9 * max(1.0, abs(_opt_logprod(d)))
    assert np.all(s > 0) and np.all(np.isfinite(s))
    mx, rows = OS.scaled_abs_max(A, s)
    assert mx <= 1 + 1e-12
    # every row of a nonsingular matrix keeps an entry of (near) 1: the
    # scaled matched entries of 2-cycles are exactly the dual equalities
    assert rows.min() > 1e-3


def test_hungarian_scales_badly_scaled_kkt(S):
    """KKT [H A^T; A 0] with rows of A scaled by 1e-8..1e8: the matching
    pairs constraint rows with H columns (zero (2,2) block) and the scaled
    matrix has all entries <= 1 with every row's largest entry near 1."""
    A = M.kkt(5, 30, seed=1)
    rng = np.random.default_rng(2)
    f = np.ones(A.n)
    f[125:] = 10.0 ** rng.uniform(-8, 8, 30)
    col = np.repeat(np.arange(A.n), np.diff(A.ptr))
    row = A.row - 1
    B = M.CSCLower(A.n, A.ptr, A.row, A.val * f[row] * f[col])
    flag, s, m = S.hungarian_scale_sym(B)
    assert flag == 0
    d = B.to_dense()
    assert abs(_logprod(d, m) - _opt_logprod(d)) <= 1e-8 * abs(_opt_logprod(d))
    mx, rows = OS.scaled_abs_max(B, s)
    assert mx <= 1 + 1e-12 and rows.min() > 1e-3


def test_hungarian_structurally_singular(S):
    # variables 2 and 3 only connect to variable 1: a rank-2 pattern
    A = M.CSCLower(3, [1, 4, 4, 4], [1, 2, 3], [1.0, 2.0, 3.0])
    flag, s, m = S.hungarian_scale_sym(A, scale_if_singular=False)
    assert flag == -2
    flag, s, m = S.hungarian_scale_sym(A, scale_if_singular=True)
    assert flag == 1
    assert np.all(s > 0) and np.all(np.isfinite(s))
    assert (m < 0).sum() == 1


@pytest.mark.parametrize("n,density,seed,zd", [(30, 0.2, 1, 0.0), (60, 0.08, 2, 0.5), (100, 0.05, 5, 1.0)])
def test_auction_near_optimal_and_bounded(S, n, density, seed, zd):
    A, d = _random_sym(n, density, seed, zd)
    flag, s, m = S.auction_scale_sym(A)
    assert flag == 0
    assert sorted(m.tolist()) == list(range(n))
    # epsilon-optimal: within n * eps_final (1e-4) of the optimum
    assert _logprod(d, m) >= _opt_logprod(d) - n * 1e-4 - 1e-9
    mx, rows = OS.scaled_abs_max(A, s)
    assert mx <= 1 + 1e-12 and rows.min() > 1e-3


def test_match_order_pairs_adjacent(S):
    A = M.kkt(5, 40, seed=3)
    flag, order, s = S.match_order(A)
    assert flag == 0
    assert sorted(order.tolist()) == list(range(1, A.n + 1))
    _, _, m = S.hungarian_scale_sym(A)
    d = A.to_dense()
    # the matched variables of every constraint row (zero diagonal) sit next
    # to a partner joined by a nonzero entry
    pos = order - 1
    inv = np.empty(A.n, np.int64)
    inv[pos] = np.arange(A.n)
    for v in range(125, A.n):
        nb = [inv[pos[v] + k] for k in (-1, 1) if 0 <= pos[v] + k < A.n]
        assert any(d[v, w] != 0.0 for w in nb), v
    mx, _ = OS.scaled_abs_max(A, s)
    assert mx <= 1 + 1e-12


def test_analyse_matching_order_through_api(S):
    A = M.kkt(6, 60, seed=5)
    opts = S.default_options()
    opts.ordering = 2
    akeep, inf, order = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=opts)
    assert inf.flag == S.WARNING_MISSING_DIAGONAL  # the zero (2,2) block stores no diagonal
    sym = S.akeep_symbolic(akeep)
    assert sorted(sym["order"].tolist()) == list(range(1, A.n + 1))
    # every constraint variable (zero diagonal) is eliminated in a supernode
    # that also holds a variable it couples to (its matched 2x2 partner):
    # the pair stays together after the tree's renumbering
    d = A.to_dense()
    sptr = sym["sptr"]
    node_of = np.searchsorted(sptr, sym["order"], side="right") - 1
    for v in range(216, A.n):
        mates = [w for w in range(A.n) if w != v and d[v, w] != 0.0 and node_of[w] == node_of[v]]
        assert mates, v
    S.spldlt_free_akeep(akeep)
    # values are required for this ordering
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, None, options=opts)
    assert inf.flag == S.ERROR_VAL


@pytest.mark.parametrize("seed", [1, 2])
def test_equilib_oracle_fixed_point(seed):
    A, _ = _random_sym(80, 0.08, seed, 0.3)
    rng = np.random.default_rng(seed)
    f = 10.0 ** rng.uniform(-5, 5, A.n)
    col = np.repeat(np.arange(A.n), np.diff(A.ptr))
    row = A.row - 1
    B = M.CSCLower(A.n, A.ptr, A.row, A.val * f[row] * f[col])
    s, it = OS.equilib_scale_sym(B, max_iter=100)
    mx, rows = OS.scaled_abs_max(B, s)
    assert np.max(np.abs(rows - 1.0)) < 1e-7
    # the equilibrated matrix is a fixed point: one more run keeps s ~ 1
    C = M.CSCLower(B.n, B.ptr, B.row, B.val * s[row] * s[col])
    s2, it2 = OS.equilib_scale_sym(C, max_iter=100)
    assert np.max(np.abs(s2 - 1.0)) < 1e-7
