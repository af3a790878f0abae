"""Dense front LU of the unsymmetric path (the reference's lu_nopiv_factor,
lu_pp_factor, lu_tpp_factor: src/kernels/lu_nopiv.hxx:13-41,
src/kernels/lu_pp.hxx:38-73, src/kernels/lu_tpp.hxx:33-118).

CPU part: the oracle's restatement against what the mathematics fixes --
P A Q = L U + S (S the Schur complement in the trailing block) for every
mode, the partial-pivoting multipliers bounded by 1, the threshold mode's
multipliers by 1/u within the pivot column's rows, no pivoting equal to
numpy's unpivoted elimination, and the backward error of the full solve
within the reference's bar (tests/testing_lu_tpp.hxx: bwderr <= 5e-14).

GPU part: the device kernel (sylver_amd_batched_front_lu) bit for bit
equal to the oracle -- factors, row and column orders, nelim -- on the
reference's case lists (tests/drivers/kernel_test.cxx:52-69: orders 8 to
256, diagonally dominant and general; here up to the kernel's limit of 128
rows) and on a batch of ragged rectangular fronts, in FP64 and FP32.

Test matrices restate tests/common.hxx: gen_mat (entries U[-1, 1]) and
gen_unsym_diagdom (:575-586: |a_ii| + 0.1 plus the row's off-diagonal
absolute sum). The reference's unsym_backward_error (:624-680) sums signed
column entries for its norm; the bar here uses the proper scaled backward
error ||b - Ax||_inf / (||A||_inf ||x||_inf + ||b||_inf).
"""
import numpy as np
import pytest


def gen_mat(m, rng):
    return rng.uniform(-1.0, 1.0, (m, m))


def gen_unsym_diagdom(m, rng):
    a = gen_mat(m, rng)
    d = np.abs(np.diag(a)) + 0.1 + (np.abs(a).sum(axis=1) - np.abs(np.diag(a)))
    a[np.arange(m), np.arange(m)] = d
    return a


def bits(x):
    return np.asfortranarray(x).tobytes(order="F")


def reconstruct(f, ne):
    m = f.shape[0]
    L = np.tril(f[:, :ne], -1) + np.eye(m)[:, :ne]
    U = np.triu(f[:ne, :])
    S = f.copy()
    S[:, :ne] = 0
    S[:ne, :] = 0
    return L @ U + S


def lu_solve_bwd_err(a, f, rp, cp, b):
    """x from the full factorization (nelim == m) of P A Q."""
    import scipy.linalg as sl
    L = np.tril(f, -1) + np.eye(f.shape[0])
    U = np.triu(f)
    y = sl.solve_triangular(L, b[rp], lower=True, unit_diagonal=True)
    z = sl.solve_triangular(U, y, lower=False)
    x = np.empty_like(z)
    x[cp] = z
    r = b - a @ x
    return np.abs(r).max() / (np.abs(a).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())


@pytest.mark.parametrize("mode", ["nopiv", "pp", "tpp"])
@pytest.mark.parametrize("m,k", [(8, 8), (32, 32), (64, 40), (100, 100), (128, 17), (256, 256)])
def test_oracle_lu_reconstructs(O, mode, m, k):
    rng = np.random.default_rng(m * 1000 + k)
    a = gen_unsym_diagdom(m, rng) if mode == "nopiv" else gen_mat(m, rng)
    ne, rp, cp, f = O.lu_front(a, k, mode, u=0.01)
    if mode != "tpp":
        assert ne == k
        assert np.array_equal(cp, np.arange(m))
    assert sorted(rp) == list(range(m)) and sorted(cp) == list(range(m))
    pa = a[rp][:, cp]
    err = np.abs(reconstruct(f, ne) - pa).max() / np.abs(a).max()
    assert err <= (1e-13 if mode != "tpp" else 1e-11), err
    if mode == "pp":
        assert np.abs(np.tril(f[:, :ne], -1)).max() <= 1.0
    if mode == "tpp" and ne > 0:
        # each pivot passed the threshold test against its column below
        assert np.abs(np.tril(f[:, :ne], -1)).max() <= 1.0 / 0.01 * (1 + 1e-12)


def test_oracle_nopiv_matches_numpy_elimination(O):
    rng = np.random.default_rng(3)
    m = 24
    a = gen_unsym_diagdom(m, rng)
    ne, rp, cp, f = O.lu_front(a, m, "nopiv")
    # textbook Doolittle without pivoting
    w = a.copy()
    for e in range(m):
        w[e + 1:, e] /= w[e, e]
        w[e + 1:, e + 1:] -= np.outer(w[e + 1:, e], w[e, e + 1:])
    assert np.allclose(f, w, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("mode,diagdom", [("nopiv", True), ("pp", True), ("pp", False), ("tpp", True),
                                          ("tpp", False)])
@pytest.mark.parametrize("m", [8, 32, 128])
def test_oracle_full_solve_bwd_err(O, mode, diagdom, m):
    """The reference's check (tests/testing_lu_tpp.hxx: solution of A x = b
    with b = A 1 through the factors, bwderr <= 5e-14)."""
    rng = np.random.default_rng(m + (7 if diagdom else 0))
    a = gen_unsym_diagdom(m, rng) if diagdom else gen_mat(m, rng)
    b = a.sum(axis=1)
    ne, rp, cp, f = O.lu_front(a, m, mode)
    assert ne == m
    assert lu_solve_bwd_err(a, f, rp, cp, b) <= 5e-14


def test_oracle_tpp_delays_a_column(O):
    """A zero diagonal with an acceptable pivot only below the fully-summed
    rows: threshold pivoting leaves the column for later (nelim < k) where
    partial pivoting takes the row from below."""
    a = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    a[2, 0] = 5.0
    ne, rp, cp, f = O.lu_front(a, 2, "tpp", u=0.5)
    pa = a[rp][:, cp]
    assert np.allclose(reconstruct(f, ne), pa)
    ne2, rp2, _, _ = O.lu_front(a, 2, "pp")
    assert ne2 == 2 and rp2[0] == 2


# ---------------------------------------------------------------------------
# device kernel vs the oracle
# ---------------------------------------------------------------------------

def _device_lu(S, fronts, ks, mode, dtype, u=0.01):
    offs, o = [], 0
    for a in fronts:
        offs.append(o)
        o += a.size
    flat = np.concatenate([np.asfortranarray(a).ravel(order="F") for a in fronts]).astype(dtype)
    dev = S.DeviceArray(flat)
    rc, ne, perms = S.batched_front_lu(dev, [a.shape[0] for a in fronts], ks, offs, dtype, mode, u)
    out = dev.numpy()
    dev.free()
    res = []
    for a, off in zip(fronts, offs):
        m = a.shape[0]
        res.append(out[off:off + m * m].reshape((m, m), order="F"))
    return rc, ne, perms, res


# the reference's (commented) case lists, tests/drivers/kernel_test.cxx:52-69
REF_CASES = [("nopiv", 32, True), ("nopiv", 128, True), ("nopiv", 32, False), ("nopiv", 128, False),
             ("pp", 8, True), ("pp", 16, True), ("pp", 32, True), ("pp", 128, True), ("pp", 32, False),
             ("pp", 128, False), ("tpp", 8, True), ("tpp", 32, False), ("tpp", 128, False)]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("mode,m,diagdom", REF_CASES)
def test_device_lu_reference_cases(S, O, gpu, dtype, mode, m, diagdom):
    rng = np.random.default_rng(m * 10 + (1 if diagdom else 0))
    a = gen_unsym_diagdom(m, rng) if diagdom else gen_mat(m, rng)
    a = a.astype(dtype)
    rc, ne, perms, got = _device_lu(S, [a], [m], mode, dtype)
    assert rc == 0
    ne_o, rp_o, cp_o, f_o = O.lu_front(a, m, mode)
    assert ne[0] == ne_o
    assert np.array_equal(perms[0][0], rp_o) and np.array_equal(perms[0][1], cp_o)
    assert bits(got[0]) == bits(f_o)
    if dtype == np.float64 and ne_o == m and (diagdom or mode != "nopiv"):
        a64 = a.astype(np.float64)
        assert lu_solve_bwd_err(a64, got[0], perms[0][0], perms[0][1], a64.sum(axis=1)) <= 5e-14


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["nopiv", "pp", "tpp"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_device_lu_ragged_batch(S, O, gpu, mode, dtype):
    """300 fronts of orders 1..128 with 0 <= k <= m in one launch, each bit
    for bit the oracle's, and P A Q = L U + S."""
    rng = np.random.default_rng(11)
    ms = rng.integers(1, 129, 300)
    ks = np.array([rng.integers(0, mm + 1) for mm in ms])
    fronts = [(gen_unsym_diagdom(mm, rng) if mode == "nopiv" else gen_mat(mm, rng)).astype(dtype) for mm in ms]
    rc, ne, perms, got = _device_lu(S, fronts, ks, mode, dtype, u=0.1)
    assert rc == 0
    for a, k, n_d, (rp, cp), f in zip(fronts, ks, ne, perms, got):
        ne_o, rp_o, cp_o, f_o = O.lu_front(a, int(k), mode, u=0.1)
        assert n_d == ne_o
        assert np.array_equal(rp, rp_o) and np.array_equal(cp, cp_o)
        assert bits(f) == bits(f_o)
        if dtype == np.float64:
            pa = a[rp][:, cp]
            assert np.abs(reconstruct(f, n_d) - pa).max() <= 1e-10 * max(1.0, np.abs(a).max())


@pytest.mark.gpu
def test_device_lu_rejects_oversized_front(S, gpu):
    a = np.zeros((129, 129))
    rc, _, _, _ = _device_lu(S, [a], [129], "pp", np.float64)
    assert rc == S.ERROR_CALL_SEQUENCE
