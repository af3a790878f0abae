"""Synthetic sparse symmetric matrices and orderings used by the tests and the
benchmark.

All matrices are returned in the reference's input format: the lower triangle
in compressed sparse column form with 1-based indices, ``ptr`` int64 of length
n+1 and ``row`` int32 (``include/sylver/sylver.h:73-93``, the C example
``examples/C/spldlt_simple_example_c.c:30-40`` passes 1-based arrays).

Orderings follow the reference's ``order`` convention: ``order[i]`` (1-based)
is the pivot position of variable ``i+1``.
"""
from __future__ import annotations

import numpy as np


class CSCLower:
    """Lower triangle of a symmetric matrix, CSC, 1-based indices."""

    def __init__(self, n, ptr, row, val):
        self.n = int(n)
        self.ptr = np.ascontiguousarray(ptr, dtype=np.int64)
        self.row = np.ascontiguousarray(row, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)

    @property
    def nnz(self):
        return int(self.ptr[-1] - 1)

    def to_dense(self):
        a = np.zeros((self.n, self.n))
        for j in range(self.n):
            for p in range(self.ptr[j] - 1, self.ptr[j + 1] - 1):
                i = self.row[p] - 1
                a[i, j] += self.val[p]
                if i != j:
                    a[j, i] += self.val[p]
        return a

    def to_scipy_full(self):
        import scipy.sparse as sp
        n = self.n
        cols = np.repeat(np.arange(n), np.diff(self.ptr))
        rows = self.row.astype(np.int64) - 1
        low = sp.csc_matrix((self.val, (rows, cols)), shape=(n, n))
        off = sp.csc_matrix((np.where(rows != cols, self.val, 0.0), (rows, cols)), shape=(n, n))
        return (low + off.T).tocsr()


def _from_coo_lower(n, rows0, cols0, vals):
    """Build 1-based CSC lower from 0-based coordinate lists (rows >= cols)."""
    rows0 = np.asarray(rows0, dtype=np.int64)
    cols0 = np.asarray(cols0, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    key = cols0 * n + rows0
    o = np.argsort(key, kind="stable")
    rows0, cols0, vals = rows0[o], cols0[o], vals[o]
    cnt = np.bincount(cols0, minlength=n)
    ptr = np.empty(n + 1, dtype=np.int64)
    ptr[0] = 1
    np.cumsum(cnt, out=ptr[1:])
    ptr[1:] += 1
    return CSCLower(n, ptr, rows0 + 1, vals)


def laplacian_2d(nx, ny=None):
    """5-point Laplacian on an nx-by-ny grid: diagonal 4, off-diagonals -1."""
    ny = nx if ny is None else ny
    return laplacian_3d(nx, ny, 1, diag=4.0)


def laplacian_3d(nx, ny=None, nz=None, diag=None):
    """7-point Laplacian on an nx*ny*nz grid: diagonal 6, off-diagonals -1.

    Variable (x, y, z) has 0-based index x + nx*(y + ny*z)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    if diag is None:
        diag = 6.0 if nz > 1 else 4.0
    n = nx * ny * nz
    idx = np.arange(n, dtype=np.int64)
    x = idx % nx
    y = (idx // nx) % ny
    z = idx // (nx * ny)
    rows = [idx]
    cols = [idx]
    vals = [np.full(n, diag)]
    for cond, step in ((x < nx - 1, 1), (y < ny - 1, nx), (z < nz - 1, nx * ny)):
        c = idx[cond]
        rows.append(c + step)
        cols.append(c)
        vals.append(np.full(c.size, -1.0))
    return _from_coo_lower(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def grid_nd_order(nx, ny=None, nz=None, leaf=8):
    """Geometric nested dissection of a structured grid.

    Recursively cuts the longest side of the box at its middle plane; the two
    halves are ordered first, the separator plane last. Boxes with at most
    ``leaf`` points are ordered naturally. Returns ``order`` (1-based pivot
    position of each variable), the convention of ``spldlt_analyse``."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    n = nx * ny * nz
    seq = np.empty(n, dtype=np.int64)
    pos = 0
    stack = [((0, nx), (0, ny), (0, nz), False)]
    # iterative postorder: push (box, expanded) so halves are emitted before the separator
    out_chunks = []
    while stack:
        bx, by, bz, sep = stack.pop()
        dims = [bx[1] - bx[0], by[1] - by[0], bz[1] - bz[0]]
        if sep or dims[0] * dims[1] * dims[2] <= leaf or max(dims) < 3:
            xs = np.arange(bx[0], bx[1])
            ys = np.arange(by[0], by[1])
            zs = np.arange(bz[0], bz[1])
            if xs.size and ys.size and zs.size:
                zz, yy, xx = np.meshgrid(zs, ys, xs, indexing="ij")
                out_chunks.append((xx + nx * (yy + ny * zz)).ravel())
            continue
        d = int(np.argmax(dims))
        b = [bx, by, bz]
        lo, hi = b[d]
        mid = (lo + hi) // 2
        left = list(b)
        left[d] = (lo, mid)
        right = list(b)
        right[d] = (mid + 1, hi)
        sepb = list(b)
        sepb[d] = (mid, mid + 1)
        # stack is LIFO: push separator first so it is emitted last
        stack.append((sepb[0], sepb[1], sepb[2], True))
        stack.append((right[0], right[1], right[2], False))
        stack.append((left[0], left[1], left[2], False))
    for c in out_chunks:
        seq[pos:pos + c.size] = c
        pos += c.size
    assert pos == n
    order = np.empty(n, dtype=np.int32)
    order[seq] = np.arange(1, n + 1, dtype=np.int32)
    return order


def kkt(nh_side, m_rows, nnz_per_row=3, seed=0):
    """Saddle-point matrix [H A^T; A 0]: H is the 7-point Laplacian on an
    nh_side^3 grid, A is m_rows x nH with ``nnz_per_row`` entries per row at
    uniform random columns, values uniform in [-1, 1]. The (2,2) block has no
    stored entries."""
    rng = np.random.default_rng(seed)
    H = laplacian_3d(nh_side)
    nh = H.n
    n = nh + m_rows
    hcols = np.repeat(np.arange(nh, dtype=np.int64), np.diff(H.ptr))
    hrows = H.row.astype(np.int64) - 1
    arow = np.repeat(np.arange(m_rows, dtype=np.int64), nnz_per_row)
    acol = np.empty(m_rows * nnz_per_row, dtype=np.int64)
    for r in range(m_rows):
        acol[r * nnz_per_row:(r + 1) * nnz_per_row] = rng.choice(nh, nnz_per_row, replace=False)
    aval = rng.uniform(-1.0, 1.0, m_rows * nnz_per_row)
    rows = np.concatenate([hrows, nh + arow])
    cols = np.concatenate([hcols, acol])
    vals = np.concatenate([H.val, aval])
    return _from_coo_lower(n, rows, cols, vals)


def random_rhs(n, nrhs=1, seed=0):
    """Right-hand sides with entries uniform in [-1, 1], column-major n x nrhs."""
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.uniform(-1.0, 1.0, size=(n, nrhs)))


def dense_front_sizes(nfronts, seed=0, nmin=8, nmax=256, rmax=2.0):
    """Front shapes of the batched small-front workload: ncol log-uniform in
    [nmin, nmax], nrow = ncol*(1+r) with r uniform in [0, rmax]."""
    rng = np.random.default_rng(seed)
    ncol = np.exp(rng.uniform(np.log(nmin), np.log(nmax + 1), nfronts)).astype(np.int64)
    ncol = np.clip(ncol, nmin, nmax)
    r = rng.uniform(0.0, rmax, nfronts)
    nrow = np.floor(ncol * (1.0 + r)).astype(np.int64)
    nrow = np.maximum(nrow, ncol)
    return nrow.astype(np.int32), ncol.astype(np.int32)


def dense_front_values(m, seed):
    """One dense SPD front: entries uniform in [-1, 1], symmetrized, plus m
    on the diagonal. Column-major m x m."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1.0, 1.0, size=(m, m))
    a = 0.5 * (a + a.T) + m * np.eye(m)
    return np.asfortranarray(a)


def backward_error(A: CSCLower, x, b):
    """||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) per right-hand side,
    full symmetric row sums; the worst over columns is returned
    (tests/common.hxx:1003-1050)."""
    S = A.to_scipy_full()
    x = np.asarray(x).reshape(A.n, -1)
    b = np.asarray(b).reshape(A.n, -1)
    r = b - S @ x
    anorm = np.abs(S).sum(axis=1).max()
    worst = 0.0
    for k in range(x.shape[1]):
        e = np.abs(r[:, k]).max() / (anorm * np.abs(x[:, k]).max() + np.abs(b[:, k]).max())
        worst = max(worst, float(e))
    return worst
