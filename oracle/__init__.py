"""CPU oracle (test infrastructure only): ctypes access to oracle/liboracle.so.

This package restates the reference's numeric phase (posdef LL^T and APP
LDL^T with delayed pivots) on the CPU in FP64. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
may use it; the library never does. Build it with ``python oracle/build.py``.

Which function each oracle entry restates is stated in oracle/oracle.h.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from . import build as _b  # noqa: F401
            raise RuntimeError(f"oracle library missing: {LIB}; run `python oracle/build.py`")
        L = C.CDLL(LIB)
        i64p, ip, dp = C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.oracle_build_map.argtypes = [C.c_int, i64p, ip, ip, ip, C.c_int, ip, i64p, ip, i64p, i64p]
        L.oracle_symbolic_tree.argtypes = [C.c_int, ip, ip, i64p, ip, ip, ip, ip, ip]
        L.oracle_symbolic_tree.restype = C.c_int64
        L.oracle_child_maps.argtypes = [C.c_int, C.c_int, ip, ip, i64p, ip, i64p, ip]
        L.oracle_factor_posdef.argtypes = [C.c_int, C.c_int, ip, ip, i64p, ip, i64p, i64p, dp, dp,
                                           C.c_int, C.c_int, C.c_char_p, i64p, dp]
        L.oracle_factor_size.argtypes = [C.c_int, ip, i64p]
        L.oracle_factor_size.restype = C.c_int64
        L.oracle_solve_posdef.argtypes = [C.c_int, C.c_int, ip, i64p, ip, i64p, dp, C.c_int, dp,
                                          C.c_int]
        L.oracle_front_potrf.argtypes = [C.c_int, C.c_int, dp, C.c_int, C.c_int]
        L.oracle_bwd_err.argtypes = [C.c_int, i64p, ip, dp, C.c_int, dp, C.c_int, dp, C.c_int]
        L.oracle_bwd_err.restype = C.c_double
        L.oracle_factor_indef.argtypes = [C.c_int, C.c_int, ip, ip, i64p, ip, i64p, i64p, dp, dp,
                                          C.c_int, C.c_double, C.c_double, C.c_int, C.c_int]
        L.oracle_factor_indef.restype = C.c_void_p
        L.oracle_indef_stats.argtypes = [C.c_void_p, ip]
        L.oracle_indef_node.argtypes = [C.c_void_p, C.c_int, ip, ip, dp, dp]
        L.oracle_solve_indef.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, C.c_int]
        L.oracle_indef_free.argtypes = [C.c_void_p]
        L.oracle_front_ldlt.argtypes = [C.c_int, C.c_int, dp, C.c_int, C.c_double, C.c_double,
                                        C.c_int, C.c_int, ip, dp, ip]
        L.oracle_lu_front.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, dp, C.c_int, ip, ip]
        L.oracle_lu_front_f32.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_float),
                                          C.c_int, ip, ip]
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def build_map(A, order, invp, sym):
    nn = sym["nnodes"]
    nptr = np.zeros(nn + 1, np.int64)
    nlist = np.zeros(2 * max(A.nnz, 1), np.int64)
    lib().oracle_build_map(A.n, _p(A.ptr, C.c_int64), _p(A.row, C.c_int),
                           _p(np.ascontiguousarray(order, np.int32), C.c_int),
                           _p(np.ascontiguousarray(invp, np.int32), C.c_int), nn,
                           _p(sym["sptr"], C.c_int), _p(sym["rptr"], C.c_int64),
                           _p(sym["rlist"], C.c_int), _p(nptr, C.c_int64), _p(nlist, C.c_int64))
    return nptr, nlist[:2 * (nptr[-1] - 1)]


def symbolic_tree(sym):
    nn = sym["nnodes"]
    out = {k: np.zeros(nn + 1, np.int32) for k in ("parent", "first_child", "next_child", "least_desc")}
    mf = C.c_int()
    nf = lib().oracle_symbolic_tree(nn, _p(sym["sptr"], C.c_int), _p(sym["sparent"], C.c_int),
                                    _p(sym["rptr"], C.c_int64), _p(out["parent"], C.c_int),
                                    _p(out["first_child"], C.c_int), _p(out["next_child"], C.c_int),
                                    _p(out["least_desc"], C.c_int), C.byref(mf))
    out["nfactor"] = int(nf)
    out["maxfront"] = mf.value
    return out


def child_maps(n, sym):
    nn = sym["nnodes"]
    q = (np.diff(sym["rptr"]) - np.diff(sym["sptr"])).astype(np.int64)
    off = np.zeros(nn + 1, np.int64)
    cmap = np.zeros(max(int(q.sum()), 1), np.int32)
    lib().oracle_child_maps(n, nn, _p(sym["sptr"], C.c_int), _p(sym["sparent"], C.c_int),
                            _p(sym["rptr"], C.c_int64), _p(sym["rlist"], C.c_int),
                            _p(off, C.c_int64), _p(cmap, C.c_int))
    return off, cmap[:int(q.sum())]


def factor_posdef(sym, aval, scaling=None, nb=256, nthreads=0, node_mask=None):
    """Returns (status, lcol_off, lcol)."""
    nn = sym["nnodes"]
    size = lib().oracle_factor_size(nn, _p(sym["sptr"], C.c_int), _p(sym["rptr"], C.c_int64))
    lcol = np.zeros(max(size, 1))
    off = np.zeros(nn + 1, np.int64)
    mask = None
    if node_mask is not None:
        mask = np.ascontiguousarray(node_mask, dtype=np.int8).tobytes()
    aval = np.ascontiguousarray(aval, np.float64)
    sc = None if scaling is None else np.ascontiguousarray(scaling, np.float64)
    st = lib().oracle_factor_posdef(sym["n"], nn, _p(sym["sptr"], C.c_int),
                                    _p(sym["sparent"], C.c_int), _p(sym["rptr"], C.c_int64),
                                    _p(sym["rlist"], C.c_int), _p(sym["nptr"], C.c_int64),
                                    _p(sym["nlist"], C.c_int64), _p(aval, C.c_double),
                                    _p(sc, C.c_double), nb, nthreads, mask, _p(off, C.c_int64),
                                    _p(lcol, C.c_double))
    return st, off, lcol


def node_factor(sym, off, lcol, node):
    m = int(sym["rptr"][node + 1] - sym["rptr"][node])
    n = int(sym["sptr"][node + 1] - sym["sptr"][node])
    return lcol[off[node]:off[node] + m * n].reshape((m, n), order="F")


def solve_posdef(sym, off, lcol, b):
    """b: (n,) or (n, nrhs) in pivot order; returns x in pivot order."""
    x = np.asfortranarray(np.array(b, dtype=np.float64, copy=True).reshape(sym["n"], -1))
    lib().oracle_solve_posdef(sym["n"], sym["nnodes"], _p(sym["sptr"], C.c_int),
                              _p(sym["rptr"], C.c_int64), _p(sym["rlist"], C.c_int),
                              _p(off, C.c_int64), _p(lcol, C.c_double), x.shape[1],
                              _p(x, C.c_double), sym["n"])
    return x


def front_potrf(a, ncol, nb=256):
    """Partial Cholesky of a dense front (copy returned, Fortran order)."""
    a = np.asfortranarray(np.array(a, dtype=np.float64, copy=True))
    m = a.shape[0]
    info = lib().oracle_front_potrf(m, ncol, _p(a, C.c_double), m, nb)
    return info, a


def bwd_err(A, x, b):
    x = np.asfortranarray(np.asarray(x, np.float64).reshape(A.n, -1))
    b = np.asfortranarray(np.asarray(b, np.float64).reshape(A.n, -1))
    return float(lib().oracle_bwd_err(A.n, _p(A.ptr, C.c_int64), _p(A.row, C.c_int),
                                      _p(A.val, C.c_double), x.shape[1], _p(x, C.c_double), A.n,
                                      _p(b, C.c_double), A.n))


STAT_KEYS = ("flag", "num_delay", "num_neg", "num_two", "num_zero", "maxfront",
             "not_first_pass", "not_second_pass")


class IndefFactor:
    """Oracle APP LDL^T factorization of a whole tree (see oracle.h)."""

    def __init__(self, sym, aval, nb=128, u=0.01, small=1e-20, action=True, failed_tpp=True,
                 scaling=None):
        self.sym = sym
        aval = np.ascontiguousarray(aval, np.float64)
        sc = None if scaling is None else np.ascontiguousarray(scaling, np.float64)
        self.h = lib().oracle_factor_indef(
            sym["n"], sym["nnodes"], _p(sym["sptr"], C.c_int), _p(sym["sparent"], C.c_int),
            _p(sym["rptr"], C.c_int64), _p(sym["rlist"], C.c_int), _p(sym["nptr"], C.c_int64),
            _p(sym["nlist"], C.c_int64), _p(aval, C.c_double),
            None if sc is None else _p(sc, C.c_double), nb, u, small, int(action),
            int(failed_tpp))
        out = np.zeros(8, np.int32)
        lib().oracle_indef_stats(self.h, _p(out, C.c_int))
        self.stats = dict(zip(STAT_KEYS, out.tolist()))

    def node(self, i):
        dims = np.zeros(4, np.int32)
        lib().oracle_indef_node(self.h, i, _p(dims, C.c_int), None, None, None)
        m, n, ne, nd = dims.tolist()
        rows = np.zeros(max(m, 1), np.int32)
        L = np.zeros(max(m * ne, 1))
        d = np.zeros(max(2 * ne, 1))
        lib().oracle_indef_node(self.h, i, _p(dims, C.c_int), _p(rows, C.c_int),
                                _p(L, C.c_double), _p(d, C.c_double))
        return dict(m=m, n=n, nelim=ne, ndelay_in=nd, rows=rows[:m],
                    L=L[:m * ne].reshape((m, ne), order="F"), dinv=d[:2 * ne])

    def solve(self, b, job=0):
        """b in pivot order, (n,) or (n, nrhs); returns x in pivot order."""
        x = np.asfortranarray(np.array(b, dtype=np.float64, copy=True).reshape(self.sym["n"], -1))
        lib().oracle_solve_indef(self.h, job, x.shape[1], _p(x, C.c_double), self.sym["n"])
        return x

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_indef_free(self.h)
            self.h = None


def front_ldlt(a, n, nb=128, u=0.01, small=1e-20, action=True, failed_tpp=True):
    """Dense APP + TPP factorization of a full symmetric front (copy).
    Returns (nelim, nelim_first, perm, dinv, factored front)."""
    a = np.asfortranarray(np.array(a, dtype=np.float64, copy=True))
    m = a.shape[0]
    perm = np.zeros(max(n, 1), np.int32)
    dinv = np.zeros(max(2 * n, 2))
    nf = C.c_int()
    ne = lib().oracle_front_ldlt(m, n, _p(a, C.c_double), nb, u, small, int(action),
                                 int(failed_tpp), _p(perm, C.c_int), _p(dinv, C.c_double),
                                 C.byref(nf))
    return ne, nf.value, perm[:n], dinv[:2 * n], a


LU_MODES = {"nopiv": 0, "pp": 1, "tpp": 2}


def lu_front(a, k, mode="nopiv", u=0.01):
    """Dense front LU of the square m x m front `a` (a copy) over its first
    k columns (oracle_lu.cpp). Returns (nelim, rperm, cperm, factored front):
    L unit lower and U upper in the first nelim rows/columns, the Schur
    complement in the trailing block, rperm/cperm the row/column orders."""
    dt = np.float32 if np.asarray(a).dtype == np.float32 else np.float64
    a = np.asfortranarray(np.array(a, dtype=dt, copy=True))
    m = a.shape[0]
    rp = np.zeros(max(m, 1), np.int32)
    cp = np.zeros(max(m, 1), np.int32)
    if dt == np.float32:
        ne = lib().oracle_lu_front_f32(LU_MODES[mode], u, m, k, _p(a, C.c_float), m, _p(rp, C.c_int),
                                       _p(cp, C.c_int))
    else:
        ne = lib().oracle_lu_front(LU_MODES[mode], u, m, k, _p(a, C.c_double), m, _p(rp, C.c_int),
                                   _p(cp, C.c_int))
    return ne, rp[:m], cp[:m], a


def dinv_blocks(dinv, nelim):
    """Yield (start, width, 2x2-or-1x1 D^{-1} block) from the D^{-1} array."""
    i = 0
    while i < nelim:
        if i + 1 < nelim and np.isinf(dinv[2 * i + 2]):
            yield i, 2, np.array([[dinv[2 * i], dinv[2 * i + 1]], [dinv[2 * i + 1], dinv[2 * i + 3]]])
            i += 2
        else:
            yield i, 1, np.array([[dinv[2 * i]]])
            i += 1
