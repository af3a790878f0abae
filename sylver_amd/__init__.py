"""sylver_amd: MI355X-native multifrontal numeric factorization behind
SyLVER's C interface.

This module loads the in-tree native library ``libsylver_amd.so`` through
ctypes and mirrors the reference's API names and argument meaning
(include/sylver/sylver.h:73-123 in the reference). The library itself is C++
and HIP; Python is only a thin caller, used by the tests and the benchmark.

There is no CPU fallback for the numeric phase: without a HIP device
``spldlt_factorize`` returns flag SYLVER_ERROR_CUDA_UNKNOWN (-51), and if the
native library has not been built, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsylver_amd.so")

# flag codes (src/sylver_datatypes_mod.F90:14-45 in the reference)
SUCCESS = 0
ERROR_CALL_SEQUENCE = -1
ERROR_A_N_OOR = -2
ERROR_A_PTR = -3
ERROR_A_ALL_OOR = -4
ERROR_SINGULAR = -5
ERROR_NOT_POS_DEF = -6
ERROR_PTR_ROW = -7
ERROR_ORDER = -8
ERROR_VAL = -9
ERROR_X_SIZE = -10
ERROR_JOB_OOR = -11
ERROR_ALLOCATION = -50
ERROR_CUDA_UNKNOWN = -51
ERROR_UNIMPLEMENTED = -98
ERROR_UNKNOWN = -99
WARNING_IDX_OOR = 1
WARNING_DUP_IDX = 2
WARNING_DUP_AND_OOR = 3
WARNING_MISSING_DIAGONAL = 4
WARNING_MISS_DIAG_OORDUP = 5
WARNING_ANAL_SINGULAR = 6
WARNING_FACT_SINGULAR = 7

PREC_FP64 = 0
PREC_MIXED = 1
PREC_FP32 = 2
REFINE_NONE = 0
REFINE_IR = 1
REFINE_GMRES = 2
REFINE_AUTO = 3


class Inform(C.Structure):
    """sylver_inform_t (reference sylver.h:19-37)."""
    _fields_ = [
        ("flag", C.c_int), ("matrix_dup", C.c_int), ("matrix_missing_diag", C.c_int),
        ("matrix_outrange", C.c_int), ("matrix_rank", C.c_int), ("maxdepth", C.c_int),
        ("maxfront", C.c_int), ("num_delay", C.c_int), ("num_factor", C.c_long),
        ("num_flops", C.c_long), ("num_neg", C.c_int), ("num_sup", C.c_int),
        ("num_two", C.c_int), ("stat", C.c_int), ("cuda_error", C.c_int),
        ("cublas_error", C.c_int), ("unused", C.c_ubyte * 80),
    ]

    def tail(self):
        t = InformTail()
        C.memmove(C.byref(t), C.addressof(self) + Inform.unused.offset, C.sizeof(t))
        return t


class InformTail(C.Structure):
    """sylver_amd_inform_tail_t, written into sylver_inform_t.unused."""
    _fields_ = [
        ("factor_ms", C.c_double), ("solve_ms", C.c_double), ("gflops", C.c_double),
        ("bwd_err", C.c_double), ("refine_iters", C.c_int), ("refine_mode", C.c_int),
        ("precision", C.c_int), ("nlevels", C.c_int), ("upload_ms", C.c_double),
        ("pad", C.c_char * 16),
    ]


class Options(C.Structure):
    """sylver_options_t (reference sylver.h:39-71)."""
    _fields_ = [
        ("array_base", C.c_int), ("print_level", C.c_int), ("unit_diagnostics", C.c_int),
        ("unit_error", C.c_int), ("unit_warning", C.c_int), ("ordering", C.c_int),
        ("nemin", C.c_int), ("prune_tree", C.c_bool), ("min_gpu_work", C.c_long),
        ("scaling", C.c_int), ("pivot_method", C.c_int), ("small", C.c_double),
        ("u", C.c_double), ("small_subtree_threshold", C.c_long), ("nb", C.c_int),
        ("cpu_topology", C.c_int), ("action", C.c_bool), ("use_gpu", C.c_bool),
        ("gpu_perf_coeff", C.c_double), ("failed_pivot_method", C.c_int),
        ("scheduler", C.c_int),
    ]


class OptionsC(C.Structure):
    """sylver_options_c (include/sylver_amd.h; reference src/sylver_ciface.hxx:39-55)."""
    _fields_ = [("print_level", C.c_int), ("action", C.c_bool), ("small", C.c_double),
                ("u", C.c_double), ("multiplier", C.c_double),
                ("small_subtree_threshold", C.c_long), ("nb", C.c_int), ("pivot_method", C.c_int),
                ("failed_pivot_method", C.c_int), ("cpu_topology", C.c_int)]


class InformC(C.Structure):
    """sylver_inform_c (include/sylver_amd.h; reference src/sylver_ciface.hxx:73-86)."""
    _fields_ = [(k, C.c_int) for k in ("flag", "num_delay", "num_neg", "num_two", "num_zero",
                                        "maxfront", "not_first_pass", "not_second_pass")]


class ExtOptions(C.Structure):
    """sylver_amd_options_t."""
    _fields_ = [
        ("precision", C.c_int), ("bf16_min_ncol", C.c_int), ("refine", C.c_int),
        ("refine_tol", C.c_double), ("refine_max_iter", C.c_int), ("use_graph", C.c_int),
        ("tile_nb", C.c_int), ("small_front_max", C.c_int), ("record_timing", C.c_int),
        ("indef_panel_min", C.c_int), ("small_assembly", C.c_int), ("strip_trsm", C.c_int),
        ("level_assembly", C.c_int), ("reserved", C.c_int * 3),
    ]


class Stats(C.Structure):
    """sylver_amd_stats_t."""
    _fields_ = [
        ("factor_ms", C.c_double), ("upload_ms", C.c_double), ("solve_ms", C.c_double),
        ("flops", C.c_double), ("nlevels", C.c_int), ("nlaunches", C.c_int),
        ("precision", C.c_int), ("refine_iters", C.c_int), ("bwd_err", C.c_double),
        ("kind_ms", C.c_double * 16), ("kind_flops", C.c_double * 16),
        ("kind_bytes", C.c_double * 16), ("kind_launches", C.c_int * 16),
    ]

KIND_NAMES = ["zero", "scatter_a", "extend_add", "small_front", "diag_potrf", "panel_trsm",
              "schur_update", "diag_inverse", "schur_update_wide"]


_lib = None

# Every symbol declared in include/sylver/sylver.h and include/sylver_amd.h.
EXPORTED = [
    "sylver_init", "sylver_finalize", "sylver_default_options", "spldlt_analyse",
    "spldlt_factorize", "spldlt_solve", "spldlt_free_akeep", "spldlt_free_fkeep",
    "spldlt_create_symbolic_tree", "spldlt_destroy_symbolic_tree",
    "spldlt_create_numeric_tree_posdef_dbl", "spldlt_create_numeric_tree_dbl",
    "spldlt_destroy_numeric_tree_posdef_dbl", "spldlt_destroy_numeric_tree_dbl",
    "spldlt_tree_solve_fwd_posdef_dbl", "spldlt_tree_solve_bwd_posdef_dbl",
    "spldlt_tree_solve_fwd_dbl", "spldlt_tree_solve_bwd_dbl", "spldlt_tree_solve_diag_dbl",
    "spldlt_tree_solve_diag_bwd_dbl", "sylver_amd_default_options", "sylver_amd_set_options",
    "sylver_amd_get_options", "sylver_amd_fkeep_stats", "sylver_amd_akeep_sizes",
    "sylver_amd_akeep_symbolic", "sylver_amd_fkeep_device_symbolic",
    "sylver_amd_fkeep_node_factor", "sylver_amd_fkeep_node_assembled",
    "sylver_amd_fkeep_indef_node", "sylver_amd_fkeep_indef_stats",
    "sylver_amd_batched_front_potrf", "sylver_amd_batched_front_fill", "sylver_amd_device_count", "sylver_amd_dev_alloc",
    "sylver_amd_dev_free", "sylver_amd_memcpy_h2d", "sylver_amd_memcpy_d2h",
    "sylver_amd_last_batched_ms", "sylver_amd_set_front_variants", "sylver_amd_time_fronts",
    "sylver_amd_batched_front_potrf_v", "sylver_amd_batched_front_potrf_noalloc", "sylver_amd_batched_front_lu",
    "sylver_amd_mlp_init", "sylver_amd_mlp_grad", "sylver_amd_mlp_train", "sylver_amd_mlp_predict",
    "sylver_amd_hungarian_scale_sym", "sylver_amd_auction_scale_sym", "sylver_amd_match_order",
    "sylver_amd_plan_check", "sylver_amd_rb_read", "sylver_amd_rb_free",
]


def lib():
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"native library missing: {LIB_PATH}; build it with `python -m sylver_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, pvp = C.c_void_p, C.POINTER(C.c_void_p)
    ip, lp, dp = C.POINTER(C.c_int), C.POINTER(C.c_long), C.POINTER(C.c_double)
    L.sylver_init.argtypes = [C.c_int, C.c_int]
    L.sylver_finalize.argtypes = []
    L.sylver_default_options.argtypes = [C.POINTER(Options)]
    L.spldlt_analyse.argtypes = [C.c_int, ip, lp, ip, dp, pvp, C.c_bool, C.POINTER(Options),
                                 C.POINTER(Inform)]
    L.spldlt_factorize.argtypes = [C.c_bool, lp, ip, dp, dp, vp, pvp, C.POINTER(Options),
                                   C.POINTER(Inform)]
    L.spldlt_solve.argtypes = [C.c_int, C.c_int, dp, C.c_int, vp, vp, C.POINTER(Options),
                               C.POINTER(Inform)]
    L.spldlt_free_akeep.argtypes = [pvp]
    L.spldlt_free_fkeep.argtypes = [pvp]
    L.sylver_amd_default_options.argtypes = [C.POINTER(ExtOptions)]
    L.sylver_amd_set_options.argtypes = [C.POINTER(ExtOptions)]
    L.sylver_amd_get_options.argtypes = [C.POINTER(ExtOptions)]
    L.sylver_amd_fkeep_stats.argtypes = [vp, C.POINTER(Stats)]
    L.sylver_amd_akeep_sizes.argtypes = [vp, ip, ip, lp, lp]
    L.sylver_amd_akeep_symbolic.argtypes = [vp, ip, ip, ip, ip, lp, ip, lp, lp]
    L.sylver_amd_fkeep_device_symbolic.argtypes = [vp, ip, ip, lp, ip, lp, lp, ip, lp]
    L.sylver_amd_fkeep_node_factor.argtypes = [vp, C.c_int, dp]
    L.sylver_amd_fkeep_indef_node.argtypes = [vp, C.c_int, ip, ip, dp, dp]
    L.sylver_amd_fkeep_indef_stats.argtypes = [vp, ip]
    L.sylver_amd_batched_front_potrf.argtypes = [C.c_int, ip, ip, lp, vp, C.c_int, ip, vp]
    L.sylver_amd_batched_front_potrf_noalloc.argtypes = [C.c_int, ip, ip, lp, vp, C.c_int, ip, vp]
    L.sylver_amd_batched_front_lu.argtypes = [C.c_int, ip, ip, lp, vp, C.c_int, C.c_int, C.c_double, ip, ip, ip, vp]
    L.sylver_amd_batched_front_fill.argtypes = [C.c_int, ip, lp, vp, C.c_int, C.c_ulonglong]
    L.sylver_amd_device_count.argtypes = []
    L.sylver_amd_dev_alloc.argtypes = [pvp, C.c_size_t]
    L.sylver_amd_dev_free.argtypes = [vp]
    L.sylver_amd_memcpy_h2d.argtypes = [vp, vp, C.c_size_t]
    L.sylver_amd_memcpy_d2h.argtypes = [vp, vp, C.c_size_t]
    L.sylver_amd_last_batched_ms.argtypes = []
    L.sylver_amd_last_batched_ms.restype = C.c_double
    fp = C.POINTER(C.c_float)
    L.sylver_amd_set_front_variants.argtypes = [vp, C.POINTER(C.c_byte), C.c_int]
    L.sylver_amd_batched_front_potrf_v.argtypes = [C.c_int, ip, ip, lp, vp, C.c_int, C.POINTER(C.c_byte), ip]
    L.sylver_amd_time_fronts.argtypes = [C.c_int, ip, ip, C.POINTER(C.c_byte), C.c_int, C.c_int, dp]
    L.sylver_amd_mlp_init.argtypes = [fp, C.c_ulonglong]
    L.sylver_amd_mlp_grad.argtypes = [fp, fp, fp, C.c_int, fp, fp]
    L.sylver_amd_mlp_train.argtypes = [fp, fp, fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_ulonglong, fp]
    L.sylver_amd_mlp_predict.argtypes = [fp, fp, C.c_int, fp]
    L.sylver_amd_hungarian_scale_sym.argtypes = [C.c_int, lp, ip, dp, C.c_int, dp, ip]
    L.sylver_amd_auction_scale_sym.argtypes = [C.c_int, lp, ip, dp, dp, ip]
    L.sylver_amd_match_order.argtypes = [C.c_int, lp, ip, dp, ip, dp]
    L.sylver_amd_plan_check.argtypes = [vp, C.c_int, C.c_int, C.c_int, lp, C.c_char_p, C.c_int]
    L.sylver_amd_rb_read.argtypes = [C.c_char_p, C.c_int, C.c_ulonglong, ip, ip, C.POINTER(lp), C.POINTER(ip),
                                     C.POINTER(dp), C.c_char_p]
    L.sylver_amd_rb_free.argtypes = [vp]
    L.sylver_amd_rb_free.restype = None
    _lib = L
    return L


def _p(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def default_options():
    o = Options()
    lib().sylver_default_options(C.byref(o))
    return o


def default_ext_options():
    o = ExtOptions()
    lib().sylver_amd_default_options(C.byref(o))
    return o


def set_ext_options(**kw):
    """Update the process-wide extension options (precision, refine, ...)."""
    o = ExtOptions()
    lib().sylver_amd_get_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    lib().sylver_amd_set_options(C.byref(o))
    return o


def device_count():
    return int(lib().sylver_amd_device_count())


def sylver_init(ncpu=1, ngpu=1):
    lib().sylver_init(ncpu, ngpu)


def sylver_finalize():
    lib().sylver_finalize()


def spldlt_analyse(n, ptr, row, val=None, order=None, check=True, options=None):
    """Reference: spldlt_analyse (sylver.h:79-83). Returns (akeep, inform, order);
    ``order`` is the final pivot order written back by the library."""
    opts = options or default_options()
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    row = np.ascontiguousarray(row, dtype=np.int32)
    ordv = (np.ascontiguousarray(order, dtype=np.int32).copy() if order is not None
            else np.zeros(max(n, 1), dtype=np.int32))
    akeep = C.c_void_p()
    inform = Inform()
    valp = None if val is None else np.ascontiguousarray(val, dtype=np.float64)
    lib().spldlt_analyse(n, _p(ordv, C.c_int), _p(ptr, C.c_long), _p(row, C.c_int),
                         _p(valp, C.c_double), C.byref(akeep), bool(check), C.byref(opts),
                         C.byref(inform))
    return akeep, inform, ordv[:n]


def spldlt_factorize(posdef, val, akeep, fkeep=None, ptr=None, row=None, scale=None, options=None):
    """Reference: spldlt_factorize (sylver.h:97-100). Returns (fkeep, inform)."""
    opts = options or default_options()
    fk = fkeep if fkeep is not None else C.c_void_p()
    inform = Inform()
    val = np.ascontiguousarray(val, dtype=np.float64)
    pp = None if ptr is None else np.ascontiguousarray(ptr, dtype=np.int64)
    rr = None if row is None else np.ascontiguousarray(row, dtype=np.int32)
    lib().spldlt_factorize(bool(posdef), _p(pp, C.c_long), _p(rr, C.c_int), _p(val, C.c_double),
                           _p(scale, C.c_double), akeep, C.byref(fk), C.byref(opts),
                           C.byref(inform))
    return fk, inform


def spldlt_solve(x, akeep, fkeep, job=0, options=None):
    """Reference: spldlt_solve (sylver.h:102-104). x: (n,) or (n, nrhs) array,
    solved in place (Fortran order). Returns inform."""
    opts = options or default_options()
    assert x.dtype == np.float64 and (x.ndim == 1 or x.flags.f_contiguous)
    n = x.shape[0]
    nrhs = 1 if x.ndim == 1 else x.shape[1]
    inform = Inform()
    lib().spldlt_solve(job, nrhs, _p(x, C.c_double), n, akeep, fkeep, C.byref(opts),
                       C.byref(inform))
    return inform


def spldlt_free_akeep(akeep):
    lib().spldlt_free_akeep(C.byref(akeep))


def spldlt_free_fkeep(fkeep):
    lib().spldlt_free_fkeep(C.byref(fkeep))


def akeep_symbolic(akeep):
    """Host copies of the analyse outputs (1-based, reference conventions)."""
    n, nn = C.c_int(), C.c_int()
    nr, nl = C.c_long(), C.c_long()
    lib().sylver_amd_akeep_sizes(akeep, C.byref(n), C.byref(nn), C.byref(nr), C.byref(nl))
    n, nn, nr, nl = n.value, nn.value, nr.value, nl.value
    out = dict(
        order=np.zeros(n, np.int32), invp=np.zeros(n, np.int32),
        sptr=np.zeros(nn + 1, np.int32), sparent=np.zeros(nn, np.int32),
        rptr=np.zeros(nn + 1, np.int64), rlist=np.zeros(nr, np.int32),
        nptr=np.zeros(nn + 1, np.int64), nlist=np.zeros(2 * nl, np.int64))
    lib().sylver_amd_akeep_symbolic(
        akeep, _p(out["order"], C.c_int), _p(out["invp"], C.c_int), _p(out["sptr"], C.c_int),
        _p(out["sparent"], C.c_int), _p(out["rptr"], C.c_long), _p(out["rlist"], C.c_int),
        _p(out["nptr"], C.c_long), _p(out["nlist"], C.c_long))
    out["n"] = n
    out["nnodes"] = nn
    return out


def fkeep_device_symbolic(fkeep, sym):
    """The device-resident copies read back from HBM."""
    nn = sym["nnodes"]
    q = (np.diff(sym["rptr"]) - np.diff(sym["sptr"])).astype(np.int64)
    out = dict(
        sptr=np.zeros(nn + 1, np.int32), sparent=np.zeros(nn, np.int32),
        rptr=np.zeros(nn + 1, np.int64), rlist=np.zeros(len(sym["rlist"]), np.int32),
        nptr=np.zeros(nn + 1, np.int64), nlist=np.zeros(len(sym["nlist"]), np.int64),
        cmap=np.zeros(max(int(q.sum()), 1), np.int32), cmap_off=np.zeros(nn + 1, np.int64))
    flag = lib().sylver_amd_fkeep_device_symbolic(
        fkeep, _p(out["sptr"], C.c_int), _p(out["sparent"], C.c_int), _p(out["rptr"], C.c_long),
        _p(out["rlist"], C.c_int), _p(out["nptr"], C.c_long), _p(out["nlist"], C.c_long),
        _p(out["cmap"], C.c_int), _p(out["cmap_off"], C.c_long))
    if flag != 0:
        raise RuntimeError(f"device symbolic read-back failed: {flag}")
    return out


def fkeep_node_factor(fkeep, node, m, n):
    out = np.zeros((m, n), order="F")
    flag = lib().sylver_amd_fkeep_node_factor(fkeep, node, _p(out, C.c_double))
    if flag != 0:
        raise RuntimeError(f"node factor read-back failed: {flag}")
    return out


INDEF_STAT_KEYS = ("num_delay", "num_neg", "num_two", "num_zero", "maxfront",
                   "not_first_pass", "not_second_pass")


def fkeep_indef_stats(fkeep):
    out = np.zeros(8, np.int32)
    flag = lib().sylver_amd_fkeep_indef_stats(fkeep, _p(out, C.c_int))
    if flag != 0:
        raise RuntimeError(f"indefinite statistics read-back failed: {flag}")
    return dict(zip(INDEF_STAT_KEYS, out[:7].tolist()))


def fkeep_indef_node(fkeep, node):
    """Indefinite factors of one node: dict(m, n, nelim, ndelay_in, rows, L, dinv)."""
    dims = np.zeros(4, np.int32)
    flag = lib().sylver_amd_fkeep_indef_node(fkeep, node, _p(dims, C.c_int), None, None, None)
    if flag != 0:
        raise RuntimeError(f"indefinite node read-back failed: {flag}")
    m, n, ne, nd = dims.tolist()
    rows = np.zeros(max(m, 1), np.int32)
    L = np.zeros((max(m, 1), max(ne, 1)), order="F")
    d = np.zeros(max(2 * ne, 1))
    lib().sylver_amd_fkeep_indef_node(fkeep, node, _p(dims, C.c_int), _p(rows, C.c_int),
                                      _p(L, C.c_double), _p(d, C.c_double))
    return dict(m=m, n=n, nelim=ne, ndelay_in=nd, rows=rows[:m], L=L[:m, :ne], dinv=d[:2 * ne])


def fkeep_stats(fkeep):
    s = Stats()
    lib().sylver_amd_fkeep_stats(fkeep, C.byref(s))
    return s


class SpLDLT:
    """Analyse once, factorize (possibly many times) and solve: the call
    sequence of the reference's drivers (drivers/spldlt_test.F90:123-150)."""

    def __init__(self, A, order=None, options=None, check=True):
        self.A = A
        self.options = options or default_options()
        if order is not None:
            self.options.ordering = 0
        self.akeep, inf, self.order = spldlt_analyse(A.n, A.ptr, A.row, A.val, order=order,
                                                     check=check, options=self.options)
        self.analyse_inform = inf
        if inf.flag < 0:
            raise RuntimeError(f"spldlt_analyse failed: flag {inf.flag}")
        self.fkeep = None

    def factorize(self, posdef=True, val=None):
        val = self.A.val if val is None else val
        self.fkeep, inf = spldlt_factorize(posdef, val, self.akeep, self.fkeep,
                                           options=self.options)
        self.factor_inform = inf
        return inf

    def solve(self, b, job=0):
        x = np.asfortranarray(np.array(b, dtype=np.float64, copy=True))
        inf = spldlt_solve(x, self.akeep, self.fkeep, job=job, options=self.options)
        self.solve_inform = inf
        return x, inf

    def free(self):
        if self.fkeep is not None:
            spldlt_free_fkeep(self.fkeep)
            self.fkeep = None
        if self.akeep is not None:
            spldlt_free_akeep(self.akeep)
            self.akeep = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceArray:
    """A device buffer owned through the library's HIP runtime."""

    def __init__(self, host: np.ndarray):
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        if lib().sylver_amd_dev_alloc(C.byref(self.ptr), self.nbytes) != 0:
            raise MemoryError("device allocation failed")
        h = np.ascontiguousarray(host)
        if lib().sylver_amd_memcpy_h2d(self.ptr, h.ctypes.data_as(C.c_void_p), self.nbytes) != 0:
            raise RuntimeError("host->device copy failed")

    @classmethod
    def empty(cls, count, dtype):
        """An uninitialized device buffer of `count` elements."""
        self = cls.__new__(cls)
        self.shape, self.dtype = (int(count),), np.dtype(dtype)
        self.nbytes = int(count) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        if lib().sylver_amd_dev_alloc(C.byref(self.ptr), max(self.nbytes, 1)) != 0:
            raise MemoryError("device allocation failed")
        return self

    def read(self, off, count):
        """Copy `count` elements starting at element `off` to the host."""
        out = np.empty(int(count), self.dtype)
        src = C.c_void_p(self.ptr.value + int(off) * self.dtype.itemsize)
        if lib().sylver_amd_memcpy_d2h(out.ctypes.data_as(C.c_void_p), src, out.nbytes) != 0:
            raise RuntimeError("device->host copy failed")
        return out

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if lib().sylver_amd_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes) != 0:
            raise RuntimeError("device->host copy failed")
        return out

    def free(self):
        if self.ptr:
            lib().sylver_amd_dev_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def batched_front_fill(fronts_flat: "DeviceArray", nrow, offsets, dtype, seed):
    """Synthetic SPD fronts (see sylver_amd_batched_front_fill)."""
    nrow = np.ascontiguousarray(nrow, np.int32)
    offsets = np.ascontiguousarray(offsets, np.int64)
    return lib().sylver_amd_batched_front_fill(
        len(nrow), _p(nrow, C.c_int), _p(offsets, C.c_long), fronts_flat.ptr,
        0 if np.dtype(dtype) == np.float64 else 1, int(seed))


def batched_front_potrf(fronts_flat: "DeviceArray", nrow, ncol, offsets, dtype):
    """Factor many dense fronts in one launch (device data); returns
    (rc, info array, device ms)."""
    nrow = np.ascontiguousarray(nrow, np.int32)
    ncol = np.ascontiguousarray(ncol, np.int32)
    offsets = np.ascontiguousarray(offsets, np.int64)
    info = DeviceArray(np.zeros(len(nrow), np.int32))
    rc = lib().sylver_amd_batched_front_potrf(
        len(nrow), _p(nrow, C.c_int), _p(ncol, C.c_int), _p(offsets, C.c_long), fronts_flat.ptr,
        0 if np.dtype(dtype) == np.float64 else 1, C.cast(info.ptr, C.POINTER(C.c_int)), None)
    out = info.numpy()
    info.free()
    return rc, out, float(lib().sylver_amd_last_batched_ms())


def batched_front_potrf_noalloc(fronts_flat: "DeviceArray", nrow, ncol, offsets, dtype, info: "DeviceArray"):
    """sylver_amd_batched_front_potrf without device allocations (sizes
    staged in the library's fixed per-device array); info is the caller's
    device array of at least len(nrow) int32. Returns the flag."""
    nrow = np.ascontiguousarray(nrow, np.int32)
    ncol = np.ascontiguousarray(ncol, np.int32)
    offsets = np.ascontiguousarray(offsets, np.int64)
    return lib().sylver_amd_batched_front_potrf_noalloc(
        len(nrow), _p(nrow, C.c_int), _p(ncol, C.c_int), _p(offsets, C.c_long), fronts_flat.ptr,
        0 if np.dtype(dtype) == np.float64 else 1, C.cast(info.ptr, C.POINTER(C.c_int)), None)


LU_MODES = {"nopiv": 0, "pp": 1, "tpp": 2}


def batched_front_lu(fronts_flat: "DeviceArray", m, k, offsets, dtype, mode="nopiv", u=0.01):
    """Dense front LU of many square fronts in one launch (lu.hip). Returns
    (flag, nelim, list of (rperm, cperm) per front)."""
    m = np.ascontiguousarray(m, np.int32)
    k = np.ascontiguousarray(k, np.int32)
    offsets = np.ascontiguousarray(offsets, np.int64)
    tot = int(m.sum())
    rp = np.zeros(max(tot, 1), np.int32)
    cp = np.zeros(max(tot, 1), np.int32)
    ne = np.zeros(max(len(m), 1), np.int32)
    rc = lib().sylver_amd_batched_front_lu(
        len(m), _p(m, C.c_int), _p(k, C.c_int), _p(offsets, C.c_long), fronts_flat.ptr,
        0 if np.dtype(dtype) == np.float64 else 1, LU_MODES[mode], float(u), _p(rp, C.c_int), _p(cp, C.c_int),
        _p(ne, C.c_int), None)
    perms, o = [], 0
    for mm in m:
        perms.append((rp[o:o + mm].copy(), cp[o:o + mm].copy()))
        o += int(mm)
    return rc, ne[:len(m)].copy(), perms


def batched_front_potrf_v(fronts_flat: "DeviceArray", nrow, ncol, offsets, dtype, variant=None):
    """The same factorization through the launch plan with a per-front
    variant (1 one workgroup, 0 tiled path, -1 / None default rule)."""
    nrow = np.ascontiguousarray(nrow, np.int32)
    ncol = np.ascontiguousarray(ncol, np.int32)
    offsets = np.ascontiguousarray(offsets, np.int64)
    vp = None
    if variant is not None:
        v = np.ascontiguousarray(variant, np.int8)
        vp = _p(v, C.c_byte)
    info = DeviceArray(np.zeros(len(nrow), np.int32))
    rc = lib().sylver_amd_batched_front_potrf_v(
        len(nrow), _p(nrow, C.c_int), _p(ncol, C.c_int), _p(offsets, C.c_long), fronts_flat.ptr,
        0 if np.dtype(dtype) == np.float64 else 1, vp, C.cast(info.ptr, C.POINTER(C.c_int)))
    out = info.numpy()
    info.free()
    return rc, out, float(lib().sylver_amd_last_batched_ms())


# ---- autotuner and learned cost model (see sylver_amd/costmodel.py) ----
MLP_PARAMS = 8 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2


def set_front_variants(akeep, variant):
    """Per-front kernel variant of later posdef factorizations (1 one-workgroup
    kernel, 0 tiled path, -1 default rule); None clears."""
    if variant is None:
        rc = lib().sylver_amd_set_front_variants(akeep, None, 0)
    else:
        v = np.ascontiguousarray(variant, np.int8)
        rc = lib().sylver_amd_set_front_variants(akeep, _p(v, C.c_byte), len(v))
    if rc != 0:
        raise RuntimeError(f"sylver_amd_set_front_variants failed: {rc}")


def time_fronts(m, n, variant=None, precision=PREC_FP64, reps=3):
    """Device ms of factoring the synthetic fronts (m[i], n[i]) together."""
    m = np.ascontiguousarray(m, np.int32)
    n = np.ascontiguousarray(n, np.int32)
    vp = None
    if variant is not None:
        v = np.ascontiguousarray(variant, np.int8)
        vp = _p(v, C.c_byte)
    ms = C.c_double(0.0)
    rc = lib().sylver_amd_time_fronts(len(m), _p(m, C.c_int), _p(n, C.c_int), vp, int(precision), int(reps),
                                      C.byref(ms))
    if rc != 0:
        raise RuntimeError(f"sylver_amd_time_fronts failed: {rc}")
    return ms.value


def mlp_init(seed):
    p = np.zeros(MLP_PARAMS, np.float32)
    lib().sylver_amd_mlp_init(_p(p, C.c_float), int(seed))
    return p


def mlp_grad(params, x, y):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    g = np.zeros(MLP_PARAMS, np.float32)
    loss = np.zeros(1, np.float32)
    rc = lib().sylver_amd_mlp_grad(_p(params, C.c_float), _p(x, C.c_float), _p(y, C.c_float), len(x),
                                   _p(g, C.c_float), _p(loss, C.c_float))
    if rc != 0:
        raise RuntimeError(f"sylver_amd_mlp_grad failed: {rc}")
    return g, float(loss[0])


def mlp_train(params, x, y, epochs, batch, lr, seed):
    """Returns (trained params, per-epoch loss)."""
    p = np.array(params, np.float32, copy=True)
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    hist = np.zeros(epochs, np.float32)
    rc = lib().sylver_amd_mlp_train(_p(p, C.c_float), _p(x, C.c_float), _p(y, C.c_float), len(x), int(epochs),
                                    int(batch), float(lr), int(seed), _p(hist, C.c_float))
    if rc != 0:
        raise RuntimeError(f"sylver_amd_mlp_train failed: {rc}")
    return p, hist


def mlp_predict(params, x):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros((len(x), 2), np.float32)
    rc = lib().sylver_amd_mlp_predict(_p(params, C.c_float), _p(x, C.c_float), len(x), _p(out, C.c_float))
    if rc != 0:
        raise RuntimeError(f"sylver_amd_mlp_predict failed: {rc}")
    return out



# ---- scalings and matching-based ordering (host code; see include/sylver_amd.h) ----
def hungarian_scale_sym(A, scale_if_singular=True):
    """Returns (flag, scaling, match) in the caller's variable order."""
    s = np.zeros(A.n)
    m = np.zeros(A.n, np.int32)
    flag = lib().sylver_amd_hungarian_scale_sym(
        A.n, _p(np.ascontiguousarray(A.ptr, np.int64), C.c_long), _p(np.ascontiguousarray(A.row, np.int32), C.c_int),
        _p(np.ascontiguousarray(A.val, np.float64), C.c_double), int(scale_if_singular), _p(s, C.c_double),
        _p(m, C.c_int))
    return flag, s, m


def auction_scale_sym(A):
    s = np.zeros(A.n)
    m = np.zeros(A.n, np.int32)
    flag = lib().sylver_amd_auction_scale_sym(
        A.n, _p(np.ascontiguousarray(A.ptr, np.int64), C.c_long), _p(np.ascontiguousarray(A.row, np.int32), C.c_int),
        _p(np.ascontiguousarray(A.val, np.float64), C.c_double), _p(s, C.c_double), _p(m, C.c_int))
    return flag, s, m


def match_order(A):
    """Returns (flag, order (1-based positions), scaling)."""
    order = np.zeros(A.n, np.int32)
    s = np.zeros(A.n)
    flag = lib().sylver_amd_match_order(
        A.n, _p(np.ascontiguousarray(A.ptr, np.int64), C.c_long), _p(np.ascontiguousarray(A.row, np.int32), C.c_int),
        _p(np.ascontiguousarray(A.val, np.float64), C.c_double), _p(order, C.c_int), _p(s, C.c_double))
    return flag, order, s


def plan_check(akeep, small_front_max=0, bf16_min_ncol=0, num_cus=256):
    """Host-side index check of the device launch plan (no device needed).
    Returns (stats dict, first violation text)."""
    out = np.zeros(10, np.int64)
    why = C.create_string_buffer(256)
    rc = lib().sylver_amd_plan_check(akeep, int(small_front_max), int(bf16_min_ncol), int(num_cus),
                                     _p(out, C.c_long), why, 256)
    if rc != 0:
        raise RuntimeError(f"sylver_amd_plan_check failed: {rc} {why.value.decode()}")
    keys = ("violations", "factor_launches", "fwd_launches", "bwd_launches", "arena_elems",
            "exchange_slots", "levels", "cb_slots", "strip_tasks", "diag_tasks")
    return dict(zip(keys, out.tolist())), why.value.decode()
