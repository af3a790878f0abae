"""Oracle for the scalings (TEST INFRASTRUCTURE ONLY: used by tests/ to check
the library; never called by it).

equilib_scale_sym restates the norm equilibration the reference reaches
through SPRAL's equilib_scale_sym (called at
src/spldlt_factorize_mod.F90:808-835; SPRAL itself is not in the reference's
tree, so the loop is stated from Ruiz's method, RAL-TR-2001-034, with
SPRAL's defaults of 10 sweeps and tolerance 1e-8): s = 1; per sweep, for
every stored lower entry a_rc the scaled value |s_r a_rc s_c| (formed as
(s_r |a|) s_c) updates the running maxima of rows r and c; then
s_i /= sqrt(max_i) where max_i > 0; stop after a sweep whose maxima are all
within tol of 1.

The matching-based scalings are checked against scipy's
linear_sum_assignment (the assignment problem they solve) in the tests.
"""
from __future__ import annotations

import numpy as np


def equilib_scale_sym(A, max_iter=10, tol=1e-8):
    """A: lower CSC, 1-based (A.n, A.ptr, A.row, A.val). Returns
    (scaling in the caller's order, sweeps run)."""
    n = A.n
    col = np.repeat(np.arange(n), np.diff(A.ptr))
    row = A.row.astype(np.int64) - 1
    a = np.abs(A.val)
    s = np.ones(n)
    it = 0
    for it in range(1, max_iter + 1):
        v = (s[row] * a) * s[col]
        mx = np.zeros(n)
        np.maximum.at(mx, row, v)
        np.maximum.at(mx, col, v)
        pos = mx > 0
        s = np.where(pos, s / np.sqrt(np.where(pos, mx, 1.0)), s)
        if np.max(np.abs(1.0 - mx)) < tol:
            break
    return s, it


def scaled_abs_max(A, s):
    """Largest |s_i a_ij s_j| over the stored entries, and the per-row maxima
    of the full symmetric scaled matrix."""
    n = A.n
    col = np.repeat(np.arange(n), np.diff(A.ptr))
    row = A.row.astype(np.int64) - 1
    v = np.abs(A.val) * s[row] * s[col]
    mx = np.zeros(n)
    np.maximum.at(mx, row, v)
    np.maximum.at(mx, col, v)
    return float(v.max()) if len(v) else 0.0, mx
