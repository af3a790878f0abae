"""Writes tests/golden/small_cases.json: the reference's small test matrices
with expected results fixed by closed forms (numpy), used to pin the oracle
and the device path.

Matrices (reference sources, read as text):
- simple_mat: tests/sylver_test_mod.F90:120-155, 4x4 SPD, right-hand side
  A*1 so that x = 1;
- simple_mat_indef: tests/sylver_test_mod.F90:320-355, 4x4 indefinite
  (inertia from the eigenvalues);
- c_example: examples/C/spldlt_simple_example_c.c:30-40, 3x3 tridiagonal
  [2 -1; -1 2 -1; -1 2], b = 1, x = [1.5, 2, 1.5];
- chol_dd_51: the 51x51 diagonally dominant block of
  src/kernels/test/chol_kernels.cxx (entries uniform, diagonal |a_jj| +
  max(m, n)), here from numpy seed 51; expected L = numpy Cholesky;
- lap2d_4 and lap3d_3: Laplacians with the dense Cholesky factor of the
  permuted matrix for the natural order.

Run: python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def dense_from_lower(n, ptr, row, val):
    a = np.zeros((n, n))
    for j in range(n):
        for p in range(ptr[j] - 1, ptr[j + 1] - 1):
            i = row[p] - 1
            a[i, j] += val[p]
            if i != j:
                a[j, i] += val[p]
    return a


def csc_case(name, n, ptr, row, val, b=None):
    a = dense_from_lower(n, ptr, row, val)
    if b is None:
        b = a @ np.ones(n)
    x = np.linalg.solve(a, b)
    ev = np.linalg.eigvalsh(a)
    return dict(name=name, n=n, ptr=list(map(int, ptr)), row=list(map(int, row)),
                val=list(map(float, val)), b=list(map(float, b)), x=list(map(float, x)),
                num_neg=int((ev < 0).sum()), num_pos=int((ev > 0).sum()))


def lap(nx, ny, nz):
    n = nx * ny * nz
    rows, cols, vals = [], [], []
    diag = 6.0 if nz > 1 else 4.0
    for k in range(n):
        x, y, z = k % nx, (k // nx) % ny, k // (nx * ny)
        rows.append(k); cols.append(k); vals.append(diag)
        if x < nx - 1: rows.append(k + 1); cols.append(k); vals.append(-1.0)
        if y < ny - 1: rows.append(k + nx); cols.append(k); vals.append(-1.0)
        if z < nz - 1: rows.append(k + nx * ny); cols.append(k); vals.append(-1.0)
    o = np.lexsort((rows, cols))
    rows = np.array(rows)[o]; cols = np.array(cols)[o]; vals = np.array(vals)[o]
    ptr = np.concatenate([[1], 1 + np.cumsum(np.bincount(cols, minlength=n))])
    return n, ptr, rows + 1, vals


def main():
    cases = []
    cases.append(csc_case("simple_mat", 4, [1, 4, 5, 7, 8], [1, 2, 4, 2, 3, 4, 4],
                          [10.0, 2.0, 3.0, 10.0, 10.0, 4.0, 10.0]))
    cases.append(csc_case("simple_mat_indef", 4, [1, 4, 5, 7, 8], [1, 2, 4, 2, 3, 4, 4],
                          [1.0, 2.0, 3.0, 1.0, 1.0, 4.0, 1.0]))
    cases.append(csc_case("c_example", 3, [1, 3, 5, 6], [1, 2, 2, 3, 3],
                          [2.0, -1.0, 2.0, -1.0, 2.0], b=np.ones(3)))
    for name, dims in (("lap2d_4", (4, 4, 1)), ("lap3d_3", (3, 3, 3))):
        n, ptr, row, val = lap(*dims)
        c = csc_case(name, n, ptr, row, val)
        a = dense_from_lower(n, ptr, row, val)
        c["chol_natural"] = np.linalg.cholesky(a).ravel(order="F").tolist()
        cases.append(c)
    rng = np.random.default_rng(51)
    m = 51
    a = rng.uniform(-1.0, 1.0, (m, m))
    a = np.tril(a) + np.tril(a, -1).T
    for j in range(m):
        a[j, j] = abs(a[j, j]) + m
    chol = dict(name="chol_dd_51", m=m, a=a.ravel(order="F").tolist(),
                l=np.linalg.cholesky(a).ravel(order="F").tolist())
    with open(os.path.join(HERE, "small_cases.json"), "w") as f:
        json.dump(dict(cases=cases, chol_dd=chol), f)
    print("wrote", os.path.join(HERE, "small_cases.json"))


if __name__ == "__main__":
    main()
