/*
 * CPU oracle: a restatement, in plain C++ FP64, of the reference's numeric
 * phase for the posdef path. TEST INFRASTRUCTURE ONLY: it is the checker of
 * the device path (tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may use it) and is never called by the library.
 *
 * Every function cites the reference code it follows. Indices follow the
 * reference's conventions (1-based rlist/sptr/sparent/rptr/nptr/nlist).
 */
#ifndef SYLVER_ORACLE_H
#define SYLVER_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* build_map: src/spldlt_analyse_mod.F90:130-232. */
void oracle_build_map(int n, const int64_t *ptr, const int *row, const int *perm,
                      const int *invp, int nnodes, const int *sptr, const int64_t *rptr,
                      const int *rlist, int64_t *nptr, int64_t *nlist);

/* SymbolicTree constructor: src/SymbolicTree.cxx:7-102. Outputs per node
 * (0-based, nnodes+1 entries incl. the virtual root): parent, first_child,
 * next_child, least_desc; returns nfactor = sum nrow*ncol and *maxfront. */
int64_t oracle_symbolic_tree(int nnodes, const int *sptr, const int *sparent, const int64_t *rptr,
                             int *parent, int *first_child, int *next_child, int *least_desc,
                             int *maxfront);

/* Child maps: src/NumericTreePosdef.hxx:219-238. For each node c with a
 * parent, cmap[cmap_off[c] + k] = position in the parent's row list of row
 * rlist_c[ncol_c + k] (0-based). cmap_off has nnodes+1 entries. */
void oracle_child_maps(int n, int nnodes, const int *sptr, const int *sparent,
                       const int64_t *rptr, const int *rlist, int64_t *cmap_off, int *cmap);

/* Multifrontal LL^T factorization (NumericTreePosdef::factor_mf,
 * src/NumericTreePosdef.hxx:133-353; factor_front_posdef src/factor.hxx:19-148
 * with block size nb; init_a_block src/kernels/assemble.hxx:161-214;
 * assemble_block / assemble_contrib_block :230-517). Factor of node i is
 * written to lcol + lcol_off[i] as nrow x ncol column-major (ld = nrow).
 * lcol_off has nnodes+1 entries (filled by the oracle). Returns 0, or -6 if
 * a pivot is not positive. nthreads <= 0 uses all OpenMP threads.
 * node_mask (may be NULL) restricts the work to the flagged nodes, whose
 * children must be flagged too (used for bounded CPU timing samples). */
int oracle_factor_posdef(int n, int nnodes, const int *sptr, const int *sparent,
                         const int64_t *rptr, const int *rlist, const int64_t *nptr,
                         const int64_t *nlist, const double *aval, const double *scaling,
                         int nb, int nthreads, const char *node_mask, int64_t *lcol_off,
                         double *lcol);

/* Size of the lcol array oracle_factor_posdef needs. */
int64_t oracle_factor_size(int nnodes, const int *sptr, const int64_t *rptr);

/* Forward and backward solves (NumericTreePosdef::solve_fwd/solve_bwd,
 * src/NumericTreePosdef.hxx:357-450; cholesky_solve_fwd/bwd
 * src/kernels/llt.hxx:17-48). x in pivot order, ldx x nrhs. */
void oracle_solve_posdef(int n, int nnodes, const int *sptr, const int64_t *rptr,
                         const int *rlist, const int64_t *lcol_off, const double *lcol, int nrhs,
                         double *x, int ldx);

/* Dense partial Cholesky of one front (factorize_diag_block / solve_block /
 * update_block semantics, src/kernels/factor.hxx:34-263): a is m x m
 * column-major (ld = lda); columns [0, ncol) are factored, the trailing
 * block receives the Schur complement (lower). Returns 0 or the 1-based
 * column of the first non-positive pivot. */
int oracle_front_potrf(int m, int ncol, double *a, int lda, int nb);

/* Scaled backward error, worst over right-hand sides, with full symmetric
 * row sums (tests/common.hxx:1003-1050, tests/sylver_test_mod.F90:198-246):
 * ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf). A is the lower
 * triangle, 1-based CSC. */
double oracle_bwd_err(int n, const int64_t *ptr, const int *row, const double *val, int nrhs,
                      const double *x, int ldx, const double *b, int ldb);

#ifdef __cplusplus
}
#endif
#endif
