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

/* ---- indefinite path (oracle_indef.cpp) ---------------------------------
 * APP LDL^T with delayed pivots: factorize_indef_app_notask
 * (src/factor_indef.hxx:600-860), permute_failed (:1190-1289),
 * factor_front_indef_failed (src/factor_failed.hxx:29-157),
 * assemble_delays (src/kernels/assemble.hxx:926-961), the delay-aware
 * assembly maps (src/assemble.hxx:152-190) and the inertia count
 * (src/NumericTree.hxx:150-179). The pivot rule inside a diagonal block and
 * the TPP second pass are this oracle's own statement of published methods
 * (Bunch-Parlett complete pivoting; threshold partial pivoting with the
 * Duff-Reid 2x2 test): the reference delegates both to SPRAL, which is not
 * in its tree. scaling (pivot order, may be NULL) scales the scattered
 * entries as init_a_block does. Returns an opaque handle (free with
 * oracle_indef_free). */
void *oracle_factor_indef(int n, int nnodes, const int *sptr, const int *sparent,
                          const int64_t *rptr, const int *rlist, const int64_t *nptr,
                          const int64_t *nlist, const double *aval, const double *scaling,
                          int nb, double u, double small, int action, int failed_tpp);
/* out[8] = {flag, num_delay, num_neg, num_two, num_zero, maxfront,
 * not_first_pass, not_second_pass}. */
void oracle_indef_stats(const void *h, int *out);
/* dims[4] = {m, n, nelim, ndelay_in}; rows (m, 1-based pivot-order index of
 * each front row), L (m x nelim, ld m), dinv (2 nelim); any may be NULL. */
void oracle_indef_node(const void *h, int node, int *dims, int *rows, double *L, double *dinv);
/* job 0 full, 1 fwd, 2 diag, 3 bwd, 4 diag+bwd (src/NumericTree.hxx:408-532). */
void oracle_solve_indef(const void *h, int job, int nrhs, double *x, int ldx);
void oracle_indef_free(void *h);
/* One dense front (m x m full symmetric, ld m) with n fully-summed columns:
 * returns nelim (or -5 singular with action == 0); perm[n] is the position
 * permutation, dinv[2n] the D^{-1} entries, *nelim_first the APP count. */
int oracle_front_ldlt(int m, int n, double *a, int nb, double u, double small, int action,
                      int failed_tpp, int *perm, double *dinv, int *nelim_first);

/* Dense front LU of an m x m column-major front over its first k columns
 * (src/kernels/lu_nopiv.hxx, lu_pp.hxx, lu_tpp.hxx); mode 0 no pivoting, 1
 * partial pivoting, 2 threshold partial pivoting (u) with column interchanges
 * among the first k columns. rperm/cperm (m) receive the row/column orders.
 * Returns nelim (or -1 for bad arguments). */
int oracle_lu_front(int mode, double u, int m, int k, double *a, int lda, int *rperm, int *cperm);
int oracle_lu_front_f32(int mode, double u, int m, int k, float *a, int lda, int *rperm, int *cperm);

#ifdef __cplusplus
}
#endif
#endif
