/*
 * sylver_amd.h -- entry points of the MI355X numeric factorization beyond the
 * reference's public header.
 *
 * Three groups:
 *  1. The internal numeric-tree seam of the reference (the C ABI that the
 *     reference's Fortran factor_core and solve call into), with the same
 *     names and argument meaning; these are where the device path plugs in.
 *  2. Library-owned extension knobs (precision tier, refinement) and
 *     statistics, kept out of sylver_options_t so its ABI is unchanged.
 *  3. Read-back of the symbolic structures (host copies and the device
 *     resident copies) and the batched dense front factorization, used by
 *     the parity tests and the benchmark.
 * All functions are plain C: pointers, sizes, no exceptions.
 */
#ifndef SYLVER_AMD_EXT_H
#define SYLVER_AMD_EXT_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#include "sylver/sylver.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- flag codes (src/sylver_datatypes_mod.F90:14-45) ---- */
enum {
   SYLVER_SUCCESS = 0,
   SYLVER_ERROR_CALL_SEQUENCE = -1,
   SYLVER_ERROR_A_N_OOR = -2,
   SYLVER_ERROR_A_PTR = -3,
   SYLVER_ERROR_A_ALL_OOR = -4,
   SYLVER_ERROR_SINGULAR = -5,
   SYLVER_ERROR_NOT_POS_DEF = -6,
   SYLVER_ERROR_PTR_ROW = -7,
   SYLVER_ERROR_ORDER = -8,
   SYLVER_ERROR_VAL = -9,
   SYLVER_ERROR_X_SIZE = -10,
   SYLVER_ERROR_JOB_OOR = -11,
   SYLVER_ERROR_NOT_LLT = -13,
   SYLVER_ERROR_NOT_LDLT = -14,
   SYLVER_ERROR_NO_SAVED_SCALING = -15,
   SYLVER_ERROR_ALLOCATION = -50,
   SYLVER_ERROR_CUDA_UNKNOWN = -51, /* HIP runtime failure (or no HIP device) */
   SYLVER_ERROR_CUBLAS_UNKNOWN = -52,
   SYLVER_ERROR_UNIMPLEMENTED = -98,
   SYLVER_ERROR_UNKNOWN = -99,
   SYLVER_WARNING_IDX_OOR = 1,
   SYLVER_WARNING_DUP_IDX = 2,
   SYLVER_WARNING_DUP_AND_OOR = 3,
   SYLVER_WARNING_MISSING_DIAGONAL = 4,
   SYLVER_WARNING_MISS_DIAG_OORDUP = 5,
   SYLVER_WARNING_ANAL_SINGULAR = 6,
   SYLVER_WARNING_FACT_SINGULAR = 7,
   SYLVER_WARNING_MATCH_ORD_NO_SCALE = 8
};

/* Flags of this library beyond the reference's set (the reference has no
 * refinement): spldlt_solve with low-precision factors returns
 * SYLVER_AMD_WARNING_REFINE when refinement stopped after refine_max_iter
 * steps above refine_tol (x is the last iterate, inform tail bwd_err its
 * backward error), and SYLVER_AMD_ERROR_REFINE when a step produced a
 * non-finite backward error (x is not to be used). */
enum {
   SYLVER_AMD_WARNING_REFINE = 9,
   SYLVER_AMD_ERROR_REFINE = -60
};

/* ---- 1. internal numeric-tree seam ----
 * Reference: src/sylver_ciface.hxx:39-55 (options_c) and :73-86 (inform_c);
 * Fortran mirror src/sylver_ciface_mod.F90:9-32. */
typedef struct {
   int print_level;
   bool action;
   double small;
   double u;
   double multiplier;
   long small_subtree_threshold;
   int nb;
   int pivot_method;
   int failed_pivot_method;
   int cpu_topology;
} sylver_options_c;

typedef struct {
   int flag;
   int num_delay;
   int num_neg;
   int num_two;
   int num_zero;
   int maxfront;
   int not_first_pass;
   int not_second_pass;
} sylver_inform_c;

/* Reference: src/SymbolicTree.cxx:110-137. Builds the symbolic tree from the
 * analyse outputs (1-based arrays, kept by pointer: they must outlive the
 * tree). Subtree partitioning is a CPU-runtime concept of the reference; on
 * this path every node runs on the device, so nsubtrees may be 0. */
void *spldlt_create_symbolic_tree(void *akeep, int n, int nnodes,
                                  int const *sptr, int const *sparent,
                                  long const *rptr, int const *rlist,
                                  long const *nptr, long const *nlist,
                                  int nsubtrees, int const *subtrees,
                                  int const *small, int const *contrib_dest,
                                  int const *exec_loc);
void spldlt_destroy_symbolic_tree(void *symbolic_tree);

/* Reference: src/NumericTreePosdef.cxx:19-35 and src/NumericTree.cxx:27-44.
 * Runs the whole numeric factorization on the device; returns the numeric
 * tree, or NULL with inform->flag set. aval are the values of the (cleaned)
 * input matrix addressed by nlist; scaling may be NULL. child_contrib is
 * unused (no CPU subtrees on this path). */
void *spldlt_create_numeric_tree_posdef_dbl(void *fkeep, void *symbolic_tree,
                                            double *aval, double const *scaling,
                                            void **child_contrib,
                                            sylver_options_c *options,
                                            sylver_inform_c *inform);
void *spldlt_create_numeric_tree_dbl(bool posdef, void *fkeep,
                                     void *symbolic_tree, double *aval,
                                     double const *scaling, void **child_contrib,
                                     sylver_options_c *options,
                                     sylver_inform_c *inform);
void spldlt_destroy_numeric_tree_posdef_dbl(void *numeric_tree);
void spldlt_destroy_numeric_tree_dbl(bool posdef, void *numeric_tree);

/* Reference: src/NumericTreePosdef.cxx:46-83, src/NumericTree.cxx:56-136.
 * x is in pivot order (already permuted), column-major ldx x nrhs, host
 * memory. Return a flag. */
int spldlt_tree_solve_fwd_posdef_dbl(void const *numeric_tree, int nrhs, double *x, int ldx);
int spldlt_tree_solve_bwd_posdef_dbl(void const *numeric_tree, int nrhs, double *x, int ldx);
int spldlt_tree_solve_fwd_dbl(bool posdef, void const *numeric_tree, int nrhs, double *x, int ldx);
int spldlt_tree_solve_bwd_dbl(bool posdef, void const *numeric_tree, int nrhs, double *x, int ldx);
int spldlt_tree_solve_diag_dbl(bool posdef, void const *numeric_tree, int nrhs, double *x, int ldx);
int spldlt_tree_solve_diag_bwd_dbl(bool posdef, void const *numeric_tree, int nrhs, double *x, int ldx);

/* ---- 2. extension knobs and statistics (no counterpart in the reference) ---- */
enum {
   SYLVER_AMD_PREC_FP64 = 0,   /* fronts, updates and solves in FP64 */
   SYLVER_AMD_PREC_MIXED = 1,  /* FP32 fronts; BF16 (power-of-two block scaled) updates
                                  on fronts wider than bf16_min_ncol; FP32 accumulation;
                                  FP64 refinement in spldlt_solve */
   SYLVER_AMD_PREC_FP32 = 2    /* FP32 fronts and updates; FP64 refinement */
};
enum {
   SYLVER_AMD_REFINE_NONE = 0,
   SYLVER_AMD_REFINE_IR = 1,     /* classical iterative refinement */
   SYLVER_AMD_REFINE_GMRES = 2,  /* GMRES-IR (FP64 GMRES preconditioned by the factors) */
   SYLVER_AMD_REFINE_AUTO = 3    /* IR, falling back to GMRES-IR when IR stalls */
};

typedef struct {
   int precision;        /* SYLVER_AMD_PREC_* (default FP64) */
   int bf16_min_ncol;    /* BF16 updates on fronts with more pivots than this (1024) */
   int refine;           /* SYLVER_AMD_REFINE_* (default AUTO; ignored for FP64 factors) */
   double refine_tol;    /* target scaled backward error (1e-15) */
   int refine_max_iter;  /* (30) */
   int use_graph;        /* capture the level schedule in a hipGraph (1) */
   int tile_nb;          /* block size of the large-front path (0 = auto) */
   int small_front_max;  /* fronts with nrow <= this use the one-workgroup kernel (0 = auto) */
   int record_timing;    /* record per-level device timings (0) */
   int indef_panel_min;  /* indefinite fronts with at least this many fully-summed columns
                            use the multi-workgroup APP panel (0 = auto: 256) */
   int small_assembly;   /* one-workgroup posdef fronts: 1 assembled (zero, A entries,
                            children) inside their factor launch, one launch per level;
                            0 separate zero / scatter / extend-add launches (default:
                            at 80^3 the large fronts' extend-add launches stay, and the
                            fused form measured 0.65 ms slower in both precisions) */
   int strip_trsm;       /* posdef large fronts, levels of at most 64 of them: 1 the panel
                            TRSM by substitution in 64-row strips after a factor-only
                            diagonal launch, the blocks' inverses (for the solves) formed
                            by the factorization's last launch (default); 0 the TRSM by
                            the inverse formed in the diagonal launch; 2 1 for FP64
                            factors, 0 for FP32 ones */
   int level_assembly;   /* posdef fronts not assembled inside their factor launch: 1 a
                            level's assembly in one launch, each workgroup owning a run of
                            one front's columns (zero, original entries, every child's
                            columns landing there, children in list order -- bitwise the
                            separate launches) (default); 0 separate zero / scatter
                            launches and one extend-add launch per child rank */
   int reserved[3];
} sylver_amd_options_t;

void sylver_amd_default_options(sylver_amd_options_t *opt);
/* Process-wide extension options used by later spldlt_factorize/solve calls. */
void sylver_amd_set_options(sylver_amd_options_t const *opt);
void sylver_amd_get_options(sylver_amd_options_t *opt);

/* Written into sylver_inform_t.unused by spldlt_factorize and spldlt_solve. */
typedef struct {
   double factor_ms;      /* device time of the numeric factorization */
   double solve_ms;       /* device time of the last solve (incl. refinement) */
   double gflops;         /* num_flops / factor time */
   double bwd_err;        /* last scaled backward error seen by refinement (or -1) */
   int refine_iters;      /* refinement iterations of the last solve */
   int refine_mode;       /* SYLVER_AMD_REFINE_* actually used */
   int precision;         /* SYLVER_AMD_PREC_* of the factors */
   int nlevels;           /* levels of the device schedule */
   double upload_ms;      /* host->device copy of the values */
   char pad[16];
} sylver_amd_inform_tail_t;

typedef struct {
   double factor_ms;
   double upload_ms;
   double solve_ms;
   double flops;           /* reference flop model over all nodes */
   int nlevels;
   int nlaunches;
   int precision;
   int refine_iters;
   double bwd_err;
   /* Per launch kind (0 zero, 1 scatter A, 2 extend-add, 3 one-workgroup
    * fronts, 4 diagonal POTRF/TRTRI, 5 panel TRSM, 6 Schur update, 7 inverse
    * of small-front diagonal blocks), filled when the factorization ran with
    * record_timing != 0 (eager launches bracketed by HIP events on the
    * library's stream): summed device time, algorithmic flops, launches. */
   double kind_ms[16];
   double kind_flops[16];
   double kind_bytes[16];
   int kind_launches[16];
} sylver_amd_stats_t;

int sylver_amd_fkeep_stats(void const *fkeep, sylver_amd_stats_t *stats);

/* ---- 3. read-back for parity tests ---- */
/* Sizes: n, nnodes, length of rlist, number of nlist pairs. */
int sylver_amd_akeep_sizes(void const *akeep, int *n, int *nnodes, long *nrlist, long *nnlist);
/* Host copies of the analyse outputs (any pointer may be NULL). */
int sylver_amd_akeep_symbolic(void const *akeep, int *order, int *invp, int *sptr,
                              int *sparent, long *rptr, int *rlist, long *nptr,
                              long *nlist);
/* Copies of the device-resident structures, read back from HBM: sptr,
 * sparent, rptr, rlist, nptr, nlist (same sizes as above) and the child
 * maps (cmap, length sum over nodes of nrow-ncol; cmap_off, nnodes+1). */
int sylver_amd_fkeep_device_symbolic(void const *fkeep, int *sptr, int *sparent,
                                     long *rptr, int *rlist, long *nptr, long *nlist,
                                     int *cmap, long *cmap_off);
/* Factor of node `node` (0-based): L (nrow x ncol, column-major, ld = nrow) as
 * FP64 on the host (converted from FP32 storage if needed). */
int sylver_amd_fkeep_node_factor(void const *fkeep, int node, double *l);
/* Contribution-block memory of node `node` before its parent consumed it is
 * not kept; the assembled front of a node (nrow x nrow, before
 * factorization) is returned when the factorization ran with
 * record_timing == 2 (debug mode). */
int sylver_amd_fkeep_node_assembled(void const *fkeep, int node, double *front);
/* Indefinite factors of node `node` (0-based) of an fkeep of
 * spldlt_factorize(posdef = false) or of a numeric tree of
 * spldlt_create_numeric_tree_dbl(posdef = false): dims[4] = {m, n, nelim,
 * ndelay_in} (front rows and fully-summed columns including delayed ones,
 * eliminated pivots, delayed columns received); rows (m ints): 1-based
 * pivot-order index of each front row after pivoting (the reference's
 * NumericFront::perm followed by the row list, src/NumericTree.hxx:423-431);
 * l (m x nelim, ld m): the columns of L (unit diagonal implied); dinv
 * (2*nelim): D^{-1} as 1x1 [d, 0] and 2x2 [a, b, inf, c]
 * (src/kernels/ldlt_app.hxx:166-171). Any pointer may be NULL. */
int sylver_amd_fkeep_indef_node(void const *fkeep, int node, int *dims, int *rows, double *l,
                                double *dinv);
/* out[8] = {num_delay, num_neg, num_two, num_zero, maxfront,
 * not_first_pass, not_second_pass, 0} of the last indefinite factorization
 * (the counters of the reference's inform_c, src/sylver_ciface.hxx:73-86). */
int sylver_amd_fkeep_indef_stats(void const *fkeep, int *out);

/* Host-side check of the device launch plan (no device needed): builds the
 * plan of the analysis with the given knobs (0 = defaults) and num_cus
 * compute units, and checks every index its launches use against the
 * buffers they address (fronts in the arena, inverse blocks, task lists,
 * child maps, exchange and partial-sum slots). out[10] = {violations, factor
 * launches, forward-solve launches, backward-solve launches, arena
 * elements, exchange slots, levels, partial-sum slots, strip-TRSM
 * workgroups, diagonal-block tasks}; why (why_len bytes,
 * may be NULL) receives the first violation. Returns 0 or an error flag. */
int sylver_amd_plan_check(void const *akeep, int small_front_max, int bf16_min_ncol, int num_cus,
                          long *out, char *why, int why_len);

/* Batched dense front factorization: for each front f, the nrow[f] x nrow[f]
 * column-major block at data + offset[f] (elements, ld = nrow[f], device
 * memory) is partially Cholesky factorized over its first ncol[f] columns:
 * POTRF of the leading block, TRSM of the rows below, SYRK of the trailing
 * block (the contribution block, lower triangle updated in place).
 * dtype: 0 = double, 1 = float. info (device, nfronts ints) receives 0 or the
 * 1-based column of the first non-positive pivot. stream may be NULL.
 * Returns 0 or a SYLVER_ERROR_* code. */
int sylver_amd_batched_front_potrf(int nfronts, int const *nrow_host, int const *ncol_host,
                                   long const *offset_host, void *data_dev, int dtype,
                                   int *info_dev, void *stream);

/* Dense front LU of the unsymmetric path (the reference's kernels
 * lu_nopiv_factor, src/kernels/lu_nopiv.hxx:13-41; lu_pp_factor,
 * src/kernels/lu_pp.hxx:38-73; lu_tpp_factor, src/kernels/lu_tpp.hxx:33-118),
 * batched: for each front f the m[f] x m[f] column-major block at
 * data + offset[f] (elements, ld = m[f], device memory; 1 <= m[f] <= 128) is
 * factored over its first k[f] columns in place -- L unit lower and U upper
 * in the eliminated rows and columns, the Schur complement in the trailing
 * block. mode: 0 no pivoting, 1 partial pivoting (rows interchanged over the
 * whole front), 2 threshold partial pivoting with threshold u and column
 * interchanges among the first k[f] columns (a column that has no
 * acceptable pivot is left for later and may stay uneliminated). rperm and
 * cperm (host, sum of m[f] ints, front after front) receive each front's
 * row and column orders (0-based: row i of the factored front is row
 * rperm[i] of the input); nelim (host, nfronts ints) the eliminated columns.
 * dtype: 0 = double, 1 = float. Synchronous on stream (may be NULL); the
 * kernel's device time is then sylver_amd_last_batched_ms().
 * Returns 0 or a SYLVER_ERROR_* code. */
int sylver_amd_batched_front_lu(int nfronts, int const *m_host, int const *k_host,
                                long const *offset_host, void *data_dev, int dtype, int mode,
                                double u, int *rperm_host, int *cperm_host, int *nelim_host,
                                void *stream);

/* sylver_amd_batched_front_potrf without any device allocation. No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The per-front sizes and offsets are staged in a fixed array of
 * device memory (1 MiB per device, held until the process ends), so nfronts
 * is at most 65536 (larger: SYLVER_ERROR_CALL_SEQUENCE). Same arguments,
 * kernel and results as sylver_amd_batched_front_potrf; synchronous on
 * stream; does not time itself (sylver_amd_last_batched_ms is unchanged).
 * Calls that overlap on one device share that array and race. */
int sylver_amd_batched_front_potrf_noalloc(int nfronts, int const *nrow_host, int const *ncol_host,
                                           long const *offset_host, void *data_dev, int dtype,
                                           int *info_dev, void *stream);

/* Synthetic input of the batched small-front workload (no counterpart in
 * the reference): fills each front f (nrow[f] x nrow[f], ld nrow[f], at
 * data + offset[f] elements, device memory) with a symmetric matrix whose
 * entries are uniform in [-1, 1] from a counter-based hash of (seed, f, i, j),
 * plus nrow[f] on the diagonal (SPD by diagonal dominance). dtype as above. */
int sylver_amd_batched_front_fill(int nfronts, int const *nrow_host, long const *offset_host,
                                  void *data, int dtype, unsigned long long seed);

/* ---- Autotuner and learned cost model (no counterpart in the reference:
 * its scheduler decides by static rules). ---- */

/* Per-front kernel variant of later posdef factorizations with this akeep:
 * variant[i] = 1 one-workgroup kernel, 0 tiled large-front path, -1 the
 * default rule (nrow <= small_front_max); variant == NULL clears. nnodes
 * must equal the analysis's node count. Returns 0 or
 * SYLVER_ERROR_CALL_SEQUENCE. */
int sylver_amd_set_front_variants(void *akeep, signed char const *variant, int nnodes);

/* The batched partial factorization of sylver_amd_batched_front_potrf
 * through the launch plan: per front the one-workgroup kernel (variant[f] =
 * 1) or the tiled large-front path (0); -1 or variant == NULL the default
 * rule (nrow <= 256: one workgroup). Fronts of one variant share launches.
 * Synchronous; info[f] (device) as above; the device time is returned by
 * sylver_amd_last_batched_ms. */
int sylver_amd_batched_front_potrf_v(int nfronts, int const *nrow_host, int const *ncol_host,
                                     long const *offset_host, void *data, int dtype,
                                     signed char const *variant_host, int *info);

/* Device time (ms, median of reps after one warm-up) of factoring k
 * independent synthetic SPD fronts (m[i] rows, n[i] pivots, values from the
 * batched front generator) with the given variants (NULL = default rule) in
 * precision SYLVER_AMD_PREC_FP64 or _FP32: the autotuner's measurement. */
int sylver_amd_time_fronts(int k, int const *m, int const *n, signed char const *variant, int precision,
                           int reps, double *ms);

/* The cost model's 8 -> 64 -> 64 -> 2 MLP (ReLU, FP32, 4866 parameters:
 * W1[8][64], b1[64], W2[64][64], b2[64], W3[64][2], b3[2]); x is n x 8,
 * y and out n x 2, row-major, host memory. init: Glorot-uniform weights
 * from a seeded stream; grad: gradient of sum (y_pred - y)^2 / (2 b) at
 * params over the batch (device forward/backward kernels); train: Adam,
 * seeded shuffles, loss_hist[epoch] (epochs <= 1024); predict: forward. */
int sylver_amd_mlp_init(float *params, unsigned long long seed);
int sylver_amd_mlp_grad(float const *params, float const *x, float const *y, int b, float *grads,
                        float *loss);
int sylver_amd_mlp_train(float *params, float const *x, float const *y, int n, int epochs, int batch,
                         float lr, unsigned long long seed, float *loss_hist);
int sylver_amd_mlp_predict(float const *params, float const *x, int n, float *out);

/* ---- Scalings and matching-based ordering (host code) ----
 * The reference calls SPRAL for these (src/spldlt_factorize_mod.F90:734-835:
 * hungarian_scale_sym, auction_scale_sym; src/spldlt_analyse_mod.F90:789-820:
 * match_order_metis); spldlt_factorize selects them with options.scaling =
 * 1 / 2 and spldlt_analyse with options.ordering = 2. Exposed with SPRAL's
 * argument meaning for direct use and testing. The matrix is the lower
 * triangle, 1-based CSC; outputs are in the caller's variable order:
 * scaling[n] > 0 with every |scaling_i a_ij scaling_j| <= 1; match[n] (may
 * be NULL) the variable matched to each variable, -1 for none (0-based).
 * hungarian: returns 0, 1 (structurally singular, scaled anyway when
 * scale_if_singular) or -2 (structurally singular, not scaled). auction:
 * returns 0 or 1 (some variables unmatched within its iteration limit).
 * match_order: order[n] = 1-based pivot position of each variable, matched
 * pairs adjacent; returns 0 or 1 (structurally singular). */
int sylver_amd_hungarian_scale_sym(int n, long const *ptr, int const *row, double const *val,
                                   int scale_if_singular, double *scaling, int *match);
int sylver_amd_auction_scale_sym(int n, long const *ptr, int const *row, double const *val,
                                 double *scaling, int *match);
int sylver_amd_match_order(int n, long const *ptr, int const *row, double const *val, int *order,
                           double *scaling);

/* Number of HIP devices visible (0 when none); never fails. */
int sylver_amd_device_count(void);

/* ---- Rutherford-Boeing input (replaces SPRAL's rb_read, which the
 * reference's drivers call: drivers/spldlt_test.F90:86-95; SPRAL is not
 * part of the reference tree). ---- */
#define SYLVER_AMD_RB_WARNING_MADE_UP 1  /* pattern-only file: values made up (or zeros) */
#define SYLVER_AMD_RB_ERROR_ARGS -1
#define SYLVER_AMD_RB_ERROR_OPEN -2      /* file cannot be opened */
#define SYLVER_AMD_RB_ERROR_FORMAT -3    /* malformed header, format or data */
#define SYLVER_AMD_RB_ERROR_TYPE -4      /* complex, elemental or non-square matrix */
#define SYLVER_AMD_RB_ERROR_NO_VALUES -5 /* values = 1 and the file holds a pattern only */
#define SYLVER_AMD_RB_ERROR_ALLOC -6

/* Reads a Rutherford-Boeing (or Harwell-Boeing) assembled real or pattern
 * matrix into the lower triangle in 1-based CSC (ptr: n+1 entries, row and
 * val: ptr[n]-1 entries; malloc'ed, release each with sylver_amd_rb_free).
 * Symmetric types keep their entries (mirrored below the diagonal if stored
 * above); unsymmetric types keep row >= column. values: 0 read values if
 * present (zeros for a pattern), 1 values required, 2 values made up for a
 * pattern-only file (symmetric, diagonally dominant, from seed), as the
 * reference driver's rb_options%values = 2. type (may be NULL) receives the
 * 3-letter matrix type. Returns 0, SYLVER_AMD_RB_WARNING_MADE_UP or a
 * negative SYLVER_AMD_RB_ERROR_* code. */
int sylver_amd_rb_read(const char *filename, int values, unsigned long long seed, int *nrow, int *ncol,
                       long **ptr, int **row, double **val, char *type);
void sylver_amd_rb_free(void *p);

/* Device memory helpers for callers without their own HIP runtime (the
 * tests and benchmark use them instead of a second framework's allocator).
 * Return 0 or SYLVER_ERROR_CUDA_UNKNOWN. */
int sylver_amd_dev_alloc(void **ptr, size_t bytes);
int sylver_amd_dev_free(void *ptr);
int sylver_amd_memcpy_h2d(void *dst, const void *src, size_t bytes);
int sylver_amd_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* Device time (ms, HIP events) of the last sylver_amd_batched_front_potrf call. */
double sylver_amd_last_batched_ms(void);

#ifdef __cplusplus
}
#endif

#endif /* SYLVER_AMD_EXT_H */
