/*
 * sylver.h -- public C interface of the MI355X numeric factorization.
 *
 * Drop-in for the reference header include/sylver/sylver.h: the two structs
 * keep the reference's field order, types and sizes (sylver.h:19-71, which in
 * turn must match the Fortran bind(C) types in
 * src/interfaces/C/sylver_ciface.F90:9-69), and the eight entry points keep
 * their names, argument lists and semantics (sylver.h:73-123).
 *
 * Index convention: ptr/row/order are 1-based whatever array_base says. The
 * reference computes `cindexed` but never converts
 * (src/interfaces/C/sylver_ciface.F90:86,572-588), and its C example passes
 * 1-based arrays (examples/C/spldlt_simple_example_c.c:37-38).
 *
 * Errors are never thrown across this boundary: inform->flag < 0 is an error,
 * > 0 a warning, with the reference's codes
 * (src/sylver_datatypes_mod.F90:14-45).
 */
#ifndef SYLVER_AMD_SYLVER_H
#define SYLVER_AMD_SYLVER_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same layout as the reference sylver_inform_t (sylver.h:19-37). The
 * 80-byte tail is reserved by the reference for expansion; this library
 * writes the fields of sylver_amd_inform_tail_t (sylver_amd.h) there. */
typedef struct {
   int flag;
   int matrix_dup;
   int matrix_missing_diag;
   int matrix_outrange;
   int matrix_rank;
   int maxdepth;
   int maxfront;
   int num_delay;
   long num_factor;
   long num_flops;
   int num_neg;
   int num_sup;
   int num_two;
   int stat;
   int cuda_error;   /* HIP runtime error code of the last failing call */
   int cublas_error; /* unused: no vendor BLAS on this path */
   char unused[80];
} sylver_inform_t;

/* Same layout as the reference sylver_options_t (sylver.h:39-71). */
typedef struct {
   int array_base;
   int print_level;
   int unit_diagnostics;
   int unit_error;
   int unit_warning;
   int ordering;     /* 0: user order, 1: nested dissection (built here; the reference calls METIS),
                       2: matching-based (compressed 2x2 candidates, then nested dissection) */
   int nemin;        /* supernode amalgamation parameter (default 32) */
   bool prune_tree;
   long min_gpu_work;
   int scaling;      /* <= 0: the caller's `scale` if given; 1: matching (shortest augmenting
                       paths); 2: matching (auction); 3: the scaling saved by ordering = 2;
                       >= 4: norm equilibration */
   int pivot_method; /* 1 APP aggressive, 2 APP block (default), 3 TPP */
   double small;
   double u;
   long small_subtree_threshold;
   int nb;
   int cpu_topology;
   bool action;
   bool use_gpu;
   double gpu_perf_coeff;
   int failed_pivot_method;
   int scheduler;
} sylver_options_t;

/* Reference: sylver.h:73 -> sylver_mod.F90:8-26 (StarPU init). Here there is
 * no runtime to start: ncpu/ngpu are recorded, and the numeric phase runs on
 * the current HIP device (one GPU). Streams, graphs and device buffers belong
 * to the akeep/fkeep handles that spldlt_analyse/spldlt_factorize create. */
void sylver_init(int ncpu, int ngpu);

/* Reference: sylver.h:75. Waits for outstanding device work; device memory is
 * released by spldlt_free_akeep / spldlt_free_fkeep, not here. */
void sylver_finalize(void);

/* Reference: sylver.h:77 -> sylver_ciface.F90 default options
 * (defaults from src/sylver_datatypes_mod.F90:97-198). */
void sylver_default_options(sylver_options_t *options);

/* Reference: sylver.h:79-83 -> spldlt_analyse (src/spldlt_analyse_mod.F90:587-890).
 * Symbolic analysis on the CPU: input check, ordering, elimination tree,
 * supernodes with nemin amalgamation, row lists, build_map. `order` is read
 * when options->ordering == 0 and written with the final pivot order. */
void spldlt_analyse(int n, int *order, long const *ptr, int const *row,
                    double const *val, void **akeep, bool check,
                    sylver_options_t const *options, sylver_inform_t *inform);

/* Reference: sylver.h:97-100 -> spldlt_factorize (src/spldlt_factorize_mod.F90:569-898).
 * Numeric factorization on the GPU. ptr/row may be NULL (the pattern of the
 * analyse call is used). posdef selects LL^T, otherwise APP LDL^T. */
void spldlt_factorize(bool posdef, long const *ptr, int const *row,
                      double const *val, double *scale, void *akeep,
                      void **fkeep, sylver_options_t const *options,
                      sylver_inform_t *inform);

/* Reference: sylver.h:102-104 -> spldlt_solve (src/spldlt_factorize_mod.F90:1169-1286).
 * job 0 = full solve (the only job the reference's solve honours,
 * spldlt_factorize_mod.F90:968-1031); x is ldx x nrhs column-major, in/out. */
void spldlt_solve(int job, int nrhs, double *x, int ldx, void *akeep,
                  void *fkeep, sylver_options_t const *options,
                  sylver_inform_t *inform);

/* Reference: sylver.h:121-123. */
void spldlt_free_akeep(void **akeep);
void spldlt_free_fkeep(void **fkeep);

#ifdef __cplusplus
}
#endif

#endif /* SYLVER_AMD_SYLVER_H */
