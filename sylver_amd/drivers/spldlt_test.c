/*
 * spldlt_test: the reference's sparse test driver (drivers/spldlt_test.F90)
 * on this library's C interface. Reads a Rutherford-Boeing matrix, builds
 * right-hand sides b = A * ones, then analyse -> factorize -> solve, and
 * prints what the reference driver prints: the factorize and solve wall
 * times (system clock around each call, :133-148), the forward error against
 * x = 1, the scaled backward error of every right-hand side, and the
 * factorization statistics (:152-166).
 *
 * Options are the reference driver's (proc_args, :170-300): --mat FILE,
 * --nrhs N, --posdef / --indef, --scale=none|mc64|auction|mc77,
 * --ordering=mc64-metis, --nemin N, --nb N, --ncpu N, --ngpu N, --u X,
 * --print-level N, --failed-pivot-method=tpp|pass, --prune-tree,
 * --no-prune-tree; the CPU-runtime ones (--gpu-perf-coeff, --flat-topology,
 * --numa-topology, --sched=lws|hp) are accepted and have no effect here.
 * Extensions of this library: --precision=fp64|fp32|mixed (factor precision
 * tier; low-precision factors are refined in FP64 by spldlt_solve),
 * --refine=none|ir|gmres|auto.
 *
 * The backward error uses full symmetric row sums for ||A||_inf (the
 * reference driver's matrix_inf_norm sums the stored lower triangle only,
 * :376-387). Exit status: 0 success, 1 usage, 2 read failure, 3 a call
 * returned an error flag.
 */
#include <math.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "sylver/sylver.h"
#include "sylver_amd.h"

static double now_s(void) {
   struct timespec t;
   clock_gettime(CLOCK_MONOTONIC, &t);
   return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static int need_arg(int i, int argc, const char* what) {
   if (i >= argc) {
      printf("Missing value after %s\n", what);
      exit(1);
   }
   return i;
}

int main(int argc, char** argv) {
   sylver_options_t options;
   sylver_inform_t inform;
   sylver_default_options(&options);
   sylver_amd_options_t xo;
   sylver_amd_get_options(&xo);
   bool pos_def = false;
   options.ordering = 1;
   options.scaling = 0;
   options.prune_tree = true;
   options.print_level = 1;
   int nrhs = 1, ncpu = 1, ngpu = 0;
   const char* matfile = "matrix.rb";

   for (int i = 1; i < argc; ++i) {
      const char* a = argv[i];
      if (!strcmp(a, "--scale=none")) {
         options.scaling = 0;
         printf(" Set scaling to None\n");
      } else if (!strcmp(a, "--scale=mc64")) {
         options.scaling = 1;
         printf(" Set scaling to MC64\n");
      } else if (!strcmp(a, "--scale=auction")) {
         options.scaling = 2;
         printf(" Set scaling to Auction\n");
      } else if (!strcmp(a, "--scale=mc77")) {
         options.scaling = 4;
         printf(" Set scaling to MC77\n");
      } else if (!strcmp(a, "--ordering=mc64-metis")) {
         options.ordering = 2;
         options.scaling = 3;
         printf(" Using matching-based ordering (scaling overwritten)\n");
      } else if (!strcmp(a, "--nrhs")) {
         nrhs = atoi(argv[need_arg(++i, argc, a)]);
         printf(" solving for %d right-hand sides\n", nrhs);
      } else if (!strcmp(a, "--indef")) {
         pos_def = false;
         printf(" Matrix assumed indefinite\n");
      } else if (!strcmp(a, "--posdef")) {
         pos_def = true;
         printf(" Matrix assumed positive definite\n");
      } else if (!strcmp(a, "--nemin")) {
         options.nemin = atoi(argv[need_arg(++i, argc, a)]);
         printf(" Supernode amalgamation nemin = %d\n", options.nemin);
      } else if (!strcmp(a, "--nb")) {
         options.nb = atoi(argv[need_arg(++i, argc, a)]);
         printf(" CPU block size = %d\n", options.nb);
      } else if (!strcmp(a, "--ncpu")) {
         ncpu = atoi(argv[need_arg(++i, argc, a)]);
         printf(" Number of CPUs = %d\n", ncpu);
      } else if (!strcmp(a, "--ngpu")) {
         ngpu = atoi(argv[need_arg(++i, argc, a)]);
         printf(" Number of GPUs = %d\n", ngpu);
      } else if (!strcmp(a, "--mat")) {
         matfile = argv[need_arg(++i, argc, a)];
         printf(" Matrix = %s\n", matfile);
      } else if (!strcmp(a, "--failed-pivot-method=tpp")) {
         options.failed_pivot_method = 1;
         printf(" Failed pivot method TPP\n");
      } else if (!strcmp(a, "--failed-pivot-method=pass")) {
         options.failed_pivot_method = 2;
         printf(" Failed pivot method PASS\n");
      } else if (!strcmp(a, "--gpu-perf-coeff")) {
         options.gpu_perf_coeff = atof(argv[need_arg(++i, argc, a)]);
         printf(" GPU Performance coefficient = %g (no effect: one device)\n", options.gpu_perf_coeff);
      } else if (!strcmp(a, "--prune-tree")) {
         options.prune_tree = true;
         printf(" Enable tree pruning\n");
      } else if (!strcmp(a, "--no-prune-tree")) {
         options.prune_tree = false;
         printf(" Disable tree pruning\n");
      } else if (!strcmp(a, "--print-level")) {
         options.print_level = atoi(argv[need_arg(++i, argc, a)]);
         printf(" Print level = %d\n", options.print_level);
      } else if (!strcmp(a, "--u")) {
         options.u = atof(argv[need_arg(++i, argc, a)]);
         printf(" Pivoting threshold u = %g\n", options.u);
      } else if (!strcmp(a, "--flat-topology") || !strcmp(a, "--numa-topology") || !strcmp(a, "--sched=lws") ||
                 !strcmp(a, "--sched=hp")) {
         printf(" %s: CPU runtime option, no effect here\n", a);
      } else if (!strncmp(a, "--precision=", 12)) {
         const char* v = a + 12;
         xo.precision = !strcmp(v, "fp32") ? SYLVER_AMD_PREC_FP32
                        : !strcmp(v, "mixed") ? SYLVER_AMD_PREC_MIXED
                                              : SYLVER_AMD_PREC_FP64;
         printf(" Factor precision tier = %s\n", v);
      } else if (!strncmp(a, "--refine=", 9)) {
         const char* v = a + 9;
         xo.refine = !strcmp(v, "none") ? SYLVER_AMD_REFINE_NONE
                     : !strcmp(v, "ir") ? SYLVER_AMD_REFINE_IR
                     : !strcmp(v, "gmres") ? SYLVER_AMD_REFINE_GMRES
                                           : SYLVER_AMD_REFINE_AUTO;
         printf(" Refinement = %s\n", v);
      } else {
         printf(" Unrecognised command line argument: %s\n", a);
         return 1;
      }
   }
   sylver_amd_set_options(&xo);
   if (nrhs < 1) nrhs = 1;
   printf(" [spldlt_test]    ncpu: %d\n", ncpu);
   printf(" [spldlt_test]      nb: %d\n", options.nb);
   printf(" [spldlt_test] scaling: %d\n", options.scaling);

   printf("Reading...\n");
   int m = 0, n = 0;
   long* ptr = NULL;
   int* row = NULL;
   double* val = NULL;
   char type[4] = {0, 0, 0, 0};
   const int rb_flag = sylver_amd_rb_read(matfile, 2, 0ull, &m, &n, &ptr, &row, &val, type);
   if (rb_flag < 0) {
      printf(" Rutherford-Boeing read failed with error %d\n", rb_flag);
      return 2;
   }
   printf("ok\n");
   const long nz = ptr[n] - 1;

   /* b = A * ones: full symmetric row sums of the lower triangle */
   double* rhs = calloc((size_t)n * nrhs, sizeof(double));
   double* soln = malloc(sizeof(double) * (size_t)n * nrhs);
   for (int j = 0; j < n; ++j)
      for (long p = ptr[j] - 1; p < ptr[j + 1] - 1; ++p) {
         const int i = row[p] - 1;
         for (int r = 0; r < nrhs; ++r) {
            rhs[i + (size_t)r * n] += val[p];
            if (i != j) rhs[j + (size_t)r * n] += val[p];
         }
      }

   sylver_init(ncpu, ngpu);
   void *akeep = NULL, *fkeep = NULL;
   spldlt_analyse(n, NULL, ptr, row, val, &akeep, true, &options, &inform);
   if (inform.flag < 0) {
      printf(" spldlt_analyse flag %d\n", inform.flag);
      return 3;
   }
   double t0 = now_s();
   spldlt_factorize(pos_def, ptr, row, val, NULL, akeep, &fkeep, &options, &inform);
   double t1 = now_s();
   if (inform.flag < 0) {
      printf(" spldlt_factorize flag %d\n", inform.flag);
      return 3;
   }
   printf("ok\n");
   printf(" Factor took %14.6e\n", t1 - t0);
   const sylver_inform_t finf = inform;

   printf("Solve...\n");
   memcpy(soln, rhs, sizeof(double) * (size_t)n * nrhs);
   t0 = now_s();
   spldlt_solve(0, nrhs, soln, n, akeep, fkeep, &options, &inform);
   t1 = now_s();
   if (inform.flag < 0) {
      printf(" spldlt_solve flag %d\n", inform.flag);
      return 3;
   }
   printf("ok\n");
   printf(" Solve took %14.6e\n", t1 - t0);
   const sylver_amd_inform_tail_t* tail = (const sylver_amd_inform_tail_t*)inform.unused;
   if (tail->refine_iters > 0)
      printf(" refinement: %d iterations (mode %d), bwd error %10.2e\n", tail->refine_iters, tail->refine_mode,
             tail->bwd_err);
   int ist[8] = {0, 0, 0, 0, 0, 0, 0, 0};
   if (!pos_def) sylver_amd_fkeep_indef_stats(fkeep, ist);
   spldlt_free_fkeep(&fkeep);
   spldlt_free_akeep(&akeep);
   sylver_finalize();

   int bad = 0;
   double fwd = 0.0;
   for (int i = 0; i < n; ++i) {
      const double e = fabs(soln[i] - 1.0);
      if (!(e < 1e-6)) ++bad;
      if (e > fwd || e != e) fwd = e;
   }
   printf(" number bad cmp = %d\n", bad);
   printf(" fwd error || ||_inf = %14.6e\n", fwd);

   /* scaled backward error per right-hand side */
   double* res = calloc((size_t)n * nrhs, sizeof(double));
   double* rown = calloc((size_t)n, sizeof(double));
   for (int j = 0; j < n; ++j)
      for (long p = ptr[j] - 1; p < ptr[j + 1] - 1; ++p) {
         const int i = row[p] - 1;
         rown[i] += fabs(val[p]);
         if (i != j) rown[j] += fabs(val[p]);
         for (int r = 0; r < nrhs; ++r) {
            res[i + (size_t)r * n] += val[p] * soln[j + (size_t)r * n];
            if (i != j) res[j + (size_t)r * n] += val[p] * soln[i + (size_t)r * n];
         }
      }
   double normA = 0.0;
   for (int i = 0; i < n; ++i) normA = rown[i] > normA ? rown[i] : normA;
   printf(" bwd error scaled = ");
   for (int r = 0; r < nrhs; ++r) {
      double rmax = 0.0, xmax = 0.0, bmax = 0.0;
      for (int i = 0; i < n; ++i) {
         const size_t k = i + (size_t)r * n;
         const double d = fabs(res[k] - rhs[k]);
         rmax = d > rmax || d != d ? d : rmax;
         xmax = fabs(soln[k]) > xmax || soln[k] != soln[k] ? fabs(soln[k]) : xmax;
         bmax = fabs(rhs[k]) > bmax ? fabs(rhs[k]) : bmax;
      }
      const double den = normA * xmax + bmax;
      printf(" %14.6e", den > 0.0 ? rmax / den : rmax);
   }
   printf("\n");
   printf("nfact:%10.2e\n", (double)finf.num_factor);
   printf("nflop:%10.2e\n", (double)finf.num_flops);
   printf("delay:%10d\n", finf.num_delay);
   printf("2x2piv:%10d\n", finf.num_two);
   printf("maxfront:%10d\n", finf.maxfront);
   printf("not_first_pass:%10d\n", ist[5]);
   printf("not_second_pass:%10d\n", ist[6]);
   printf("num_neg:%10d\n", finf.num_neg);
   printf("nnz(lower):%10ld type:%s\n", nz, type);
   free(res);
   free(rown);
   free(rhs);
   free(soln);
   sylver_amd_rb_free(ptr);
   sylver_amd_rb_free(row);
   sylver_amd_rb_free(val);
   return 0;
}
