/*
 * A plain C caller of the public API, written against the reference's
 * include path ("sylver/sylver.h") and call sequence
 * (examples/C/spldlt_simple_example_c.c:30-95 in the reference): the 3x3
 * tridiagonal matrix [2 -1 0; -1 2 -1; 0 -1 2] as a 1-based lower CSC,
 * treated as indefinite (posdef = 0), ordering computed by the library
 * (order = NULL), one right-hand side of ones, no scaling.
 *
 * Exit status: 0 when x = [1.5, 2, 1.5] to 1e-14; 2 when a call returned an
 * error flag (its value is printed); 3 when the solution is wrong.
 */
#include <math.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>

#include "sylver/sylver.h"

int main(void) {
   const int n = 3, nrhs = 1;
   long ptr[4] = {1, 3, 5, 6};
   int row[5] = {1, 2, 2, 3, 3};
   double val[5] = {2.0, -1.0, 2.0, -1.0, 2.0};
   double x[3] = {1.0, 1.0, 1.0};
   const double want[3] = {1.5, 2.0, 1.5};
   void *akeep = NULL, *fkeep = NULL;
   sylver_options_t options;
   sylver_inform_t inform;

   sylver_default_options(&options);
   options.ordering = 1;
   options.scaling = 0;
   sylver_init(1, 1);

   spldlt_analyse(n, NULL, ptr, row, val, &akeep, true, &options, &inform);
   if (inform.flag < 0) {
      printf("spldlt_analyse flag %d\n", inform.flag);
      return 2;
   }
   spldlt_factorize(false, NULL, NULL, val, NULL, akeep, &fkeep, &options, &inform);
   if (inform.flag < 0) {
      printf("spldlt_factorize flag %d\n", inform.flag);
      spldlt_free_akeep(&akeep);
      return 2;
   }
   spldlt_solve(0, nrhs, x, n, akeep, fkeep, &options, &inform);
   if (inform.flag < 0) {
      printf("spldlt_solve flag %d\n", inform.flag);
      return 2;
   }
   sylver_finalize();
   printf("x = %.17g %.17g %.17g\n", x[0], x[1], x[2]);
   spldlt_free_fkeep(&fkeep);
   spldlt_free_akeep(&akeep);
   if (akeep != NULL || fkeep != NULL) return 3;
   for (int i = 0; i < n; ++i)
      if (!(fabs(x[i] - want[i]) <= 1e-14)) return 3;
   return 0;
}
