// CPU oracle: plain C++ FP64 restatement of the reference's posdef numeric
// phase. TEST INFRASTRUCTURE ONLY -- the checker of the device path and the
// CPU baseline of the benchmark; the library never calls it.
//
// Dense kernels are written out as loops (the reference calls BLAS/LAPACK
// through src/kernels/wrappers.cxx:11-257); the operation order per entry
// follows the reference's tiled algorithm.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Lower Cholesky of an n x n block in place (host_potrf lower). Returns 0 or
// the 1-based column of the first non-positive pivot.
int potrf_lower(int n, double* a, int lda) {
   for (int j = 0; j < n; ++j) {
      double d = a[j + (int64_t)j * lda];
      for (int k = 0; k < j; ++k) d -= a[j + (int64_t)k * lda] * a[j + (int64_t)k * lda];
      if (!(d > 0.0)) return j + 1;
      d = std::sqrt(d);
      a[j + (int64_t)j * lda] = d;
      for (int i = j + 1; i < n; ++i) {
         double s = a[i + (int64_t)j * lda];
         for (int k = 0; k < j; ++k) s -= a[i + (int64_t)k * lda] * a[j + (int64_t)k * lda];
         a[i + (int64_t)j * lda] = s / d;
      }
   }
   return 0;
}

// B := B * L^{-T} for the m x n block B (host_trsm right, lower, transposed).
void trsm_rlt(int m, int n, const double* l, int ldl, double* b, int ldb) {
   for (int j = 0; j < n; ++j) {
      double* bj = b + (int64_t)j * ldb;
      for (int k = 0; k < j; ++k) {
         const double lk = l[j + (int64_t)k * ldl];
         const double* bk = b + (int64_t)k * ldb;
         for (int i = 0; i < m; ++i) bj[i] -= bk[i] * lk;
      }
      const double inv = 1.0 / l[j + (int64_t)j * ldl];
      for (int i = 0; i < m; ++i) bj[i] *= inv;
   }
}

// C -= A * B^T over an m x nc tile; rows i < diag_off + j are skipped when
// `lower` (the diagonal tile of a symmetric update).
void gemm_nt_sub(int m, int nc, int k, const double* A, int lda, const double* B, int ldb,
                 double* C, int ldc, bool lower) {
   for (int j = 0; j < nc; ++j) {
      double* cj = C + (int64_t)j * ldc;
      const int i0 = lower ? j : 0;
      for (int l = 0; l < k; ++l) {
         const double bjl = B[j + (int64_t)l * ldb];
         const double* al = A + (int64_t)l * lda;
         for (int i = i0; i < m; ++i) cj[i] -= al[i] * bjl;
      }
   }
}

// Partial right-looking tiled Cholesky of an m x m front over its first ncol
// columns: per block column POTRF of the diagonal block, TRSM of the rows
// below, update of every trailing tile (src/factor.hxx:40-142 task order).
int front_potrf(int m, int ncol, double* a, int lda, int nb, bool parallel) {
   for (int j0 = 0; j0 < ncol; j0 += nb) {
      const int jb = std::min(nb, ncol - j0);
      int info = potrf_lower(jb, a + j0 + (int64_t)j0 * lda, lda);
      if (info) return j0 + info;
      const int r0 = j0 + jb;
      if (r0 >= m) continue;
      trsm_rlt(m - r0, jb, a + j0 + (int64_t)j0 * lda, lda, a + r0 + (int64_t)j0 * lda, lda);
      const int ntile = (m - r0 + nb - 1) / nb;
      const double* P = a + (int64_t)j0 * lda;  // panel columns j0..j0+jb, all rows
#pragma omp parallel for schedule(dynamic, 1) if (parallel)
      for (int t = 0; t < ntile; ++t) {
         const int k0 = r0 + t * nb;
         const int kb = std::min(nb, m - k0);
         gemm_nt_sub(m - k0, kb, jb, P + k0, lda, P + k0, lda, a + k0 + (int64_t)k0 * lda, lda,
                     true);
      }
   }
   return 0;
}

}  // namespace

extern "C" {

// src/spldlt_analyse_mod.F90:130-232.
void oracle_build_map(int n, const int64_t* ptr, const int* row, const int* perm, const int* invp,
                      int nnodes, const int* sptr, const int64_t* rptr, const int* rlist,
                      int64_t* nptr, int64_t* nlist) {
   // transpose of A-lower without diagonal (rows -> list of columns)
   std::vector<int64_t> start(n + 2, 0);
   int64_t nz = ptr[n] - 1;
   std::vector<int> tcol(std::max<int64_t>(nz, 1));
   std::vector<int64_t> torig(std::max<int64_t>(nz, 1));
   for (int c = 1; c <= n; ++c)
      for (int64_t p = ptr[c - 1]; p < ptr[c]; ++p)
         if (row[p - 1] != c) start[row[p - 1]]++;
   int64_t acc = 0;
   for (int r = 0; r <= n + 1; ++r) {
      int64_t t = start[r];
      start[r] = acc;
      acc += t;
   }
   std::vector<int64_t> fill(start.begin(), start.end());
   for (int c = 1; c <= n; ++c)
      for (int64_t p = ptr[c - 1]; p < ptr[c]; ++p) {
         int r = row[p - 1];
         if (r == c) continue;
         tcol[fill[r]] = c;
         torig[fill[r]++] = p;
      }
   std::vector<int> where(n + 1, 0);
   int64_t pp = 0;
   for (int nd = 0; nd < nnodes; ++nd) {
      const int nrow = (int)(rptr[nd + 1] - rptr[nd]);
      nptr[nd] = pp + 1;
      for (int64_t q = rptr[nd]; q < rptr[nd + 1]; ++q) where[rlist[q - 1]] = (int)(q - rptr[nd]) + 1;
      for (int piv = sptr[nd]; piv < sptr[nd + 1]; ++piv) {
         const int var = invp[piv - 1];
         for (int64_t q = start[var]; q < start[var + 1]; ++q) {
            int k = std::abs(perm[tcol[q] - 1]);
            if (k < piv) continue;
            nlist[2 * pp] = torig[q];
            nlist[2 * pp + 1] = (int64_t)(piv - sptr[nd]) * nrow + where[k];
            ++pp;
         }
      }
      for (int piv = sptr[nd]; piv < sptr[nd + 1]; ++piv) {
         const int var = invp[piv - 1];
         for (int64_t q = ptr[var - 1]; q < ptr[var]; ++q) {
            int k = std::abs(perm[row[q - 1] - 1]);
            if (k < piv) continue;
            nlist[2 * pp] = q;
            nlist[2 * pp + 1] = (int64_t)(piv - sptr[nd]) * nrow + where[k];
            ++pp;
         }
      }
   }
   nptr[nnodes] = pp + 1;
}

// src/SymbolicTree.cxx:7-102.
int64_t oracle_symbolic_tree(int nnodes, const int* sptr, const int* sparent, const int64_t* rptr,
                             int* parent, int* first_child, int* next_child, int* least_desc,
                             int* maxfront) {
   for (int i = 0; i <= nnodes; ++i) {
      least_desc[i] = i;
      first_child[i] = -1;
      next_child[i] = -1;
      parent[i] = -1;
   }
   int mf = 0;
   for (int i = 0; i < nnodes; ++i) {
      parent[i] = sparent[i] - 1;
      mf = std::max(mf, (int)(rptr[i + 1] - rptr[i]));
      least_desc[parent[i]] = std::min(least_desc[parent[i]], least_desc[i]);
   }
   for (int i = 0; i < nnodes; ++i) {
      int p = std::min(parent[i], nnodes);
      next_child[i] = first_child[p];
      first_child[p] = i;
   }
   int64_t nf = 0;
   for (int i = 0; i < nnodes; ++i) nf += (rptr[i + 1] - rptr[i]) * (int64_t)(sptr[i + 1] - sptr[i]);
   if (maxfront) *maxfront = mf;
   return nf;
}

// src/NumericTreePosdef.hxx:219-238.
void oracle_child_maps(int n, int nnodes, const int* sptr, const int* sparent, const int64_t* rptr,
                       const int* rlist, int64_t* cmap_off, int* cmap) {
   cmap_off[0] = 0;
   for (int i = 0; i < nnodes; ++i) {
      int64_t q = (rptr[i + 1] - rptr[i]) - (sptr[i + 1] - sptr[i]);
      cmap_off[i + 1] = cmap_off[i] + q;
   }
   std::vector<int> map(n + 1, 0);
   std::vector<int> par(nnodes + 1), fc(nnodes + 1), nc(nnodes + 1), ld(nnodes + 1);
   int mf;
   oracle_symbolic_tree(nnodes, sptr, sparent, rptr, par.data(), fc.data(), nc.data(), ld.data(), &mf);
   for (int p = 0; p < nnodes; ++p) {
      const int* rl = rlist + rptr[p] - 1;
      const int m = (int)(rptr[p + 1] - rptr[p]);
      for (int i = 0; i < m; ++i) map[rl[i]] = i;
      for (int c = fc[p]; c != -1; c = nc[c]) {
         const int* rc = rlist + rptr[c] - 1;
         const int ncol = sptr[c + 1] - sptr[c];
         for (int64_t k = 0; k < cmap_off[c + 1] - cmap_off[c]; ++k)
            cmap[cmap_off[c] + k] = map[rc[ncol + k]];
      }
   }
}

int64_t oracle_factor_size(int nnodes, const int* sptr, const int64_t* rptr) {
   int64_t s = 0;
   for (int i = 0; i < nnodes; ++i) s += (rptr[i + 1] - rptr[i]) * (int64_t)(sptr[i + 1] - sptr[i]);
   return s;
}

int oracle_front_potrf(int m, int ncol, double* a, int lda, int nb) {
   return front_potrf(m, ncol, a, lda, nb, false);
}

// src/NumericTreePosdef.hxx:133-353 with init_a_block
// (src/kernels/assemble.hxx:161-214), assemble_block (:230-286) before the
// front factorization and assemble_contrib_block (:343-517) after it.
int oracle_factor_posdef(int n, int nnodes, const int* sptr, const int* sparent, const int64_t* rptr,
                         const int* rlist, const int64_t* nptr, const int64_t* nlist,
                         const double* aval, const double* scaling, int nb, int nthreads,
                         const char* node_mask, int64_t* lcol_off, double* lcol) {
#ifdef _OPENMP
   if (nthreads > 0) omp_set_num_threads(nthreads);
   const int nthr = omp_get_max_threads();
#else
   const int nthr = 1;
#endif
   std::vector<int> par(nnodes + 1), fc(nnodes + 1), nc(nnodes + 1), ld(nnodes + 1);
   int mf;
   oracle_symbolic_tree(nnodes, sptr, sparent, rptr, par.data(), fc.data(), nc.data(), ld.data(), &mf);
   lcol_off[0] = 0;
   for (int i = 0; i < nnodes; ++i)
      lcol_off[i + 1] = lcol_off[i] + (rptr[i + 1] - rptr[i]) * (int64_t)(sptr[i + 1] - sptr[i]);
   // levels by height so that independent nodes can run concurrently; the
   // result of each node does not depend on this schedule
   std::vector<int> h(nnodes, 0);
   int nlev = 0;
   for (int i = 0; i < nnodes; ++i) {
      if (par[i] < nnodes) h[par[i]] = std::max(h[par[i]], h[i] + 1);
      nlev = std::max(nlev, h[i] + 1);
   }
   std::vector<std::vector<int>> levels(nlev);
   for (int i = 0; i < nnodes; ++i)
      if (!node_mask || node_mask[i]) levels[h[i]].push_back(i);
   std::vector<std::vector<double>> cb(nnodes);  // contribution blocks (q x q, ld q)
   std::vector<std::vector<int>> maps(nthr, std::vector<int>(n + 1, 0));
   int status = 0;

   auto process = [&](int ni, bool inner_parallel, std::vector<int>& map) -> int {
      const int m = (int)(rptr[ni + 1] - rptr[ni]);
      const int ncol = sptr[ni + 1] - sptr[ni];
      const int* rl = rlist + rptr[ni] - 1;
      std::vector<double> F((size_t)m * m, 0.0);
      // init_a_block
      for (int64_t p = nptr[ni]; p < nptr[ni + 1]; ++p) {
         int64_t src = nlist[2 * (p - 1)] - 1;
         int64_t dest = nlist[2 * (p - 1) + 1] - 1;
         int c = (int)(dest / m), r = (int)(dest % m);
         double v = aval[src];
         if (scaling) v = scaling[rl[r] - 1] * v * scaling[rl[c] - 1];
         F[(size_t)c * m + r] = v;
      }
      for (int i = 0; i < m; ++i) map[rl[i]] = i;
      // fully-summed part of the children's contributions
      for (int c = fc[ni]; c != -1; c = nc[c]) {
         const int cm = (int)(rptr[c + 1] - rptr[c]), cn = sptr[c + 1] - sptr[c], q = cm - cn;
         if (q <= 0) continue;
         const int* crl = rlist + rptr[c] - 1 + cn;
         const double* src = cb[c].data();
         for (int j = 0; j < q; ++j) {
            int col = map[crl[j]];
            if (col >= ncol) continue;
            for (int i = j; i < q; ++i) F[(size_t)col * m + map[crl[i]]] += src[(size_t)j * q + i];
         }
      }
      int info = front_potrf(m, ncol, F.data(), m, nb, inner_parallel);
      if (info) return -6;
      // contribution-block part of the children's contributions
      for (int c = fc[ni]; c != -1; c = nc[c]) {
         const int cm = (int)(rptr[c + 1] - rptr[c]), cn = sptr[c + 1] - sptr[c], q = cm - cn;
         if (q <= 0) continue;
         const int* crl = rlist + rptr[c] - 1 + cn;
         const double* src = cb[c].data();
         for (int j = 0; j < q; ++j) {
            int col = map[crl[j]];
            if (col < ncol) continue;
            for (int i = j; i < q; ++i) F[(size_t)col * m + map[crl[i]]] += src[(size_t)j * q + i];
         }
         std::vector<double>().swap(cb[c]);
      }
      double* L = lcol + lcol_off[ni];
      for (int c = 0; c < ncol; ++c)
         for (int r = 0; r < m; ++r) L[(size_t)c * m + r] = (r >= c) ? F[(size_t)c * m + r] : 0.0;
      const int q = m - ncol;
      if (q > 0) {
         cb[ni].assign((size_t)q * q, 0.0);
         for (int j = 0; j < q; ++j)
            for (int i = j; i < q; ++i) cb[ni][(size_t)j * q + i] = F[(size_t)(ncol + j) * m + ncol + i];
      }
      return 0;
   };

   for (int L = 0; L < nlev && status == 0; ++L) {
      std::vector<int>& lv = levels[L];
      if ((int)lv.size() >= 2 * nthr) {
         int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : bad)
         for (int k = 0; k < (int)lv.size(); ++k) {
#ifdef _OPENMP
            std::vector<int>& map = maps[omp_get_thread_num()];
#else
            std::vector<int>& map = maps[0];
#endif
            int r = process(lv[k], false, map);
            bad = std::min(bad, r);
         }
         status = bad;
      } else {
         for (int k = 0; k < (int)lv.size() && status == 0; ++k) status = process(lv[k], true, maps[0]);
      }
   }
   return status;
}

// src/NumericTreePosdef.hxx:357-450 with cholesky_solve_fwd/bwd
// (src/kernels/llt.hxx:17-48): TRSV on the pivot block then GEMV on the rows
// below (forward), the reverse for the backward solve.
void oracle_solve_posdef(int n, int nnodes, const int* sptr, const int64_t* rptr, const int* rlist,
                         const int64_t* lcol_off, const double* lcol, int nrhs, double* x, int ldx) {
   (void)n;
   int maxm = 0;
   for (int i = 0; i < nnodes; ++i) maxm = std::max(maxm, (int)(rptr[i + 1] - rptr[i]));
   std::vector<double> xl(maxm);
   for (int r = 0; r < nrhs; ++r) {
      double* xr = x + (int64_t)r * ldx;
      for (int ni = 0; ni < nnodes; ++ni) {
         const int m = (int)(rptr[ni + 1] - rptr[ni]), nc = sptr[ni + 1] - sptr[ni];
         const int* rl = rlist + rptr[ni] - 1;
         const double* L = lcol + lcol_off[ni];
         for (int i = 0; i < m; ++i) xl[i] = xr[rl[i] - 1];
         for (int j = 0; j < nc; ++j) {
            xl[j] /= L[j + (size_t)j * m];
            for (int i = j + 1; i < nc; ++i) xl[i] -= L[i + (size_t)j * m] * xl[j];
         }
         for (int j = 0; j < nc; ++j)
            for (int i = nc; i < m; ++i) xl[i] -= L[i + (size_t)j * m] * xl[j];
         for (int i = 0; i < m; ++i) xr[rl[i] - 1] = xl[i];
      }
      for (int ni = nnodes - 1; ni >= 0; --ni) {
         const int m = (int)(rptr[ni + 1] - rptr[ni]), nc = sptr[ni + 1] - sptr[ni];
         const int* rl = rlist + rptr[ni] - 1;
         const double* L = lcol + lcol_off[ni];
         for (int i = 0; i < m; ++i) xl[i] = xr[rl[i] - 1];
         for (int j = 0; j < nc; ++j) {
            double s = xl[j];
            for (int i = nc; i < m; ++i) s -= L[i + (size_t)j * m] * xl[i];
            xl[j] = s;
         }
         for (int j = nc - 1; j >= 0; --j) {
            double s = xl[j];
            for (int i = j + 1; i < nc; ++i) s -= L[i + (size_t)j * m] * xl[i];
            xl[j] = s / L[j + (size_t)j * m];
         }
         for (int i = 0; i < nc; ++i) xr[rl[i] - 1] = xl[i];
      }
   }
}

// tests/sylver_test_mod.F90:198-246 (compute_resid), worst column.
double oracle_bwd_err(int n, const int64_t* ptr, const int* row, const double* val, int nrhs,
                      const double* x, int ldx, const double* b, int ldb) {
   std::vector<double> res((size_t)n), work((size_t)n);
   double worst = 0.0;
   for (int r = 0; r < nrhs; ++r) {
      const double* xr = x + (int64_t)r * ldx;
      const double* br = b + (int64_t)r * ldb;
      for (int i = 0; i < n; ++i) {
         res[i] = br[i];
         work[i] = 0.0;
      }
      for (int k = 0; k < n; ++k)
         for (int64_t p = ptr[k] - 1; p < ptr[k + 1] - 1; ++p) {
            int i = row[p] - 1;
            if (i < 0 || i >= n) continue;
            double a = val[p];
            res[i] -= a * xr[k];
            work[i] += std::fabs(a);
            if (i == k) continue;
            res[k] -= a * xr[i];
            work[k] += std::fabs(a);
         }
      double anorm = 0, bnorm = 0, xnorm = 0, rnorm = 0;
      for (int k = 0; k < n; ++k) {
         anorm = std::max(anorm, work[k]);
         bnorm = std::max(bnorm, std::fabs(br[k]));
         xnorm = std::max(xnorm, std::fabs(xr[k]));
         rnorm = std::max(rnorm, std::fabs(res[k]));
         if (std::isnan(res[k])) return res[k];
      }
      worst = std::max(worst, rnorm / (anorm * xnorm + bnorm));
   }
   return worst;
}

}  // extern "C"
