// Device kernels of the indefinite (APP LDL^T) numeric phase on gfx950.
//
// A front of the indefinite path is m x m column-major (ld) in the arena with
// n fully-summed columns (own pivots, then the delayed columns of the
// children, then the remaining rows of the row list). During the panel
// factorization the fully-summed n x n block is kept symmetric in both
// triangles, so that a symmetric exchange of two pivot candidates is a swap
// of two columns (all m rows) and of two rows (n columns); rows >= n are kept
// in column storage only. The contribution block (rows and columns >= n) is
// updated afterwards by the MFMA GEMM with every eliminated pivot.
//
// Pivoting decisions follow the reference's APP semantics
// (src/factor_indef.hxx:600-860, src/kernels/ldlt_app.hxx:144-255,
// include/sylver/kernels/Column.hxx:87-101, src/factor_failed.hxx:29-157):
// per block of kAppNB columns a complete-pivoting (Bunch-Parlett) LDL^T of
// the diagonal block, the a-posteriori threshold test |l| <= 1/u on every
// uneliminated row, restore of the failed columns, update of the
// uneliminated fully-summed entries; then failed columns are moved behind the
// eliminated ones and retried by threshold partial pivoting over all rows.
// Every product-sum is written with explicit roundings (fnma / dot2 with
// contraction off) in the same order as the CPU oracle, so that a front is
// factored bit for bit like the oracle's.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace sylver_amd {

constexpr int kAppNB = 64;  // APP block size of the indefinite panel

struct IndefDesc {
   int m, n, ld, nelim;      // front rows, fully-summed columns, leading dim, eliminated
   int ncol, nrow, ndelay_in, parent;
   int first_child, next_child, delay_col0, nelim_first;
   int64_t off;              // front offset in the arena (elements)
   int64_t rlist_off, amap_off, num_a, cmap_off;
   int64_t var_off;          // m ints: 1-based pivot-order index of each front row
   int64_t pre_off;          // n ints: pre-pivot position of each fully-summed row
   int64_t d_off;            // 2 n doubles: D^{-1} (1x1 [d, 0], 2x2 [a, b, inf, c])
   int64_t ws_off;           // workspace (doubles) of the panel kernel / CB update
   int64_t x_off;            // solve work vectors: at x_off * nrhs, 2 m doubles per right-hand side
   int64_t lf_off;           // L (m x nelim, ld m) in the factor store after the level
   int64_t cb_off;           // trailing block (rows/cols nelim..m-1, ld m - nelim) in the
   double* cb_big;           //   contribution store, or in its own allocation when cb_big
};

// Trailing block of a factored front: its delayed columns and contribution
// block, (m - nelim) square, lower triangle, kept until the parent assembled it.
__device__ __forceinline__ const double* child_cb(const IndefDesc& c, const double* cbs) {
   return c.cb_big ? c.cb_big : cbs + c.cb_off;
}

__device__ __forceinline__ double fnma(double a, double b, double c) { return __builtin_fma(-a, b, c); }
__device__ __forceinline__ double dot2(double a, double b, double c, double d) {
   return __builtin_fma(a, b, c * d);
}

// Block-wide arg-max: largest value, smallest index among equal values.
// `sv`/`si` are LDS scratch of 4 entries (one per wave).
__device__ inline void block_argmax(double& v, int& idx, double* sv, int* si) {
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(v, o);
      const int oi = __shfl_xor(idx, o);
      if (ov > v || (ov == v && oi < idx)) {
         v = ov;
         idx = oi;
      }
   }
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   __syncthreads();
   if (lane == 0) {
      sv[wave] = v;
      si[wave] = idx;
   }
   __syncthreads();
   v = sv[0];
   idx = si[0];
#pragma unroll
   for (int w = 1; w < kThreads / 64; ++w)
      if (sv[w] > v || (sv[w] == v && si[w] < idx)) {
         v = sv[w];
         idx = si[w];
      }
   __syncthreads();
}

// Two block-wide arg-max reductions at once (same rule as block_argmax);
// `sv`/`si` hold 8 entries.
__device__ inline void block_argmax2(double& v0, int& i0, double& v1, int& i1, double* sv, int* si) {
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) {
      const double a0 = __shfl_xor(v0, o), a1 = __shfl_xor(v1, o);
      const int j0 = __shfl_xor(i0, o), j1 = __shfl_xor(i1, o);
      if (a0 > v0 || (a0 == v0 && j0 < i0)) {
         v0 = a0;
         i0 = j0;
      }
      if (a1 > v1 || (a1 == v1 && j1 < i1)) {
         v1 = a1;
         i1 = j1;
      }
   }
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   __syncthreads();
   if (lane == 0) {
      sv[wave] = v0;
      si[wave] = i0;
      sv[4 + wave] = v1;
      si[4 + wave] = i1;
   }
   __syncthreads();
   v0 = sv[0];
   i0 = si[0];
   v1 = sv[4];
   i1 = si[4];
#pragma unroll
   for (int w = 1; w < kThreads / 64; ++w) {
      if (sv[w] > v0 || (sv[w] == v0 && si[w] < i0)) {
         v0 = sv[w];
         i0 = si[w];
      }
      if (sv[4 + w] > v1 || (sv[4 + w] == v1 && si[4 + w] < i1)) {
         v1 = sv[4 + w];
         i1 = si[4 + w];
      }
   }
}

__device__ inline double block_max(double v, double* sv) {
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   __syncthreads();
   if (lane == 0) sv[wave] = v;
   __syncthreads();
   v = sv[0];
#pragma unroll
   for (int w = 1; w < kThreads / 64; ++w) v = fmax(v, sv[w]);
   __syncthreads();
   return v;
}

// Symmetric exchange of fully-summed positions p and q of the front.
__device__ inline void front_swap(double* F, int ld, int m, int n, int p, int q, int* var, int* pre) {
   if (p == q) return;
   for (int k = threadIdx.x; k < m; k += kThreads) {
      const double a = F[k + (int64_t)p * ld], b = F[k + (int64_t)q * ld];
      F[k + (int64_t)p * ld] = b;
      F[k + (int64_t)q * ld] = a;
   }
   __syncthreads();
   for (int k = threadIdx.x; k < n; k += kThreads) {
      const double a = F[p + (int64_t)k * ld], b = F[q + (int64_t)k * ld];
      F[p + (int64_t)k * ld] = b;
      F[q + (int64_t)k * ld] = a;
   }
   if (threadIdx.x == 0) {
      int t = var[p];
      var[p] = var[q];
      var[q] = t;
      t = pre[p];
      pre[p] = pre[q];
      pre[q] = t;
   }
   __syncthreads();
}

// x := x * D^{-1} for one pivot with the reference's zero-pivot handling
// (apply_pivot, src/kernels/ldlt_app.hxx:172-205).
__device__ __forceinline__ void apply_dinv1(double d0, double small, double& x) {
   if (d0 == 0.0)
      x = (fabs(x) < small) ? 0.0 : __builtin_inf() * x;
   else
      x = x * d0;
}

// ApplyN / ApplyT of one uneliminated row (apply_pivot and check_threshold,
// src/kernels/ldlt_app.hxx:144-255): x holds the row's kb entries in the
// block columns; x := x L11^{-T} D^{-1} with the unit lower L11 of T and the
// D^{-1} entries dl. Returns the first column with |l| > 1/u (kb if none).
__device__ __forceinline__ int apply_row(double (&x)[kAppNB], const double (*T)[kAppNB + 1],
                                         const double* dl, int kb, double small, double inv_u) {
#pragma clang fp contract(off)
#pragma unroll
   for (int j = 1; j < kAppNB; ++j)
#pragma unroll
      for (int i = 0; i < j; ++i)
         if (j < kb) x[j] = fnma(x[i], T[j][i], x[j]);
   bool second = false;  // j is the second column of a 2x2 pivot
#pragma unroll
   for (int j = 0; j < kAppNB; ++j) {
      if (j >= kb) continue;
      if (second) {
         second = false;
      } else if (j + 1 < kb && __builtin_isinf(dl[2 * j + 2])) {
         const double a1 = x[j], a2 = x[(j + 1) % kAppNB];
         x[j] = dot2(dl[2 * j], a1, dl[2 * j + 1], a2);
         x[(j + 1) % kAppNB] = dot2(dl[2 * j + 1], a1, dl[2 * j + 3], a2);
         second = true;
      } else {
         apply_dinv1(dl[2 * j], small, x[j]);
      }
   }
   int fail = kb;
#pragma unroll
   for (int j = kAppNB - 1; j >= 0; --j)
      if (j < kb && fabs(x[j]) > inv_u) fail = j;
   return fail;
}

// Passed columns of a block from the first failing one: a 2x2 pivot split by
// the threshold test is dropped (Column::adjust, include/sylver/kernels/Column.hxx:87-101).
__device__ __forceinline__ int npass_adjust(int np, const double* dl) {
   if (np > 0 && !__builtin_isinf(dl[2 * (np - 1)]) && dl[2 * (np - 1) + 1] != 0.0) --np;
   return np;
}

// (D L[s, :])_p of the update with pivot p of a block (calcLD): D is the
// inverse of the stored D^{-1}, 1x1 or 2x2; fm, f0, fp are L[s, p - 1],
// L[s, p], L[s, p + 1].
__device__ __forceinline__ double vs_value(const double* dl, int p, int npass, double fm, double f0, double fp) {
#pragma clang fp contract(off)
   if (p + 1 < npass && __builtin_isinf(dl[2 * p + 2])) {
      const double det = __builtin_fma(dl[2 * p], dl[2 * p + 3], -(dl[2 * p + 1] * dl[2 * p + 1]));
      const double D0 = dl[2 * p + 3] / det, D1 = -dl[2 * p + 1] / det;
      return dot2(D0, f0, D1, fp);
   }
   if (p > 0 && __builtin_isinf(dl[2 * p])) {
      const int q = p - 1;
      const double det = __builtin_fma(dl[2 * q], dl[2 * q + 3], -(dl[2 * q + 1] * dl[2 * q + 1]));
      const double D1 = -dl[2 * q + 1] / det, D3 = dl[2 * q] / det;
      return dot2(D1, fm, D3, f0);
   }
   const double D0 = (dl[2 * p] == 0.0) ? 0.0 : 1.0 / dl[2 * p];
   return D0 * f0;
}

// Bunch-Parlett LDL^T of the kb x kb diagonal block T (LDS, both triangles
// on entry) with complete pivoting inside the block (alpha = (1 + sqrt 17) / 8):
// 1x1 pivots [d^{-1}, 0] and 2x2 pivots [a, b, inf, c] into dl, L below the
// diagonal of T. `swap_local(p, q)` exchanges positions p and q of the block
// (T and whatever the caller keeps in step). A pivot below `small` ends the
// block: with action the rest is zero pivots, without it the return is 1.
// All threads call; the result is uniform.
template <typename Swap>
__device__ int bp_block(double (*T)[kAppNB + 1], double* dl, int kb, double small, int action,
                        double* red_v, int* red_i, Swap&& swap_local) {
#pragma clang fp contract(off)
   // thread (ti, tj): row s + ti (or s + 1 + ti ...) of the trailing block,
   // columns tj, tj + 4, ... of it
   const int tid = threadIdx.x, ti = tid & 63, tj = tid >> 6;
   for (int s = 0; s < kb;) {
      double v0 = -1.0, v1 = -1.0;
      int i0 = 0x7fffffff, i1 = 0x7fffffff;
      {
         const int i = s + ti;
         if (i < kb)
            for (int j = s + tj; j <= i; j += 4) {
               const double v = fabs(T[i][j]);
               const int idx = j * kb + i;
               if (v > v0 || (v == v0 && idx < i0)) {
                  v0 = v;
                  i0 = idx;
               }
               if (i == j && (v > v1 || (v == v1 && i < i1))) {
                  v1 = v;
                  i1 = i;
               }
            }
      }
      block_argmax2(v0, i0, v1, i1, red_v, red_i);
      if (v0 < small) {
         if (!action) return 1;
         for (int k = s + tid; k < kb; k += kThreads) {
            dl[2 * k] = 0.0;
            dl[2 * k + 1] = 0.0;
         }
         {
            const int i = s + ti;
            if (i < kb)
               for (int j = s + tj; j < i; j += 4) T[i][j] = 0.0;
         }
         __syncthreads();
         break;
      }
      // the trailing update also writes the mirror entry, so that the
      // trailing block stays symmetric for the next exchanges
      if (v1 >= 0.6403882032022076 * v0) {
         swap_local(s, i1);
         const double dv = T[s][s];
         const double d0 = 1.0 / dv;
         if (tid == 0) {
            dl[2 * s] = d0;
            dl[2 * s + 1] = 0.0;
         }
         {
            const int i = s + 1 + ti;
            if (i < kb)
               for (int j = s + 1 + tj; j <= i; j += 4) {
                  const double li = T[i][s] * d0;
                  const double x = fnma(li, T[j][s], T[i][j]);
                  T[i][j] = x;
                  if (j < i) T[j][i] = x;
               }
         }
         __syncthreads();
         for (int i = s + 1 + tid; i < kb; i += kThreads) T[i][s] = T[i][s] * d0;
         s += 1;
      } else {
         int j0 = i0 / kb, ii0 = i0 % kb;
         swap_local(s, j0);
         if (ii0 == s) ii0 = j0;
         swap_local(s + 1, ii0);
         const double a = T[s][s], b = T[s + 1][s], c = T[s + 1][s + 1];
         const double det = __builtin_fma(a, c, -(b * b));
         const double d0 = c / det, d1 = -b / det, d3 = a / det;
         if (tid == 0) {
            dl[2 * s] = d0;
            dl[2 * s + 1] = d1;
            dl[2 * s + 2] = __builtin_inf();
            dl[2 * s + 3] = d3;
         }
         {
            const int i = s + 2 + ti;
            if (i < kb)
               for (int j = s + 2 + tj; j <= i; j += 4) {
                  const double a1 = T[i][s], a2 = T[i][s + 1];
                  const double l0 = dot2(d0, a1, d1, a2), l1 = dot2(d1, a1, d3, a2);
                  const double x = fnma(l1, T[j][s + 1], fnma(l0, T[j][s], T[i][j]));
                  T[i][j] = x;
                  if (j < i) T[j][i] = x;
               }
         }
         __syncthreads();
         for (int i = s + 2 + tid; i < kb; i += kThreads) {
            const double a1 = T[i][s], a2 = T[i][s + 1];
            T[i][s] = dot2(d0, a1, d1, a2);
            T[i][s + 1] = dot2(d1, a1, d3, a2);
         }
         if (tid == 0) T[s + 1][s] = 0.0;
         s += 2;
      }
      __syncthreads();
   }
   return 0;
}

// Second pass on the failed columns nelim..n-1 (ldlt_tpp_factor semantics,
// src/factor_failed.hxx:29-157): threshold partial pivoting over all rows,
// 1x1 when |a_pp| >= u max|a_ip|, else a 2x2 with the largest off-diagonal
// row when the Duff-Reid growth test passes; zero pivots below `small`
// (with action). One workgroup; nelim is advanced. Returns 1 when a zero
// pivot stops the factorization with action == 0.
__device__ int tpp_pass(double* F, int ld, int m, int n, int& nelim, int* var, int* pre, double* D, double* TL,
                        double u, double small, int action, double* red_v, int* red_i) {
#pragma clang fp contract(off)
   const int tid = threadIdx.x;
   while (nelim < n) {
      int piv = -1, piv_t = -1, w = 0;
      for (int p = nelim; p < n; ++p) {
         double amax = 0.0;
         int t = 0x7fffffff;
         for (int i = nelim + tid; i < m; i += kThreads) {
            if (i == p) continue;
            const double v = fabs(F[i + (int64_t)p * ld]);
            if (v > amax || (v == amax && v > 0.0 && i < t)) {
               amax = v;
               t = i;
            }
         }
         block_argmax(amax, t, red_v, red_i);
         if (!(amax > 0.0)) t = -1;
         const double app = fabs(F[p + (int64_t)p * ld]);
         if (fmax(app, amax) < small) {
            if (!action) return 1;
            piv = p;
            w = 0;  // zero pivot
            break;
         }
         if (app >= u * amax) {
            piv = p;
            w = 1;
            break;
         }
         if (t >= 0 && t < n) {
            const double a11 = F[p + (int64_t)p * ld], a21 = F[t + (int64_t)p * ld],
                         a22 = F[t + (int64_t)t * ld];
            const double det = __builtin_fma(a11, a22, -(a21 * a21));
            if (det != 0.0) {
               const double d11 = a22 / det, d21 = -a21 / det, d22 = a11 / det;
               double pm = 0.0, tm = 0.0;
               for (int i = nelim + tid; i < m; i += kThreads) {
                  if (i == p || i == t) continue;
                  pm = fmax(pm, fabs(F[i + (int64_t)p * ld]));
                  tm = fmax(tm, fabs(F[i + (int64_t)t * ld]));
               }
               pm = block_max(pm, red_v);
               tm = block_max(tm, red_v);
               const double g = fmax(dot2(fabs(d11), pm, fabs(d21), tm), dot2(fabs(d21), pm, fabs(d22), tm));
               if (u * g <= 1.0) {
                  piv = p;
                  piv_t = t;
                  w = 2;
                  break;
               }
            }
         }
      }
      if (piv < 0) break;
      front_swap(F, ld, m, n, piv, nelim, var, pre);
      if (w == 0) {
         for (int i = nelim + 1 + tid; i < m; i += kThreads) {
            F[i + (int64_t)nelim * ld] = 0.0;
            if (i < n) F[nelim + (int64_t)i * ld] = 0.0;
         }
         if (tid == 0) {
            D[2 * nelim] = 0.0;
            D[2 * nelim + 1] = 0.0;
         }
         __syncthreads();
         nelim += 1;
         continue;
      }
      double d[4];
      if (w == 2) {
         int t = piv_t;
         if (t == nelim) t = piv;
         front_swap(F, ld, m, n, t, nelim + 1, var, pre);
         const double b11 = F[nelim + (int64_t)nelim * ld], b21 = F[nelim + 1 + (int64_t)nelim * ld],
                      b22 = F[nelim + 1 + (int64_t)(nelim + 1) * ld];
         const double dt = __builtin_fma(b11, b22, -(b21 * b21));
         d[0] = b22 / dt;
         d[1] = -b21 / dt;
         d[2] = __builtin_inf();
         d[3] = b11 / dt;
      } else {
         d[0] = 1.0 / F[nelim + (int64_t)nelim * ld];
         d[1] = 0.0;
      }
      // l = v D^{-1} for every row below; v kept for the update
      double* lv = TL;             // 2m
      double* vv = TL + 2 * (int64_t)m;
      const int r0 = nelim + w;
      for (int r = r0 + tid; r < m; r += kThreads) {
         const double a1 = F[r + (int64_t)nelim * ld];
         const double a2 = (w == 2) ? F[r + (int64_t)(nelim + 1) * ld] : 0.0;
         vv[2 * r] = a1;
         vv[2 * r + 1] = a2;
         if (w == 2) {
            lv[2 * r] = dot2(d[0], a1, d[1], a2);
            lv[2 * r + 1] = dot2(d[1], a1, d[3], a2);
         } else {
            double x = a1;
            apply_dinv1(d[0], small, x);
            lv[2 * r] = x;
            lv[2 * r + 1] = 0.0;
         }
      }
      __syncthreads();
      for (int64_t e = tid; e < (int64_t)(m - r0) * (n - r0); e += kThreads) {
         const int r = r0 + (int)(e % (m - r0)), s = r0 + (int)(e / (m - r0));
         double x = fnma(lv[2 * r], vv[2 * s], F[r + (int64_t)s * ld]);
         if (w == 2) x = fnma(lv[2 * r + 1], vv[2 * s + 1], x);
         F[r + (int64_t)s * ld] = x;
      }
      __syncthreads();
      for (int64_t e = tid; e < (int64_t)(n - r0) * (n - r0); e += kThreads) {
         const int i = r0 + (int)(e % (n - r0)), j = r0 + (int)(e / (n - r0));
         if (i > j) F[j + (int64_t)i * ld] = F[i + (int64_t)j * ld];
      }
      for (int r = r0 + tid; r < m; r += kThreads) {
         F[r + (int64_t)nelim * ld] = lv[2 * r];
         if (r < n) F[nelim + (int64_t)r * ld] = lv[2 * r];
         if (w == 2) {
            F[r + (int64_t)(nelim + 1) * ld] = lv[2 * r + 1];
            if (r < n) F[nelim + 1 + (int64_t)r * ld] = lv[2 * r + 1];
         }
      }
      if (tid == 0) {
         D[2 * nelim] = d[0];
         D[2 * nelim + 1] = d[1];
         if (w == 2) {
            D[2 * nelim + 2] = d[2];
            D[2 * nelim + 3] = d[3];
            F[nelim + 1 + (int64_t)nelim * ld] = 0.0;
            F[nelim + (int64_t)(nelim + 1) * ld] = 0.0;
         }
      }
      __syncthreads();
      nelim += w;
   }
   return 0;
}

// ---------------------------------------------------------------------------
// Panel factorization of one front per workgroup: APP over blocks of kAppNB
// columns, permute_failed, TPP on the failed columns. Writes nelim,
// nelim_first, the D^{-1} entries, and permutes var/pre. status[node] gets
// -5 on a zero pivot with action == 0.
// Workspace per node (doubles): backup m*kAppNB | permute scratch m*n |
// TPP scratch 4m | flags and lists 2m ints (stored in m doubles).
// ---------------------------------------------------------------------------
static __global__ __launch_bounds__(kThreads) void k_app_front(const int* __restrict__ list,
                                                        IndefDesc* __restrict__ nodes,
                                                        double* __restrict__ arena,
                                                        int* __restrict__ vars, int* __restrict__ pres,
                                                        double* __restrict__ dinv,
                                                        double* __restrict__ wsbase, double u,
                                                        double small, int action, int failed_tpp,
                                                        int* __restrict__ status, int* __restrict__ out) {
#pragma clang fp contract(off)
   constexpr int NB = kAppNB;
   __shared__ double T[NB][NB + 1];
   __shared__ double VS[NB][NB + 1];
   __shared__ double dl[2 * NB];
   __shared__ double red_v[8];
   __shared__ int red_i[8];
   __shared__ int s_fail, s_npass, s_bad;
   const int node = list[blockIdx.x];
   const IndefDesc nd = nodes[node];
   const int m = nd.m, n = nd.n, ld = nd.ld, tid = threadIdx.x;
   if (n <= 0) {
      if (tid == 0) {
         nodes[node].nelim = 0;
         nodes[node].nelim_first = 0;
         out[2 * blockIdx.x] = 0;
         out[2 * blockIdx.x + 1] = 0;
      }
      return;
   }
   double* F = arena + nd.off;
   int* var = vars + nd.var_off;
   int* pre = pres + nd.pre_off;
   double* D = dinv + nd.d_off;
   double* Bk = wsbase + nd.ws_off;
   double* PS = Bk + (int64_t)m * NB;
   double* TL = PS + (int64_t)m * n;           // TPP scratch: l (2m) and v (2m)
   int* EL = reinterpret_cast<int*>(TL + 4 * (int64_t)m);  // m flags
   int* NP = EL + m;                                       // m ints: new positions
   if (tid == 0) s_bad = 0;
   // full symmetric fully-summed block; nothing eliminated yet
   for (int64_t e = tid; e < (int64_t)n * n; e += kThreads) {
      const int i = (int)(e % n), j = (int)(e / n);
      if (i < j) F[i + (int64_t)j * ld] = F[j + (int64_t)i * ld];
   }
   for (int i = tid; i < m; i += kThreads) EL[i] = 0;
   for (int i = tid; i < 2 * n; i += kThreads) D[i] = 0.0;
   __syncthreads();
   int nelim = 0;
   for (int c0 = 0; c0 < n; c0 += NB) {
      const int kb = min(NB, n - c0);
      for (int64_t e = tid; e < (int64_t)kb * m; e += kThreads) {
         const int i = (int)(e % m), j = (int)(e / m);
         Bk[i + (int64_t)j * m] = F[i + (int64_t)(c0 + j) * ld];
      }
      for (int e = tid; e < kb * kb; e += kThreads) {
         const int i = e % kb, j = e / kb;
         T[i][j] = F[c0 + i + (int64_t)(c0 + j) * ld];
      }
      __syncthreads();
      // ---- Bunch-Parlett LDL^T of the diagonal block (in LDS) ----
      auto swap_local = [&](int p, int q) {
         if (p == q) return;
         for (int k = tid; k < kb; k += kThreads) {
            double a = T[k][p];
            T[k][p] = T[k][q];
            T[k][q] = a;
         }
         __syncthreads();
         for (int k = tid; k < kb; k += kThreads) {
            double a = T[p][k];
            T[p][k] = T[q][k];
            T[q][k] = a;
         }
         for (int k = tid; k < m; k += kThreads) {
            double a = Bk[k + (int64_t)p * m];
            Bk[k + (int64_t)p * m] = Bk[k + (int64_t)q * m];
            Bk[k + (int64_t)q * m] = a;
         }
         __syncthreads();
         for (int k = tid; k < kb; k += kThreads) {
            double a = Bk[c0 + p + (int64_t)k * m];
            Bk[c0 + p + (int64_t)k * m] = Bk[c0 + q + (int64_t)k * m];
            Bk[c0 + q + (int64_t)k * m] = a;
         }
         __syncthreads();
         front_swap(F, ld, m, n, c0 + p, c0 + q, var, pre);
      };
      if (bp_block(T, dl, kb, small, action, red_v, red_i, swap_local)) {
         if (tid == 0) s_bad = 1;
         __syncthreads();
      }
      if (s_bad) break;
      // ---- ApplyN / ApplyT on every other uneliminated row + threshold ----
      if (tid == 0) s_fail = kb;
      __syncthreads();
      const double inv_u = 1.0 / u;
      for (int r = tid; r < m; r += kThreads) {
         if (EL[r] || (r >= c0 && r < c0 + kb)) continue;
         double x[NB];
#pragma unroll
         for (int j = 0; j < NB; ++j) x[j] = (j < kb) ? F[r + (int64_t)(c0 + j) * ld] : 0.0;
         const int fail = apply_row(x, T, dl, kb, small, inv_u);
         if (fail < kb) atomicMin(&s_fail, fail);
#pragma unroll
         for (int j = 0; j < NB; ++j)
            if (j < kb) {
               F[r + (int64_t)(c0 + j) * ld] = x[j];
               if (r < n) F[c0 + j + (int64_t)r * ld] = x[j];
            }
      }
      __syncthreads();
      if (tid == 0) {
         s_npass = npass_adjust(s_fail, dl);
      }
      __syncthreads();
      const int npass = s_npass;
      // L11 for the passed columns (rows of the block below each pivot)
      for (int e = tid; e < kb * kb; e += kThreads) {
         const int i = e % kb, j = e / kb;
         if (j < npass && i > j) {
            F[c0 + i + (int64_t)(c0 + j) * ld] = T[i][j];
            F[c0 + j + (int64_t)(c0 + i) * ld] = T[i][j];
         }
      }
      // failed columns restored from the backup on every uneliminated row
      for (int64_t e = tid; e < (int64_t)(kb - npass) * m; e += kThreads) {
         const int i = (int)(e % m), j = npass + (int)(e / m);
         if (EL[i] || (i >= c0 && i < c0 + npass)) continue;
         const double b = Bk[i + (int64_t)j * m];
         F[i + (int64_t)(c0 + j) * ld] = b;
         if (i < n) F[c0 + j + (int64_t)i * ld] = b;
      }
      __syncthreads();
      // ---- update of the uneliminated fully-summed columns ----
      if (npass > 0) {
         for (int s0 = 0; s0 < n; s0 += NB) {
            const int sc = min(NB, n - s0);
            // VS[t][p] = (D L[s, :])_p for column s = s0 + t
            for (int e = tid; e < sc * npass; e += kThreads) {
               const int t = e % sc, p = e / sc;
               const int s = s0 + t;
               const double* fs = F + s + (int64_t)c0 * ld;
               const double fm = p > 0 ? fs[(int64_t)(p - 1) * ld] : 0.0;
               const double fp = p + 1 < npass ? fs[(int64_t)(p + 1) * ld] : 0.0;
               VS[t][p] = vs_value(dl, p, npass, fm, fs[(int64_t)p * ld], fp);
            }
            __syncthreads();
            for (int r = tid; r < m; r += kThreads) {
               if (EL[r] || (r >= c0 && r < c0 + npass)) continue;
               double lr[NB];
#pragma unroll
               for (int p = 0; p < NB; ++p) lr[p] = (p < npass) ? F[r + (int64_t)(c0 + p) * ld] : 0.0;
               for (int t = 0; t < sc; ++t) {
                  const int s = s0 + t;
                  if (EL[s] || (s >= c0 && s < c0 + npass)) continue;
                  double x = F[r + (int64_t)s * ld];
#pragma unroll
                  for (int p = 0; p < NB; ++p)
                     if (p < npass) x = fnma(lr[p], VS[t][p], x);
                  F[r + (int64_t)s * ld] = x;
               }
            }
            __syncthreads();
         }
         // symmetric copy of the updated fully-summed block
         for (int64_t e = tid; e < (int64_t)n * n; e += kThreads) {
            const int i = (int)(e % n), j = (int)(e / n);
            if (i > j && !EL[i] && !EL[j] && !(i >= c0 && i < c0 + npass) && !(j >= c0 && j < c0 + npass))
               F[j + (int64_t)i * ld] = F[i + (int64_t)j * ld];
         }
      }
      __syncthreads();
      for (int j = tid; j < npass; j += kThreads) {
         EL[c0 + j] = 1;
         D[2 * (c0 + j)] = dl[2 * j];
         D[2 * (c0 + j) + 1] = dl[2 * j + 1];
      }
      nelim += npass;
      __syncthreads();
   }
   if (s_bad) {
      if (tid == 0) {
         status[node] = -5;
         out[2 * blockIdx.x] = 0;
         out[2 * blockIdx.x + 1] = 0;
      }
      return;
   }
   const int nelim_first = nelim;
   // ---- permute_failed: eliminated positions first, then failed ones ----
   if (nelim < n) {
      if (tid == 0) {
         int k = 0;
         for (int p = 0; p < n; ++p)
            if (EL[p]) NP[k++] = p;
         for (int p = 0; p < n; ++p)
            if (!EL[p]) NP[k++] = p;
      }
      __syncthreads();
      for (int64_t e = tid; e < (int64_t)m * n; e += kThreads) {
         const int i = (int)(e % m), k = (int)(e / m);
         const int oi = (i < n) ? NP[i] : i;
         PS[i + (int64_t)k * m] = F[oi + (int64_t)NP[k] * ld];
      }
      // var / pre / D through the backup area (free now)
      int* ivar = reinterpret_cast<int*>(Bk);
      int* ipre = ivar + n;
      double* dd = Bk + n;
      for (int k = tid; k < n; k += kThreads) {
         ivar[k] = var[NP[k]];
         ipre[k] = pre[NP[k]];
         dd[2 * k] = D[2 * NP[k]];
         dd[2 * k + 1] = D[2 * NP[k] + 1];
      }
      __syncthreads();
      for (int64_t e = tid; e < (int64_t)m * n; e += kThreads) {
         const int i = (int)(e % m), k = (int)(e / m);
         F[i + (int64_t)k * ld] = PS[i + (int64_t)k * m];
      }
      for (int k = tid; k < n; k += kThreads) {
         var[k] = ivar[k];
         pre[k] = ipre[k];
         D[2 * k] = dd[2 * k];
         D[2 * k + 1] = dd[2 * k + 1];
      }
      __syncthreads();
   }
   // ---- second pass: threshold partial pivoting on the failed columns ----
   if (nelim < n && (m == n || failed_tpp)) {
      if (tpp_pass(F, ld, m, n, nelim, var, pre, D, TL, u, small, action, red_v, red_i)) {
         if (tid == 0) {
            status[node] = -5;
            out[2 * blockIdx.x] = nelim;
            out[2 * blockIdx.x + 1] = nelim_first;
         }
         return;
      }
   }
   if (tid == 0) {
      nodes[node].nelim = nelim;
      nodes[node].nelim_first = nelim_first;
      out[2 * blockIdx.x] = nelim;
      out[2 * blockIdx.x + 1] = nelim_first;
   }
}

// ---------------------------------------------------------------------------
// Multi-workgroup panel of one large front. The same APP steps as
// k_app_front, split into launches per block of kAppNB columns, each launch
// spread over the whole GPU except the two small bookkeeping ones:
//   k_lp_diag    Bunch-Parlett of the diagonal block (one workgroup), the
//                block permutation recorded instead of applied;
//   k_lp_perm    permutation of the block's rows and columns, backup;
//   k_lp_apply   ApplyN/ApplyT + threshold on every uneliminated row;
//   k_lp_finish  passed count, L11, D^{-1}, failed-column list;
//   k_lp_update  restore of the failed columns and update of every
//                uneliminated fully-summed column (lower triangle, mirrored),
//                128 x 128 tiles, one FMA chain per entry in pivot order.
// Then k_lp_np / k_lp_pgather / k_lp_pscatter (permute_failed) and k_lp_tail
// (the second pass). Every entry goes through the same arithmetic, in the
// same order, as in k_app_front, so the two give identical factors.
// The panel state follows the single-workgroup workspace: block T, its
// D^{-1}, counters, the block permutation and the list of failed columns.
// ---------------------------------------------------------------------------
enum {
   kStFail = 0, kStNpass = 1, kStBad = 2, kStNelim = 3, kStNf = 4, kStNfOld = 5,
   kStW = 6,        // second pass: width of the current pivot (0 = zero pivot or none)
   kStDone = 7,     // second pass: no further pivot
   kStTppErr = 8,   // second pass: zero pivot with action == 0
   kStFirst = 9     // eliminated columns after the first pass
};

struct AppWs {
   double* Bk;  // m x kAppNB backup of the block columns
   double* PS;  // m x n permutation scratch
   double* TL;  // 4 m second-pass scratch
   int* EL;     // m eliminated flags
   int* NP;     // m new positions
   double* TG;  // kAppNB x kAppNB factored diagonal block
   double* DL;  // 2 kAppNB D^{-1} entries of the block
   int* ST;     // 16 counters (kSt*)
   int* PERM;   // kAppNB block permutation
   int* FL;     // n failed columns, ascending
   double* TS;  // 4 n second-pass column statistics
   double* DL2; // 2 kAppNB D^{-1} entries of the odd blocks (DL: even blocks)
};

__host__ __device__ inline int64_t app_ws_elems(int m, int n) {
   return (int64_t)m * (kAppNB + n + 5) + 8 + kAppNB * kAppNB + 2 * kAppNB + 8 + kAppNB / 2 + (n + 1) / 2 + 8 +
          4 * (int64_t)n + 8 + 2 * kAppNB;
}

__device__ inline AppWs app_ws(double* base, int m, int n) {
   AppWs w;
   w.Bk = base;
   w.PS = base + (int64_t)m * kAppNB;
   w.TL = w.PS + (int64_t)m * n;
   w.EL = reinterpret_cast<int*>(w.TL + 4 * (int64_t)m);
   w.NP = w.EL + m;
   w.TG = base + (int64_t)m * (kAppNB + n + 5) + 8;
   w.DL = w.TG + kAppNB * kAppNB;
   w.ST = reinterpret_cast<int*>(w.DL + 2 * kAppNB);
   w.PERM = w.ST + 16;
   w.FL = w.PERM + kAppNB;
   w.TS = w.TG + kAppNB * kAppNB + 2 * kAppNB + 8 + kAppNB / 2 + (n + 1) / 2 + 8;
   w.DL2 = w.TS + 4 * (int64_t)n + 8;
   return w;
}

// Both triangles of the fully-summed block (64 x 64 tiles through LDS),
// cleared flags, D^{-1} and counters.
static __global__ __launch_bounds__(kThreads) void k_lp_start(int node, const IndefDesc* __restrict__ nodes,
                                                       double* __restrict__ arena, double* __restrict__ dinv,
                                                       double* __restrict__ wsbase) {
   __shared__ double tile[64][65];
   const IndefDesc nd = nodes[node];
   const int m = nd.m, n = nd.n, ld = nd.ld, tid = threadIdx.x;
   double* F = arena + nd.off;
   const AppWs w = app_ws(wsbase + nd.ws_off, m, n);
   const int64_t gid = (int64_t)blockIdx.x * kThreads + tid, G = (int64_t)gridDim.x * kThreads;
   for (int64_t i = gid; i < m; i += G) w.EL[i] = 0;
   for (int64_t i = gid; i < 2 * (int64_t)n; i += G) dinv[nd.d_off + i] = 0.0;
   if (gid < 16) w.ST[gid] = 0;
   const int nt = (n + 63) / 64;
   const int64_t ntile = (int64_t)nt * (nt + 1) / 2;
   for (int64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
      int bj = 0;  // tile column of the lower tile (bi >= bj)
      int64_t rem = t;
      while (rem >= nt - bj) {
         rem -= nt - bj;
         ++bj;
      }
      const int bi = bj + (int)rem;
      __syncthreads();
      for (int e = tid; e < 64 * 64; e += kThreads) {
         const int i = e & 63, j = e >> 6;
         const int r = bi * 64 + i, c = bj * 64 + j;
         tile[i][j] = (r < n && c < n) ? F[r + (int64_t)c * ld] : 0.0;
      }
      __syncthreads();
      for (int e = tid; e < 64 * 64; e += kThreads) {
         const int i = e & 63, j = e >> 6;  // entry (bj*64 + i, bi*64 + j) of the upper tile
         const int r = bj * 64 + i, c = bi * 64 + j;
         if (r < n && c < n && r < c) F[r + (int64_t)c * ld] = tile[j][i];
      }
   }
}

// Bunch-Parlett of the block at c0 (one workgroup): T, D^{-1} and the block
// permutation into the panel state; the threshold minimum starts at kb.
static __global__ __launch_bounds__(kThreads) void k_lp_diag(int node, int c0, int par, const IndefDesc* __restrict__ nodes,
                                                      const double* __restrict__ arena,
                                                      double* __restrict__ wsbase, double small, int action) {
#pragma clang fp contract(off)
   constexpr int NB = kAppNB;
   __shared__ double T[NB][NB + 1];
   __shared__ double dl[2 * NB];
   __shared__ double red_v[8];
   __shared__ int red_i[8];
   __shared__ int perm[NB];
   const IndefDesc nd = nodes[node];
   const int m = nd.m, n = nd.n, ld = nd.ld, tid = threadIdx.x;
   const AppWs w = app_ws(wsbase + nd.ws_off, m, n);
   if (w.ST[kStBad]) return;
   const int kb = min(NB, n - c0);
   const double* F = arena + nd.off;
   for (int e = tid; e < kb * kb; e += kThreads) {
      const int i = e % kb, j = e / kb;
      T[i][j] = F[c0 + i + (int64_t)(c0 + j) * ld];
   }
   if (tid < NB) perm[tid] = tid;
   __syncthreads();
   auto swap_local = [&](int p, int q) {
      if (p == q) return;
      for (int k = tid; k < kb; k += kThreads) {
         const double a = T[k][p];
         T[k][p] = T[k][q];
         T[k][q] = a;
      }
      __syncthreads();
      for (int k = tid; k < kb; k += kThreads) {
         const double a = T[p][k];
         T[p][k] = T[q][k];
         T[q][k] = a;
      }
      if (tid == 0) {
         const int t = perm[p];
         perm[p] = perm[q];
         perm[q] = t;
      }
      __syncthreads();
   };
   if (bp_block(T, dl, kb, small, action, red_v, red_i, swap_local)) {
      if (tid == 0) w.ST[kStBad] = 1;
      return;
   }
   for (int e = tid; e < kb * kb; e += kThreads) {
      const int i = e % kb, j = e / kb;
      w.TG[i + j * NB] = T[i][j];
   }
   double* DLp = par ? w.DL2 : w.DL;
   for (int k = tid; k < 2 * kb; k += kThreads) DLp[k] = dl[k];
   if (tid < kb) w.PERM[tid] = perm[tid];
   if (tid == 0) w.ST[kStFail] = kb;
}

// The block permutation applied to the front (what the symmetric exchanges
// of k_app_front do: columns c0.. over all rows, rows c0.. over the
// fully-summed columns, var / pre) and the backup of the permuted block
// columns. Workgroup 0 takes the diagonal block; the others one row (or one
// fully-summed column) per thread.
static __global__ __launch_bounds__(kThreads) void k_lp_perm(int node, int c0, const IndefDesc* __restrict__ nodes,
                                                      double* __restrict__ arena, int* __restrict__ vars,
                                                      int* __restrict__ pres, double* __restrict__ wsbase) {
   constexpr int NB = kAppNB;
   __shared__ int pm[NB];
   __shared__ double Tb[NB][NB + 1];
   const IndefDesc nd = nodes[node];
   const int m = nd.m, n = nd.n, ld = nd.ld, tid = threadIdx.x;
   const AppWs w = app_ws(wsbase + nd.ws_off, m, n);
   if (w.ST[kStBad]) return;
   const int kb = min(NB, n - c0);
   double* F = arena + nd.off;
   if (tid < NB) pm[tid] = tid < kb ? w.PERM[tid] : tid;
   __syncthreads();
   if (blockIdx.x == 0) {
      int* var = vars + nd.var_off;
      int* pre = pres + nd.pre_off;
      for (int e = tid; e < kb * kb; e += kThreads) {
         const int i = e % kb, j = e / kb;
         Tb[i][j] = F[c0 + pm[i] + (int64_t)(c0 + pm[j]) * ld];
      }
      int vv = 0, pp = 0;
      if (tid < kb) {
         vv = var[c0 + pm[tid]];
         pp = pre[c0 + pm[tid]];
      }
      __syncthreads();
      for (int e = tid; e < kb * kb; e += kThreads) {
         const int i = e % kb, j = e / kb;
         F[c0 + i + (int64_t)(c0 + j) * ld] = Tb[i][j];
         w.Bk[c0 + i + (int64_t)j * m] = Tb[i][j];
      }
      if (tid < kb) {
         var[c0 + tid] = vv;
         pre[c0 + tid] = pp;
      }
      return;
   }
   const int64_t items = (int64_t)m + n;  // the launch has at least two workgroups
   for (int64_t it = (int64_t)(blockIdx.x - 1) * kThreads + tid; it < items;
        it += (int64_t)max(1u, gridDim.x - 1) * kThreads) {
      double v[NB];
      if (it < m) {
         const int r = (int)it;
         if (r >= c0 && r < c0 + kb) continue;
#pragma unroll
         for (int j = 0; j < NB; ++j) v[j] = (j < kb) ? F[r + (int64_t)(c0 + pm[j]) * ld] : 0.0;
#pragma unroll
         for (int j = 0; j < NB; ++j)
            if (j < kb) {
               F[r + (int64_t)(c0 + j) * ld] = v[j];
               w.Bk[r + (int64_t)j * m] = v[j];
            }
      } else {
         const int k = (int)(it - m);
         if (k >= c0 && k < c0 + kb) continue;
#pragma unroll
         for (int j = 0; j < NB; ++j) v[j] = (j < kb) ? F[c0 + pm[j] + (int64_t)k * ld] : 0.0;
#pragma unroll
         for (int j = 0; j < NB; ++j)
            if (j < kb) F[c0 + j + (int64_t)k * ld] = v[j];
      }
   }
}

// ApplyN/ApplyT and the threshold test on the uneliminated rows outside the
// block: the failed columns so far, then rows c0 + kb .. m - 1.
static __global__ __launch_bounds__(kThreads) void k_lp_apply(int node, int c0, int par, const IndefDesc* __restrict__ nodes,
                                                       double* __restrict__ arena, double* __restrict__ wsbase,
                                                       double u, double small) {
#pragma clang fp contract(off)
   constexpr int NB = kAppNB;
   __shared__ double T[NB][NB + 1];
   __shared__ double dl[2 * NB];
   __shared__ int s_fail;
   const IndefDesc nd = nodes[node];
   const int m = nd.m, n = nd.n, ld = nd.ld, tid = threadIdx.x;
   const AppWs w = app_ws(wsbase + nd.ws_off, m, n);
   if (w.ST[kStBad]) return;
   const int kb = min(NB, n - c0);
   double* F = arena + nd.off;
   for (int e = tid; e < kb * kb; e += kThreads) {
      const int i = e % kb, j = e / kb;
      T[i][j] = w.TG[i + j * NB];
   }
   for (int k = tid; k < 2 * kb; k += kThreads) dl[k] = (par ? w.DL2 : w.DL)[k];
   if (tid == 0) s_fail = kb;
   __syncthreads();
   const int nf = w.ST[kStNf];
   const double inv_u = 1.0 / u;
   const int64_t items = (int64_t)nf + (m - c0 - kb);
   for (int64_t it = (int64_t)blockIdx.x * kThreads + tid; it < items; it += (int64_t)gridDim.x * kThreads) {
      const int r = it < nf ? w.FL[it] : c0 + kb + (int)(it - nf);
      double x[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) x[j] = (j < kb) ? F[r + (int64_t)(c0 + j) * ld] : 0.0;
      const int fail = apply_row(x, T, dl, kb, small, inv_u);
      if (fail < kb) atomicMin(&s_fail, fail);
#pragma unroll
      for (int j = 0; j < NB; ++j)
         if (j < kb) {
            F[r + (int64_t)(c0 + j) * ld] = x[j];
            if (r < n) F[c0 + j + (int64_t)r * ld] = x[j];
         }
   }
   __syncthreads();
   if (tid == 0 && s_fail < kb) atomicMin(&w.ST[kStFail], s_fail);
}

// Passed count of the block, L11 of the passed columns, their D^{-1} and
// flags; the failed columns of the block join the failed list.
static __global__ __launch_bounds__(kThreads) void k_lp_finish(int node, int c0, int par, const IndefDesc* __restrict__ nodes,
                                                        double* __restrict__ arena, double* __restrict__ dinv,
                                                        double* __restrict__ wsbase) {
   constexpr int NB = kAppNB;
   __shared__ double dl[2 * NB];
   const IndefDesc nd = nodes[node];
   const int m = nd.m, n = nd.n, ld = nd.ld, tid = threadIdx.x;
   const AppWs w = app_ws(wsbase + nd.ws_off, m, n);
   if (w.ST[kStBad]) return;
   const int kb = min(NB, n - c0);
   double* F = arena + nd.off;
   for (int k = tid; k < 2 * kb; k += kThreads) dl[k] = (par ? w.DL2 : w.DL)[k];
   const int nf = w.ST[kStNf], fail = w.ST[kStFail];
   __syncthreads();
   const int npass = npass_adjust(fail, dl);
   for (int e = tid; e < kb * kb; e += kThreads) {
      const int i = e % kb, j = e / kb;
      if (j < npass && i > j) {
         const double t = w.TG[i + j * NB];
         F[c0 + i + (int64_t)(c0 + j) * ld] = t;
         F[c0 + j + (int64_t)(c0 + i) * ld] = t;
      }
   }
   for (int j = tid; j < kb; j += kThreads) {
      if (j < npass) {
         w.EL[c0 + j] = 1;
         dinv[nd.d_off + 2 * (c0 + j)] = dl[2 * j];
         dinv[nd.d_off + 2 * (c0 + j) + 1] = dl[2 * j + 1];
      } else {
         w.FL[nf + (j - npass)] = c0 + j;
      }
   }
   __syncthreads();
   if (tid == 0) {
      w.ST[kStNpass] = npass;
      w.ST[kStNelim] += npass;
      w.ST[kStNfOld] = nf;
      w.ST[kStNf] = nf + kb - npass;
   }
}

constexpr int kLpTile = 128;  // tile of the multi-workgroup fully-summed update

// Entries of k_lp_update by mode: 0 all, 1 only those of the next block's
// diagonal block (row and column index positions in [nf, nf + kb2)), 2 all
// the others. Modes 1 and 2 split one update in two launches so that the
// next block's pivoting can start before the rest is done.
__device__ __forceinline__ bool lp_keep(int mode, int rpos, int spos, int nf, int kb2) {
   if (mode == 0) return true;
   const bool in = rpos >= nf && rpos < nf + kb2 && spos >= nf && spos < nf + kb2;
   return (mode == 1) == in;
}

// Restore of the failed block columns and update of the uneliminated
// fully-summed columns with the block's passed pivots: entry (r, s), r >= s,
// of the index space U = (failed columns, ascending; c0 + kb .. n - 1) for
// columns and U followed by rows n .. m - 1 for rows, gets
// x -= L[r, c0 + p] (D L[s, :])_p for p = 0 .. npass - 1 in that order; the
// fully-summed part is mirrored. Entries of the failed block columns start
// from the backup (the restore). 128 x 128 tiles, 8 x 8 entries per thread,
// tiles taken in turn by a grid of any size.
static __global__ __launch_bounds__(kThreads) void k_lp_update(int node, int c0, int par, int mode,
                                                        const IndefDesc* __restrict__ nodes,
                                                        double* __restrict__ arena, double* __restrict__ wsbase) {
#pragma clang fp contract(off)
   constexpr int NB = kAppNB, TT = kLpTile;
   __shared__ __align__(16) double As[NB][TT + 2];
   __shared__ __align__(16) double Bs[NB][TT + 2];
   __shared__ int ridx[TT], cidx[TT];
   __shared__ double dl[2 * NB];
   const IndefDesc nd = nodes[node];
   const int m = nd.m, n = nd.n, ld = nd.ld, tid = threadIdx.x;
   const AppWs w = app_ws(wsbase + nd.ws_off, m, n);
   if (w.ST[kStBad]) return;
   const int kb = min(NB, n - c0);
   double* F = arena + nd.off;
   const int npass = w.ST[kStNpass], nf = w.ST[kStNf], nf_old = w.ST[kStNfOld];
   const int nu = nf + (n - c0 - kb), nr = nu + (m - n);
   if (nu <= 0) return;
   const int Tr = (nr + TT - 1) / TT, Tc = (nu + TT - 1) / TT;
   const int64_t ntiles = (int64_t)Tc * Tr - (int64_t)Tc * (Tc - 1) / 2;
   // the next block's diagonal block: index positions [nf, nf + kb2)
   const int kb2 = max(0, min(NB, n - c0 - kb));
   const int t0 = nf / TT, t1 = (nf + max(kb2, 1) - 1) / TT;
   if (mode == 1 && (kb2 == 0 || npass == 0)) return;
   const int64_t nt = mode == 1 ? (t1 > t0 ? 3 : 1) : ntiles;
   for (int k = tid; k < 2 * kb; k += kThreads) dl[k] = (par ? w.DL2 : w.DL)[k];
   const int lo = c0 + npass, hi = c0 + kb;  // failed block columns
   const int tr8 = tid & 15, tc8 = tid >> 4;
   for (int64_t t = blockIdx.x; t < nt; t += gridDim.x) {
      int tc = 0, tr = 0;
      if (mode == 1) {
         tr = t == 0 ? t0 : t1;
         tc = t == 2 ? t1 : t0;
      } else {
         int64_t rem = t;
         while (rem >= Tr - tc) {
            rem -= Tr - tc;
            ++tc;
         }
         tr = tc + (int)rem;
      }
      const int R0 = tr * TT, S0 = tc * TT;
      if (npass == 0 && !(R0 < nf && R0 + TT > nf_old) && !(S0 < nf && S0 + TT > nf_old)) continue;
      __syncthreads();
      {
         const int k = (tid < TT ? R0 : S0) + (tid & (TT - 1));
         const int lim = tid < TT ? nr : nu;
         int v = -1;
         if (k < lim) v = k < nf ? w.FL[k] : (k < nu ? c0 + kb + (k - nf) : n + (k - nu));
         if (tid < TT)
            ridx[tid] = v;
         else
            cidx[tid - TT] = v;
      }
      __syncthreads();
      for (int e = tid; e < npass * TT; e += kThreads) {
         const int p = e / TT, i = e % TT;
         const int r = ridx[i], s = cidx[i];
         As[p][i] = r >= 0 ? F[r + (int64_t)(c0 + p) * ld] : 0.0;
         double v = 0.0;
         if (s >= 0) {
            const double* fs = F + s + (int64_t)c0 * ld;
            const double fm = p > 0 ? fs[(int64_t)(p - 1) * ld] : 0.0;
            const double fp = p + 1 < npass ? fs[(int64_t)(p + 1) * ld] : 0.0;
            v = vs_value(dl, p, npass, fm, fs[(int64_t)p * ld], fp);
         }
         Bs[p][i] = v;
      }
      double acc[8][8];
#pragma unroll
      for (int b = 0; b < 8; ++b) {
         const int s = cidx[tc8 * 8 + b];
#pragma unroll
         for (int a = 0; a < 8; ++a) {
            const int r = ridx[tr8 * 8 + a];
            double x = 0.0;
            if (r >= 0 && s >= 0 && r >= s && lp_keep(mode, R0 + tr8 * 8 + a, S0 + tc8 * 8 + b, nf, kb2)) {
               if (s >= lo && s < hi)
                  x = w.Bk[r + (int64_t)(s - c0) * m];
               else if (r >= lo && r < hi)
                  x = w.Bk[s + (int64_t)(r - c0) * m];
               else
                  x = F[r + (int64_t)s * ld];
            }
            acc[a][b] = x;
         }
      }
      __syncthreads();
      for (int p = 0; p < npass; ++p) {
         double av[8], bv[8];
#pragma unroll
         for (int a = 0; a < 8; ++a) av[a] = As[p][tr8 * 8 + a];
#pragma unroll
         for (int b = 0; b < 8; ++b) bv[b] = Bs[p][tc8 * 8 + b];
#pragma unroll
         for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int b = 0; b < 8; ++b) acc[a][b] = fnma(av[a], bv[b], acc[a][b]);
      }
#pragma unroll
      for (int b = 0; b < 8; ++b) {
         const int s = cidx[tc8 * 8 + b];
#pragma unroll
         for (int a = 0; a < 8; ++a) {
            const int r = ridx[tr8 * 8 + a];
            if (r >= 0 && s >= 0 && r >= s && lp_keep(mode, R0 + tr8 * 8 + a, S0 + tc8 * 8 + b, nf, kb2)) {
               F[r + (int64_t)s * ld] = acc[a][b];
               if (r < n && r > s) F[s + (int64_t)r * ld] = acc[a][b];
            }
         }
      }
   }
}

// permute_failed (src/factor_indef.hxx:1190-1331): new order = eliminated
// positions, then the failed ones, both ascending.
static __global__ __launch_bounds__(kThreads) void k_lp_np(int node, const IndefDesc* __restrict__ nodes,
                                                    double* __restrict__ wsbase) {
   const IndefDesc nd = nodes[node];
   const AppWs w = app_ws(wsbase + nd.ws_off, nd.m, nd.n);
   if (w.ST[kStBad] || w.ST[kStNelim] >= nd.n) return;
   if (threadIdx.x == 0) {
      int k = 0;
      for (int p = 0; p < nd.n; ++p)
         if (w.EL[p]) w.NP[k++] = p;
      const int nf = w.ST[kStNf];
      for (int p = 0; p < nf; ++p) w.NP[k++] = w.FL[p];
   }
}

static __global__ __launch_bounds__(kThreads) void k_lp_pgather(int node, const IndefDesc* __restrict__ nodes,
                                                         const double* __restrict__ arena,
                                                         const int* __restrict__ vars, const int* __restrict__ pres,
                                                         const double* __restrict__ dinv,
                                                         double* __restrict__ wsbase) {
   const IndefDesc nd = nodes[node];
   const int m = nd.m, n = nd.n, ld = nd.ld;
   const AppWs w = app_ws(wsbase + nd.ws_off, m, n);
   if (w.ST[kStBad] || w.ST[kStNelim] >= n) return;
   const double* F = arena + nd.off;
   const int64_t gid = (int64_t)blockIdx.x * kThreads + threadIdx.x, G = (int64_t)gridDim.x * kThreads;
   for (int64_t e = gid; e < (int64_t)m * n; e += G) {
      const int i = (int)(e % m), k = (int)(e / m);
      const int oi = (i < n) ? w.NP[i] : i;
      w.PS[i + (int64_t)k * m] = F[oi + (int64_t)w.NP[k] * ld];
   }
   int* ivar = reinterpret_cast<int*>(w.Bk);
   int* ipre = ivar + n;
   double* dd = w.Bk + n;
   for (int64_t k = gid; k < n; k += G) {
      ivar[k] = vars[nd.var_off + w.NP[k]];
      ipre[k] = pres[nd.pre_off + w.NP[k]];
      dd[2 * k] = dinv[nd.d_off + 2 * w.NP[k]];
      dd[2 * k + 1] = dinv[nd.d_off + 2 * w.NP[k] + 1];
   }
}

static __global__ __launch_bounds__(kThreads) void k_lp_pscatter(int node, const IndefDesc* __restrict__ nodes,
                                                          double* __restrict__ arena, int* __restrict__ vars,
                                                          int* __restrict__ pres, double* __restrict__ dinv,
                                                          double* __restrict__ wsbase) {
   const IndefDesc nd = nodes[node];
   const int m = nd.m, n = nd.n, ld = nd.ld;
   const AppWs w = app_ws(wsbase + nd.ws_off, m, n);
   if (w.ST[kStBad] || w.ST[kStNelim] >= n) return;
   double* F = arena + nd.off;
   const int64_t gid = (int64_t)blockIdx.x * kThreads + threadIdx.x, G = (int64_t)gridDim.x * kThreads;
   for (int64_t e = gid; e < (int64_t)m * n; e += G) {
      const int i = (int)(e % m), k = (int)(e / m);
      F[i + (int64_t)k * ld] = w.PS[i + (int64_t)k * m];
   }
   const int* ivar = reinterpret_cast<const int*>(w.Bk);
   const int* ipre = ivar + n;
   const double* dd = w.Bk + n;
   for (int64_t k = gid; k < n; k += G) {
      vars[nd.var_off + k] = ivar[k];
      pres[nd.pre_off + k] = ipre[k];
      dinv[nd.d_off + 2 * k] = dd[2 * k];
      dinv[nd.d_off + 2 * k + 1] = dd[2 * k + 1];
   }
}

// ---------------------------------------------------------------------------
// Second pass of a large front over the whole GPU, one pivot per step:
//   k_tpp_decide  (one workgroup) picks the pivot from per-column statistics
//                 exactly as tpp_pass's candidate loop does, exchanges it to
//                 the front of the remaining columns, forms D^{-1}, l = v D^{-1}
//                 (into TL) and writes l;
//   k_tpp_update  (one workgroup per remaining column) applies the pivot to
//                 its column, all rows, and recomputes the column's
//                 statistics for the next step.
// The fully-summed part is kept symmetric bit for bit (upper = lower), so a
// column's rows above the diagonal are updated from the column's own copies
// with the lower entry's operands, and no workgroup writes another's column.
// ---------------------------------------------------------------------------

// Two best (value, index) pairs, larger value first, smaller index among equals.
struct Top2 {
   double v1, v2;
   int i1, i2;
};
__device__ __forceinline__ bool t2_better(double v, int i, double w, int j) { return v > w || (v == w && i < j); }
__device__ __forceinline__ void t2_add(Top2& a, double v, int i) {
   if (t2_better(v, i, a.v1, a.i1)) {
      a.v2 = a.v1;
      a.i2 = a.i1;
      a.v1 = v;
      a.i1 = i;
   } else if (t2_better(v, i, a.v2, a.i2)) {
      a.v2 = v;
      a.i2 = i;
   }
}

// One column per workgroup: s = nelim + blockIdx.x (nelim after the pivot of
// this step). With kStW = w > 0 the column gets the pivot's update
// x -= l_r v_s (+ l2_r v2_s), rows nelim .. m - 1; then its statistics over
// those rows without the diagonal, and |F[s, s]|.
static __global__ __launch_bounds__(kThreads) void k_tpp_update(int node, const IndefDesc* __restrict__ nodes,
                                                         double* __restrict__ arena, double* __restrict__ wsbase) {
#pragma clang fp contract(off)
   __shared__ double sv1[4], sv2[4];
   __shared__ int si1[4], si2[4];
   const IndefDesc nd = nodes[node];
   const int m = nd.m, n = nd.n, ld = nd.ld, tid = threadIdx.x;
   const AppWs w = app_ws(wsbase + nd.ws_off, m, n);
   if (w.ST[kStDone]) return;
   const int r0 = w.ST[kStNelim], wd = w.ST[kStW];
   const int s = r0 + blockIdx.x;
   if (s >= n) return;
   double* col = arena + nd.off + (int64_t)s * ld;
   const double* lv = w.TL;
   const double* vv = w.TL + 2 * (int64_t)m;
   Top2 t{0.0, 0.0, 0x7fffffff, 0x7fffffff};
   for (int r = r0 + tid; r < m; r += kThreads) {
      double x = col[r];
      if (wd > 0) {
         if (r >= s) {
            x = fnma(lv[2 * r], vv[2 * s], x);
            if (wd == 2) x = fnma(lv[2 * r + 1], vv[2 * s + 1], x);
         } else {
            x = fnma(lv[2 * s], vv[2 * r], x);
            if (wd == 2) x = fnma(lv[2 * s + 1], vv[2 * r + 1], x);
         }
         col[r] = x;
      }
      if (r != s) t2_add(t, fabs(x), r);
   }
   // block reduction of the two best
   const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) {
      const double a1 = __shfl_xor(t.v1, o), a2 = __shfl_xor(t.v2, o);
      const int j1 = __shfl_xor(t.i1, o), j2 = __shfl_xor(t.i2, o);
      t2_add(t, a1, j1);
      t2_add(t, a2, j2);
   }
   if (lane == 0) {
      sv1[wave] = t.v1;
      sv2[wave] = t.v2;
      si1[wave] = t.i1;
      si2[wave] = t.i2;
   }
   __syncthreads();
   if (tid == 0) {
      Top2 a{sv1[0], sv2[0], si1[0], si2[0]};
      for (int k = 1; k < kThreads / 64; ++k) {
         t2_add(a, sv1[k], si1[k]);
         t2_add(a, sv2[k], si2[k]);
      }
      double* ts = w.TS + 4 * (int64_t)s;
      ts[0] = a.v1;
      ts[1] = a.v2;
      ts[2] = (double)a.i1;
      ts[3] = (double)a.i2;
   }
}

// Outcome of candidate p of the second pass from the column statistics
// (the candidate loop of tpp_pass): 0 none, 1 zero pivot with action == 0,
// 2 zero pivot, 3 1x1, 4 2x2 with row *t2.
__device__ inline int tpp_candidate(const double* F, int ld, int n, int p, const double* TS, double u, double small,
                                    int action, int* t2) {
#pragma clang fp contract(off)
   const double* ts = TS + 4 * (int64_t)p;
   const double amax = ts[0];
   int t = (int)ts[2];
   if (!(amax > 0.0)) t = -1;
   const double app = fabs(F[p + (int64_t)p * ld]);
   if (fmax(app, amax) < small) return action ? 2 : 1;
   if (app >= u * amax) return 3;
   if (t >= 0 && t < n) {
      const double a11 = F[p + (int64_t)p * ld], a21 = F[t + (int64_t)p * ld], a22 = F[t + (int64_t)t * ld];
      const double det = __builtin_fma(a11, a22, -(a21 * a21));
      if (det != 0.0) {
         const double d11 = a22 / det, d21 = -a21 / det, d22 = a11 / det;
         const double pm = ts[1];
         const double* tt = TS + 4 * (int64_t)t;
         const double tm = ((int)tt[2] == p) ? tt[1] : tt[0];
         const double g = fmax(dot2(fabs(d11), pm, fabs(d21), tm), dot2(fabs(d21), pm, fabs(d22), tm));
         if (u * g <= 1.0) {
            *t2 = t;
            return 4;
         }
      }
   }
   return 0;
}

static __global__ __launch_bounds__(kThreads) void k_tpp_decide(int node, const IndefDesc* __restrict__ nodes,
                                                         double* __restrict__ arena, int* __restrict__ vars,
                                                         int* __restrict__ pres, double* __restrict__ dinv,
                                                         double* __restrict__ wsbase, double u, double small,
                                                         int action) {
#pragma clang fp contract(off)
   __shared__ int s_best;
   const IndefDesc nd = nodes[node];
   const int m = nd.m, n = nd.n, ld = nd.ld, tid = threadIdx.x;
   const AppWs w = app_ws(wsbase + nd.ws_off, m, n);
   if (w.ST[kStDone]) return;
   const int nelim = w.ST[kStNelim];
   double* F = arena + nd.off;
   if (tid == 0) s_best = 0x7fffffff;
   __syncthreads();
   for (int p = nelim + tid; p < n; p += kThreads) {
      int t2 = -1;
      if (tpp_candidate(F, ld, n, p, w.TS, u, small, action, &t2)) {
         atomicMin(&s_best, p);
         break;
      }
   }
   __syncthreads();
   const int piv = s_best;
   if (piv == 0x7fffffff) {  // no pivot: the rest is delayed
      if (tid == 0) w.ST[kStDone] = 1;
      return;
   }
   int t = -1;
   const int code = tpp_candidate(F, ld, n, piv, w.TS, u, small, action, &t);
   if (code == 1) {
      if (tid == 0) {
         w.ST[kStTppErr] = 1;
         w.ST[kStDone] = 1;
      }
      return;
   }
   int* var = vars + nd.var_off;
   int* pre = pres + nd.pre_off;
   double* D = dinv + nd.d_off;
   front_swap(F, ld, m, n, piv, nelim, var, pre);
   if (code == 2) {  // zero pivot
      for (int i = nelim + 1 + tid; i < m; i += kThreads) {
         F[i + (int64_t)nelim * ld] = 0.0;
         if (i < n) F[nelim + (int64_t)i * ld] = 0.0;
      }
      if (tid == 0) {
         D[2 * nelim] = 0.0;
         D[2 * nelim + 1] = 0.0;
         w.ST[kStNelim] = nelim + 1;
         w.ST[kStW] = 0;
      }
      return;
   }
   const int wd = code == 4 ? 2 : 1;
   double d[4];
   if (wd == 2) {
      int tt = t;
      if (tt == nelim) tt = piv;
      front_swap(F, ld, m, n, tt, nelim + 1, var, pre);
      const double b11 = F[nelim + (int64_t)nelim * ld], b21 = F[nelim + 1 + (int64_t)nelim * ld],
                   b22 = F[nelim + 1 + (int64_t)(nelim + 1) * ld];
      const double dt = __builtin_fma(b11, b22, -(b21 * b21));
      d[0] = b22 / dt;
      d[1] = -b21 / dt;
      d[2] = __builtin_inf();
      d[3] = b11 / dt;
   } else {
      d[0] = 1.0 / F[nelim + (int64_t)nelim * ld];
      d[1] = 0.0;
      d[2] = d[3] = 0.0;
   }
   double* lv = w.TL;
   double* vv = w.TL + 2 * (int64_t)m;
   const int r0 = nelim + wd;
   for (int r = r0 + tid; r < m; r += kThreads) {
      const double a1 = F[r + (int64_t)nelim * ld];
      const double a2 = (wd == 2) ? F[r + (int64_t)(nelim + 1) * ld] : 0.0;
      vv[2 * r] = a1;
      vv[2 * r + 1] = a2;
      double l1, l2;
      if (wd == 2) {
         l1 = dot2(d[0], a1, d[1], a2);
         l2 = dot2(d[1], a1, d[3], a2);
      } else {
         l1 = a1;
         apply_dinv1(d[0], small, l1);
         l2 = 0.0;
      }
      lv[2 * r] = l1;
      lv[2 * r + 1] = l2;
      F[r + (int64_t)nelim * ld] = l1;
      if (r < n) F[nelim + (int64_t)r * ld] = l1;
      if (wd == 2) {
         F[r + (int64_t)(nelim + 1) * ld] = l2;
         if (r < n) F[nelim + 1 + (int64_t)r * ld] = l2;
      }
   }
   if (tid == 0) {
      D[2 * nelim] = d[0];
      D[2 * nelim + 1] = d[1];
      if (wd == 2) {
         D[2 * nelim + 2] = d[2];
         D[2 * nelim + 3] = d[3];
         F[nelim + 1 + (int64_t)nelim * ld] = 0.0;
         F[nelim + (int64_t)(nelim + 1) * ld] = 0.0;
      }
      w.ST[kStNelim] = r0;
      w.ST[kStW] = wd;
   }
}


// Start of the second pass (kStFirst = columns eliminated by the first pass;
// `allowed` = m == n or the failed-pivot method is TPP).
static __global__ void k_tpp_start(int node, const IndefDesc* __restrict__ nodes, double* __restrict__ wsbase,
                                   int allowed) {
   const IndefDesc nd = nodes[node];
   const AppWs w = app_ws(wsbase + nd.ws_off, nd.m, nd.n);
   if (threadIdx.x == 0) {
      w.ST[kStFirst] = w.ST[kStNelim];
      w.ST[kStW] = 0;
      w.ST[kStTppErr] = 0;
      w.ST[kStDone] = (!allowed || w.ST[kStBad] || w.ST[kStNelim] >= nd.n) ? 1 : 0;
   }
}

// Results of a front of the multi-workgroup panel; out slot `slot` as
// k_app_front's.
static __global__ void k_lp_tail(int node, int slot, IndefDesc* __restrict__ nodes, double* __restrict__ wsbase,
                                 int* __restrict__ status, int* __restrict__ out) {
   const IndefDesc nd = nodes[node];
   const AppWs w = app_ws(wsbase + nd.ws_off, nd.m, nd.n);
   if (threadIdx.x != 0) return;
   if (w.ST[kStBad]) {
      status[node] = -5;
      out[2 * slot] = 0;
      out[2 * slot + 1] = 0;
      return;
   }
   const int nelim = w.ST[kStNelim], nelim_first = w.ST[kStFirst];
   out[2 * slot] = nelim;
   out[2 * slot + 1] = nelim_first;
   if (w.ST[kStTppErr]) {
      status[node] = -5;
      return;
   }
   nodes[node].nelim = nelim;
   nodes[node].nelim_first = nelim_first;
}

// ---------------------------------------------------------------------------
// Assembly of the indefinite fronts of one level.
// ---------------------------------------------------------------------------
// Scatter of A (init_a_block, src/kernels/assemble.hxx:161-214, rows at or
// beyond ncol shifted by ndelay_in) and the row lists: own pivots and
// remaining rows from rlist, pre-pivot positions 0..n-1.
static __global__ __launch_bounds__(kThreads) void k_ind_init(const int* __restrict__ list,
                                                       const IndefDesc* __restrict__ nodes,
                                                       const int64_t* __restrict__ nlist,
                                                       const int* __restrict__ rlist,
                                                       const double* __restrict__ aval,
                                                       const double* __restrict__ scaling,
                                                       double* __restrict__ arena,
                                                       int* __restrict__ vars, int* __restrict__ pres) {
   const IndefDesc nd = nodes[list[blockIdx.x]];
   double* F = arena + nd.off;
   const int* rl = rlist + nd.rlist_off;
   for (int64_t p = threadIdx.x; p < nd.num_a; p += kThreads) {
      const int64_t src = nlist[2 * (nd.amap_off + p)] - 1;
      const int64_t dest = nlist[2 * (nd.amap_off + p) + 1] - 1;
      const int c = (int)(dest / nd.nrow);
      const int r0 = (int)(dest % nd.nrow);
      const int r = r0 >= nd.ncol ? r0 + nd.ndelay_in : r0;
      double v = aval[src];
      // with a scaling: rscale * a * cscale, rows and columns by their
      // symbolic row-list entries (init_a_block, src/kernels/assemble.hxx:166-177)
      if (scaling) v = scaling[rl[r0] - 1] * v * scaling[rl[c] - 1];
      F[r + (int64_t)c * nd.ld] = v;
   }
   int* var = vars + nd.var_off;
   for (int r = threadIdx.x; r < nd.m; r += kThreads) {
      if (r < nd.ncol)
         var[r] = rl[r];
      else if (r >= nd.n)
         var[r] = rl[nd.ncol + (r - nd.n)];
   }
   for (int r = threadIdx.x; r < nd.n; r += kThreads) pres[nd.pre_off + r] = r;
}

// Delayed columns of each child into its parent (assemble_delays,
// src/kernels/assemble.hxx:926-961): the lower triangle of the delayed block
// and the child's contribution rows of those columns, placed through the
// child map; the parent's row list gets the delayed variables. Children of
// one level write disjoint parent entries; grid (column chunks, children).
static __global__ __launch_bounds__(kThreads) void k_ind_delays(const int* __restrict__ children,
                                                         const IndefDesc* __restrict__ nodes,
                                                         const int* __restrict__ cmap,
                                                         double* __restrict__ arena,
                                                         int* __restrict__ vars, const double* __restrict__ cbs) {
   const IndefDesc cn = nodes[children[blockIdx.y]];
   const IndefDesc pn = nodes[cn.parent];
   const int nd_out = cn.n - cn.nelim;
   const double* C = child_cb(cn, cbs);  // front entry (nelim + i, nelim + j) at C[i + j * ldc]
   const int64_t ldc = cn.m - cn.nelim;
   double* P = arena + pn.off;
   const int* cm = cmap + cn.cmap_off;
   const int q = cn.m - cn.n;
   for (int i = blockIdx.x; i < nd_out; i += gridDim.x) {  // one delayed column per workgroup
      const int dc = cn.delay_col0 + i;
      for (int j = threadIdx.x; j < nd_out - i; j += kThreads)
         P[dc + j + (int64_t)dc * pn.ld] = C[i + j + (int64_t)i * ldc];
      for (int k = threadIdx.x; k < q; k += kThreads) {
         int r = cm[k];
         if (r >= pn.ncol) r += pn.ndelay_in;
         const double v = C[nd_out + k + (int64_t)i * ldc];
         if (r < pn.ncol)
            P[dc + (int64_t)r * pn.ld] = v;
         else
            P[r + (int64_t)dc * pn.ld] = v;
      }
   }
   if (blockIdx.x == 0)
      for (int i = threadIdx.x; i < nd_out; i += kThreads)
         vars[pn.var_off + cn.delay_col0 + i] = vars[cn.var_off + cn.nelim + i];
}

// Extend-add of the contribution blocks of the children of one rank
// (src/kernels/assemble.hxx:230-517 with the delay shift of the parent's
// map, src/assemble.hxx:184-190): one workgroup per (child, column chunk).
static __global__ __launch_bounds__(kThreads) void k_ind_ea(const int* __restrict__ children,
                                                     const IndefDesc* __restrict__ nodes,
                                                     const int* __restrict__ cmap,
                                                     double* __restrict__ arena, const double* __restrict__ cbs) {
   const IndefDesc cn = nodes[children[blockIdx.y]];
   const IndefDesc pn = nodes[cn.parent];
   const int q = cn.m - cn.n, nd_out = cn.n - cn.nelim;
   const int64_t ldc = cn.m - cn.nelim;
   const double* C = child_cb(cn, cbs) + nd_out + nd_out * ldc;  // contribution block, ld ldc
   double* P = arena + pn.off;
   const int* cm = cmap + cn.cmap_off;
   for (int j = blockIdx.x; j < q; j += gridDim.x) {
      int pj = cm[j];
      if (pj >= pn.ncol) pj += pn.ndelay_in;
      for (int i = j + threadIdx.x; i < q; i += kThreads) {
         int pi = cm[i];
         if (pi >= pn.ncol) pi += pn.ndelay_in;
         P[pi + (int64_t)pj * pn.ld] += C[i + (int64_t)j * ldc];
      }
   }
}

// LD = L[n:m, 0:nelim] D for the contribution-block update (calcLD), into the
// node's workspace (q x nelim, ld q).
static __global__ __launch_bounds__(kThreads) void k_ind_ld(const int* __restrict__ list,
                                                     const IndefDesc* __restrict__ nodes,
                                                     const double* __restrict__ arena,
                                                     const double* __restrict__ dinv,
                                                     double* __restrict__ wsbase) {
#pragma clang fp contract(off)
   const IndefDesc nd = nodes[list[blockIdx.y]];
   const int q = nd.m - nd.n, ne = nd.nelim;
   const double* F = arena + nd.off;
   const double* d = dinv + nd.d_off;
   double* W = wsbase + nd.ws_off;
   for (int j = blockIdx.x; j < ne; j += gridDim.x) {
      const bool second = j > 0 && __builtin_isinf(d[2 * j]);
      const bool first = j + 1 < ne && __builtin_isinf(d[2 * j + 2]);
      for (int i = threadIdx.x; i < q; i += kThreads) {
         double v;
         if (first) {
            const double det = __builtin_fma(d[2 * j], d[2 * j + 3], -(d[2 * j + 1] * d[2 * j + 1]));
            v = dot2(d[2 * j + 3] / det, F[nd.n + i + (int64_t)j * nd.ld], -d[2 * j + 1] / det,
                     F[nd.n + i + (int64_t)(j + 1) * nd.ld]);
         } else if (second) {
            const int k = j - 1;
            const double det = __builtin_fma(d[2 * k], d[2 * k + 3], -(d[2 * k + 1] * d[2 * k + 1]));
            v = dot2(-d[2 * k + 1] / det, F[nd.n + i + (int64_t)k * nd.ld], d[2 * k] / det,
                     F[nd.n + i + (int64_t)j * nd.ld]);
         } else {
            const double D0 = (d[2 * j] == 0.0) ? 0.0 : 1.0 / d[2 * j];
            v = D0 * F[nd.n + i + (int64_t)j * nd.ld];
         }
         W[i + (int64_t)j * q] = v;
      }
   }
}

// FP32 tier: LD as above and a copy of L21 = L[n:m, 0:nelim], both rounded
// to FP32, into the node's workspace viewed as floats from float offset
// ind32_base(ws_off): LD at [0, ldq nelim), L21 at [ldq nelim, 2 ldq nelim),
// ldq = q rounded up to 4 (16-byte aligned columns for the update's loads).
// The D^{-1} products are formed in FP64 and rounded once.
__host__ __device__ inline int64_t ind32_base(int64_t ws_off) { return (2 * ws_off + 3) / 4 * 4; }
__host__ __device__ inline int ind32_ldq(int q) { return (q + 3) / 4 * 4; }

static __global__ __launch_bounds__(kThreads) void k_ind_ld32(const int* __restrict__ list,
                                                       const IndefDesc* __restrict__ nodes,
                                                       const double* __restrict__ arena,
                                                       const double* __restrict__ dinv,
                                                       double* __restrict__ wsbase) {
#pragma clang fp contract(off)
   const IndefDesc nd = nodes[list[blockIdx.y]];
   const int q = nd.m - nd.n, ne = nd.nelim, ldq = ind32_ldq(q);
   const double* F = arena + nd.off;
   const double* d = dinv + nd.d_off;
   float* W = reinterpret_cast<float*>(wsbase) + ind32_base(nd.ws_off);
   float* L21 = W + (int64_t)ldq * ne;
   for (int j = blockIdx.x; j < ne; j += gridDim.x) {
      const bool second = j > 0 && __builtin_isinf(d[2 * j]);
      const bool first = j + 1 < ne && __builtin_isinf(d[2 * j + 2]);
      for (int i = threadIdx.x; i < ldq; i += kThreads) {
         if (i >= q) {  // padding rows: zeros, never read into C
            W[i + (int64_t)j * ldq] = 0.0f;
            L21[i + (int64_t)j * ldq] = 0.0f;
            continue;
         }
         const double lij = F[nd.n + i + (int64_t)j * nd.ld];
         double v;
         if (first) {
            const double det = __builtin_fma(d[2 * j], d[2 * j + 3], -(d[2 * j + 1] * d[2 * j + 1]));
            v = dot2(d[2 * j + 3] / det, lij, -d[2 * j + 1] / det, F[nd.n + i + (int64_t)(j + 1) * nd.ld]);
         } else if (second) {
            const int k = j - 1;
            const double det = __builtin_fma(d[2 * k], d[2 * k + 3], -(d[2 * k + 1] * d[2 * k + 1]));
            v = dot2(-d[2 * k + 1] / det, F[nd.n + i + (int64_t)k * nd.ld], d[2 * k] / det, lij);
         } else {
            const double D0 = (d[2 * j] == 0.0) ? 0.0 : 1.0 / d[2 * j];
            v = D0 * lij;
         }
         W[i + (int64_t)j * ldq] = (float)v;
         L21[i + (int64_t)j * ldq] = (float)lij;
      }
   }
}

// After a level: L of every front (m x nelim) into the factor store, and the
// trailing block (lower triangle of rows/cols nelim..m-1) into the
// contribution store for the parent's assembly. Grid (chunks, fronts).
static __global__ __launch_bounds__(kThreads) void k_ind_store(const int* __restrict__ list,
                                                        const IndefDesc* __restrict__ nodes,
                                                        const double* __restrict__ arena,
                                                        double* __restrict__ lstore, double* __restrict__ cbs,
                                                        int nnodes) {
   const IndefDesc nd = nodes[list[blockIdx.y]];
   const int m = nd.m, ne = nd.nelim;
   const double* F = arena + nd.off;
   const int64_t G = (int64_t)gridDim.x * kThreads, gid = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   double* L = lstore + nd.lf_off;
   for (int64_t e = gid; e < (int64_t)m * ne; e += G) {
      const int i = (int)(e % m), j = (int)(e / m);
      L[i + (int64_t)j * m] = F[i + (int64_t)j * nd.ld];
   }
   const int t = m - ne;
   if (nd.parent >= nnodes || t <= 0) return;
   double* C = nd.cb_big ? nd.cb_big : cbs + nd.cb_off;
   for (int64_t e = gid; e < (int64_t)t * t; e += G) {
      const int i = (int)(e % t), j = (int)(e / t);
      if (i >= j) C[i + (int64_t)j * t] = F[ne + i + (int64_t)(ne + j) * nd.ld];
   }
}

constexpr int kIsvMin = 256;  // fronts with this many eliminated columns get the blocked solve
// ... and so do fronts whose L has this many entries: one workgroup sweeping
// a tall L column by column waits on a memory round trip per column.
constexpr int64_t kIsvMinEntries = int64_t(1) << 16;
__host__ __device__ inline bool isv_blocked(int m, int nelim) {
   return nelim > 0 && (nelim >= kIsvMin || (int64_t)m * nelim >= kIsvMinEntries);
}

// ---------------------------------------------------------------------------
// Solves (NumericTree::solve_fwd / solve_diag_bwd_inner,
// src/NumericTree.hxx:408-532; ldlt_app_solve_fwd/diag/bwd,
// src/kernels/ldlt_app.cxx:53-115). One workgroup per node. Per right-hand
// side the node owns two m-vectors at x_off: the work vector in front order
// and a pre-pivot gather buffer.
// ---------------------------------------------------------------------------
// Forward: gather (own rows from b, delayed rows copied from the child that
// delayed them, contribution rows of every child added in child order),
// permute to the pivoted order, unit lower solve on [0, nelim), update of the
// rows below; the eliminated entries go to `sol`.
static __global__ __launch_bounds__(kThreads) void k_ind_fwd(const int* __restrict__ list,
                                                      const IndefDesc* __restrict__ nodes,
                                                      const double* __restrict__ lstore,
                                                      const int* __restrict__ vars,
                                                      const int* __restrict__ pres,
                                                      const int* __restrict__ cmap,
                                                      const double* __restrict__ b, int ldb,
                                                      double* __restrict__ xw,
                                                      double* __restrict__ sol, int nrhs) {
   const IndefDesc nd = nodes[list[blockIdx.x]];
   const int m = nd.m, n = nd.n, ne = nd.nelim;
   const double* F = lstore + nd.lf_off;  // L, ld m
   const int* var = vars + nd.var_off;
   const int* pre = pres + nd.pre_off;
   for (int c = 0; c < nrhs; ++c) {
      double* xl = xw + nd.x_off * nrhs + (int64_t)c * 2 * m;
      double* xp = xl + m;
      for (int r = threadIdx.x; r < m; r += kThreads) xp[r] = 0.0;
      __syncthreads();
      // own pivots: b at the symbolic row list (pre-pivot positions < ncol)
      for (int r = threadIdx.x; r < n; r += kThreads) {
         const int p = pre[r];
         if (p < nd.ncol) xp[p] = b[var[r] - 1 + (int64_t)c * ldb];
      }
      __syncthreads();
      for (int ch = nd.first_child; ch != -1;) {
         const IndefDesc cn = nodes[ch];
         const double* cx = xw + cn.x_off * nrhs + (int64_t)c * 2 * cn.m;
         const int nd_out = cn.n - cn.nelim;
         for (int i = threadIdx.x; i < nd_out; i += kThreads) xp[cn.delay_col0 + i] = cx[cn.nelim + i];
         const int* cm = cmap + cn.cmap_off;
         for (int k = threadIdx.x; k < cn.m - cn.n; k += kThreads) {
            int r = cm[k];
            if (r >= nd.ncol) r += nd.ndelay_in;
            xp[r] += cx[cn.n + k];
         }
         __syncthreads();
         ch = cn.next_child;
      }
      for (int r = threadIdx.x; r < m; r += kThreads) xl[r] = (r < n) ? xp[pre[r]] : xp[r];
      __syncthreads();
      if (isv_blocked(m, ne)) continue;  // blocked elimination by k_isv_* (the loop's barriers stay uniform)
      for (int j = 0; j < ne; ++j) {
         const double y = xl[j];
         for (int r = j + 1 + threadIdx.x; r < m; r += kThreads)
            xl[r] = fnma(F[r + (int64_t)j * m], y, xl[r]);
         __syncthreads();
      }
      for (int r = threadIdx.x; r < ne; r += kThreads) sol[var[r] - 1 + (int64_t)c * ldb] = xl[r];
      __syncthreads();
   }
}

// Diagonal: sol at the eliminated variables of every node times D^{-1}.
static __global__ __launch_bounds__(kThreads) void k_ind_diag(const IndefDesc* __restrict__ nodes,
                                                       int nnodes, const int* __restrict__ vars,
                                                       const double* __restrict__ dinv,
                                                       double* __restrict__ sol, int ld, int nrhs) {
#pragma clang fp contract(off)
   const int node = blockIdx.x;
   if (node >= nnodes) return;
   const IndefDesc nd = nodes[node];
   const int* var = vars + nd.var_off;
   const double* d = dinv + nd.d_off;
   for (int64_t e = threadIdx.x; e < (int64_t)nd.nelim * nrhs; e += kThreads) {
      const int i = (int)(e % nd.nelim), c = (int)(e / nd.nelim);
      double* x = sol + (int64_t)c * ld;
      if (i > 0 && __builtin_isinf(d[2 * i])) continue;  // second half: done by the first
      if (i + 1 < nd.nelim && __builtin_isinf(d[2 * i + 2])) {
         const double a = x[var[i] - 1], bb = x[var[i + 1] - 1];
         x[var[i] - 1] = dot2(d[2 * i], a, d[2 * i + 1], bb);
         x[var[i + 1] - 1] = dot2(d[2 * i + 1], a, d[2 * i + 3], bb);
      } else {
         x[var[i] - 1] = x[var[i] - 1] * d[2 * i];
      }
   }
}

// Backward: x_k = L11^{-T} (x_k - L21^T x_rest) with x_rest read from `sol`
// (ancestors are done, levels run top-down).
static __global__ __launch_bounds__(kThreads) void k_ind_bwd(const int* __restrict__ list,
                                                      const IndefDesc* __restrict__ nodes,
                                                      const double* __restrict__ lstore,
                                                      const int* __restrict__ vars,
                                                      double* __restrict__ xw,
                                                      double* __restrict__ sol, int ld, int nrhs) {
#pragma clang fp contract(off)
   __shared__ double red[kThreads / 64];
   const IndefDesc nd = nodes[list[blockIdx.x]];
   const int m = nd.m, ne = nd.nelim;
   const double* F = lstore + nd.lf_off;  // L, ld m
   const int* var = vars + nd.var_off;
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   for (int c = 0; c < nrhs; ++c) {
      double* xl = xw + nd.x_off * nrhs + (int64_t)c * 2 * m;
      const double* x = sol + (int64_t)c * ld;
      for (int r = threadIdx.x; r < m; r += kThreads) xl[r] = x[var[r] - 1];
      __syncthreads();
      if (isv_blocked(m, ne)) continue;  // blocked elimination by k_isv_*
      for (int j = ne - 1; j >= 0; --j) {
         double s = 0.0;
         for (int r = j + 1 + threadIdx.x; r < m; r += kThreads) s = __builtin_fma(F[r + (int64_t)j * m], xl[r], s);
#pragma unroll
         for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
         if (lane == 0) red[wave] = s;
         __syncthreads();
         if (threadIdx.x == 0) xl[j] = xl[j] - (red[0] + red[1] + red[2] + red[3]);
         __syncthreads();
      }
      for (int r = threadIdx.x; r < ne; r += kThreads) sol[var[r] - 1 + (int64_t)c * ld] = xl[r];
      __syncthreads();
   }
}


// ---------------------------------------------------------------------------
// Blocked solves of a front with many eliminated columns or a large L (isv_blocked),
// 64 columns per step: the per-node kernels above gather (forward) or load
// (backward) the work vector xl, these do the elimination, k_isv_out writes
// the eliminated variables. Forward: the same fnma chain per row, in column
// order, as k_ind_fwd; backward: block sums in a fixed order.
// ---------------------------------------------------------------------------
// Forward, triangle of block j0 (one workgroup; one wave per right-hand side).
static __global__ __launch_bounds__(kThreads) void k_isv_fdiag(int node, int j0, const IndefDesc* __restrict__ nodes,
                                                        const double* __restrict__ lstore,
                                                        double* __restrict__ xw, int nrhs) {
#pragma clang fp contract(off)
   const IndefDesc nd = nodes[node];
   const int m = nd.m, jb = min(64, nd.nelim - j0);
   const double* L = lstore + nd.lf_off;
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   for (int c = wave; c < nrhs; c += kThreads / 64) {
      double* xl = xw + nd.x_off * nrhs + (int64_t)c * 2 * m;
      double x = lane < jb ? xl[j0 + lane] : 0.0;
      for (int j = 0; j < jb; ++j) {
         const double y = __shfl(x, j);
         const double l = (lane > j && lane < jb) ? L[j0 + lane + (int64_t)(j0 + j) * m] : 0.0;
         if (lane > j && lane < jb) x = fnma(l, y, x);
      }
      if (lane < jb) xl[j0 + lane] = x;
   }
}

// Forward, rows below block j0: x_r -= L[r, j] y_j, j in the block, in order.
static __global__ __launch_bounds__(kThreads) void k_isv_fupd(int node, int j0, const IndefDesc* __restrict__ nodes,
                                                       const double* __restrict__ lstore,
                                                       double* __restrict__ xw, int nrhs) {
#pragma clang fp contract(off)
   __shared__ double ys[64];
   const IndefDesc nd = nodes[node];
   const int m = nd.m, jb = min(64, nd.nelim - j0);
   const double* L = lstore + nd.lf_off;
   const int r = j0 + jb + blockIdx.x * kThreads + threadIdx.x;
   for (int c = 0; c < nrhs; ++c) {
      double* xl = xw + nd.x_off * nrhs + (int64_t)c * 2 * m;
      __syncthreads();
      if (threadIdx.x < jb) ys[threadIdx.x] = xl[j0 + threadIdx.x];
      __syncthreads();
      if (r >= m) continue;
      double x = xl[r];
      for (int j = 0; j < jb; ++j) x = fnma(L[r + (int64_t)(j0 + j) * m], ys[j], x);
      xl[r] = x;
   }
}

// Backward, block j0: partial sums t_j = sum L[r, j] x_r over rows r below
// the block, 256 rows per workgroup; part[(chunk * nrhs + c) * 64 + j].
static __global__ __launch_bounds__(kThreads) void k_isv_bdot(int node, int j0, const IndefDesc* __restrict__ nodes,
                                                       const double* __restrict__ lstore,
                                                       const double* __restrict__ xw, double* __restrict__ part,
                                                       int nrhs) {
   __shared__ double ws[4][64];
   const IndefDesc nd = nodes[node];
   const int m = nd.m, jb = min(64, nd.nelim - j0);
   const double* L = lstore + nd.lf_off;
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const int r = j0 + jb + blockIdx.x * kThreads + threadIdx.x;
   const bool ok = r < m;
   double lr[64];
#pragma unroll
   for (int j = 0; j < 64; ++j) lr[j] = (ok && j < jb) ? L[r + (int64_t)(j0 + j) * m] : 0.0;
   for (int c = 0; c < nrhs; ++c) {
      const double* xl = xw + nd.x_off * nrhs + (int64_t)c * 2 * m;
      const double xr = ok ? xl[r] : 0.0;
      double v[64];
#pragma unroll
      for (int j = 0; j < 64; ++j) v[j] = lr[j] * xr;
      // transpose-reduce over the wave: lane l ends with column l's sum
#pragma unroll
      for (int o = 32, h = 32; o > 0; o >>= 1, h >>= 1) {
         const bool up = lane & o;
#pragma unroll
         for (int i = 0; i < h; ++i) {
            const double send = up ? v[i] : v[i + h];
            const double recv = __shfl_xor(send, o);
            v[i] = (up ? v[i + h] : v[i]) + recv;
         }
      }
      __syncthreads();
      ws[wave][lane] = v[0];
      __syncthreads();
      if (threadIdx.x < 64)
         part[((int64_t)blockIdx.x * nrhs + c) * 64 + threadIdx.x] =
             ((ws[0][threadIdx.x] + ws[1][threadIdx.x]) + ws[2][threadIdx.x]) + ws[3][threadIdx.x];
   }
}

// Backward, triangle of block j0 (one workgroup; one wave per right-hand
// side): x_j -= t_j + sum over block rows r > j of L[r, j] x_r, descending j.
static __global__ __launch_bounds__(kThreads) void k_isv_bdiag(int node, int j0, int nchunks,
                                                        const IndefDesc* __restrict__ nodes,
                                                        const double* __restrict__ lstore,
                                                        double* __restrict__ xw, const double* __restrict__ part,
                                                        int nrhs) {
   __shared__ double Lb[64][65];
   const IndefDesc nd = nodes[node];
   const int m = nd.m, jb = min(64, nd.nelim - j0);
   const double* L = lstore + nd.lf_off;
   for (int e = threadIdx.x; e < 64 * 64; e += kThreads) {
      const int i = e & 63, j = e >> 6;
      Lb[i][j] = (i < jb && j < jb) ? L[j0 + i + (int64_t)(j0 + j) * m] : 0.0;
   }
   __syncthreads();
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   for (int c = wave; c < nrhs; c += kThreads / 64) {
      double* xl = xw + nd.x_off * nrhs + (int64_t)c * 2 * m;
      double t = 0.0;
      for (int k = 0; k < nchunks; ++k) t += part[((int64_t)k * nrhs + c) * 64 + lane];
      double x = lane < jb ? xl[j0 + lane] - t : 0.0;
      double acc = 0.0;  // sum over finished rows r > lane of L[r, lane] x_r
      for (int j = jb - 1; j >= 0; --j) {
         if (lane == j) x = x - acc;
         const double xj = __shfl(x, j);
         if (lane < j) acc = __builtin_fma(Lb[j][lane], xj, acc);
      }
      if (lane < jb) xl[j0 + lane] = x;
   }
}

// Eliminated variables of a blocked front into the solution.
static __global__ __launch_bounds__(kThreads) void k_isv_out(int node, const IndefDesc* __restrict__ nodes,
                                                      const int* __restrict__ vars, const double* __restrict__ xw,
                                                      double* __restrict__ sol, int ld, int nrhs) {
   const IndefDesc nd = nodes[node];
   const int* var = vars + nd.var_off;
   for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < (int64_t)nd.nelim * nrhs;
        e += (int64_t)gridDim.x * kThreads) {
      const int r = (int)(e % nd.nelim), c = (int)(e / nd.nelim);
      sol[var[r] - 1 + (int64_t)c * ld] = xw[nd.x_off * nrhs + (int64_t)c * 2 * nd.m + r];
   }
}

}  // namespace sylver_amd
