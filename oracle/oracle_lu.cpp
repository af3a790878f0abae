// CPU restatement of the reference's dense front LU kernels (test
// infrastructure: only tests/ load it).
//
//   lu_nopiv_factor  src/kernels/lu_nopiv.hxx:13-41
//   lu_pp_factor     src/kernels/lu_pp.hxx:38-73 (permute_rows :12-34)
//   lu_tpp_factor    src/kernels/lu_tpp.hxx:33-118 (swap_rc :11-31)
//   find_col_abs_max include/sylver/kernels/common.hxx:63-104 (first index
//                    of the largest |a|, ties keep the earlier row)
//
// Front: m x m column-major (lda >= m); the first k columns are eliminated
// (k <= m), the trailing (m - k) x (m - k) block is left as the Schur
// complement. Each elimination step at pivot e: d = 1 / a_ee, l_ie = a_ie d
// for i > e, then a_ij -= l_ie a_ej for i, j > e over the whole trailing
// square (the reference's rank-1 host_gemm). The product and subtraction are
// written as one fused multiply-add, fma(-l, u, a) -- the rounding the device
// kernel uses; the reference's BLAS rounding is not pinned (its tests check
// the backward error only). Built with contraction off so these roundings
// are exactly the ones written.
#include <cmath>
#include <cstdint>
#include <utility>

namespace {

template <typename T>
void eliminate(int m, int e, T* a, int lda) {
   const T d = T(1) / a[e * lda + e];
   for (int i = e + 1; i < m; ++i) a[e * lda + i] = a[e * lda + i] * d;
   for (int j = e + 1; j < m; ++j) {
      const T u = a[j * lda + e];
      for (int i = e + 1; i < m; ++i) a[j * lda + i] = std::fma(-a[e * lda + i], u, a[j * lda + i]);
   }
}

template <typename T>
void col_abs_max(int from, int to, const T* col, int& idx, T& val) {
   idx = from;
   val = std::fabs(col[from]);
   for (int i = from; i <= to; ++i)
      if (std::fabs(col[i]) > val) {
         val = std::fabs(col[i]);
         idx = i;
      }
}

template <typename T>
void swap_rows(int r1, int r2, int m, T* a, int lda) {
   if (r1 == r2) return;
   for (int c = 0; c < m; ++c) std::swap(a[c * lda + r1], a[c * lda + r2]);
}

template <typename T>
void swap_cols(int c1, int c2, int m, T* a, int lda) {
   if (c1 == c2) return;
   for (int r = 0; r < m; ++r) std::swap(a[c1 * lda + r], a[c2 * lda + r]);
}

// mode 0: no pivoting, 1: partial pivoting, 2: threshold partial pivoting
// with column interchanges among the first k columns (u the threshold).
// rperm / cperm (m entries) start as the identity. Returns nelim.
template <typename T>
int lu_front(int mode, double u, int m, int k, T* a, int lda, int* rperm, int* cperm) {
   for (int i = 0; i < m; ++i) rperm[i] = cperm[i] = i;
   if (mode == 0) {
      for (int e = 0; e < k; ++e) eliminate(m, e, a, lda);
      return k;
   }
   if (mode == 1) {
      for (int e = 0; e < k; ++e) {
         int p;
         T v;
         col_abs_max(e, m - 1, &a[e * lda], p, v);
         if (p != e) {
            std::swap(rperm[e], rperm[p]);
            swap_rows(e, p, m, a, lda);
         }
         eliminate(m, e, a, lda);
      }
      return k;
   }
   int nelim = 0, c = 0;
   while (c < k) {
      int idx;
      T maxc;
      col_abs_max(nelim, m - 1, &a[c * lda], idx, maxc);
      int r = -1;
      if (std::fabs(a[c * lda + c]) >= T(u) * maxc) {
         r = c;
      } else {
         int p;
         T maxp;
         col_abs_max(nelim, k - 1, &a[c * lda], p, maxp);
         if (std::fabs(a[c * lda + p]) >= T(u) * maxc) r = p;
      }
      if (r < 0) {
         ++c;
         continue;
      }
      // swap_rc(nelim, r, nelim, c): rows nelim <-> r, columns nelim <-> c
      std::swap(rperm[nelim], rperm[r]);
      std::swap(cperm[nelim], cperm[c]);
      swap_rows(nelim, r, m, a, lda);
      swap_cols(nelim, c, m, a, lda);
      eliminate(m, nelim, a, lda);
      ++nelim;
      c = nelim;
   }
   return nelim;
}

}  // namespace

extern "C" {

int oracle_lu_front(int mode, double u, int m, int k, double* a, int lda, int* rperm, int* cperm) {
   if (m < 0 || k < 0 || k > m || lda < m || mode < 0 || mode > 2) return -1;
   return lu_front<double>(mode, u, m, k, a, lda, rperm, cperm);
}

int oracle_lu_front_f32(int mode, double u, int m, int k, float* a, int lda, int* rperm, int* cperm) {
   if (m < 0 || k < 0 || k > m || lda < m || mode < 0 || mode > 2) return -1;
   return lu_front<float>(mode, u, m, k, a, lda, rperm, cperm);
}

}  // extern "C"
