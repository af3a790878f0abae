// CPU oracle, indefinite path: plain C++ FP64 restatement of the reference's
// APP LDL^T numeric phase with delayed pivots. TEST INFRASTRUCTURE ONLY -- the
// checker of the device path; the library never calls it.
//
// Algorithm (the device kernels make the same decisions):
//   front: m rows, n fully-summed columns (own + delayed from children),
//   held here as a full symmetric m x m matrix S (both triangles).
//   1. APP over blocks of nb columns (factorize_indef_app_notask,
//      src/factor_indef.hxx:600-860):
//      - back up the block column, factor its diagonal block with complete
//        (Bunch-Parlett) 1x1/2x2 pivoting inside the block (the reference
//        calls SPRAL block_ldlt, whose pivot rule is not in the tree; this
//        restatement names its own rule), zero pivots below `small`;
//      - apply the pivots to every uneliminated row outside the block
//        (ApplyN / ApplyT, src/kernels/ldlt_app.hxx:172-255) with the
//        reference's zero-pivot handling, and find the first column with an
//        entry above 1/u (check_threshold :144-165);
//      - npass = that column; drop the first half of a split 2x2
//        (Column::adjust, include/sylver/kernels/Column.hxx:87-101);
//      - restore the failed columns from the backup, update every
//        uneliminated entry of the fully-summed columns with the passed
//        pivots (UpdateN / UpdateT, ldlt_app.hxx:619-718);
//   2. permute_failed (src/factor_indef.hxx:1190-1289): eliminated columns
//      first in elimination order, failed ones after them in block order;
//   3. TPP on the failed columns over all rows (factor_front_indef_failed,
//      src/factor_failed.hxx:29-157; ldlt_tpp_factor is SPRAL's, restated
//      here as textbook threshold partial pivoting with the Duff-Reid 2x2
//      test), then the contribution block update with every eliminated
//      pivot (form_contrib, ldlt_app.hxx:735-751);
//   4. delays: the failed columns go to the parent (assemble_delays,
//      src/kernels/assemble.hxx:926-961), rows shift by ndelay_in
//      (src/assemble.hxx:152-190, init_a_block assemble.hxx:161-214).
// D^{-1} format (ldlt_app.hxx:166-171): 1x1 [d, 0], 2x2 [a, b, +Inf, c].
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace {

// Every product-sum below is written with explicit fused or separate
// roundings (this file is compiled with -ffp-contract=off) so that the
// device kernels, which use the same expressions, round identically.
inline double fnma(double a, double b, double c) { return std::fma(-a, b, c); }  // c - a b
inline double dot2(double a, double b, double c, double d) { return std::fma(a, b, c * d); }

const double kInf = std::numeric_limits<double>::infinity();
const double kAlphaBP = 0.6403882032022076;  // (1 + sqrt(17)) / 8, Bunch-Parlett

struct Sym {
   int m;
   std::vector<double> a;  // full symmetric, column-major
   double& at(int i, int j) { return a[(size_t)j * m + i]; }
   void swap(int p, int q) {  // symmetric permutation of positions p, q
      if (p == q) return;
      for (int k = 0; k < m; ++k) std::swap(at(k, p), at(k, q));
      for (int k = 0; k < m; ++k) std::swap(at(p, k), at(q, k));
   }
};

// y (len 2 or 1) = x * D^{-1} for the pivot starting at d, with the
// reference's zero-pivot handling (apply_pivot, ldlt_app.hxx:172-205).
inline void apply_dinv(const double* d, bool two, double small, double* x) {
   if (two) {
      const double a1 = x[0], a2 = x[1];
      x[0] = dot2(d[0], a1, d[1], a2);
      x[1] = dot2(d[1], a1, d[3], a2);
   } else if (d[0] == 0.0) {
      x[0] = (std::fabs(x[0]) < small) ? 0.0 : kInf * x[0];
   } else {
      x[0] *= d[0];
   }
}

struct FrontResult {
   int nelim = 0, nelim_first = 0, num_two = 0, zero = 0;
   bool singular = false;  // zero pivot met with action == false
};

// Bunch-Parlett LDL^T of the kb x kb diagonal block at positions [c0, c0+kb)
// of S, pivoting inside the block only; swaps are applied to all of S and to
// the backup rows. T holds the block (row/col relative). Writes dinv[2*kb],
// L11 into T's strict lower part, and returns false on a singular block with
// action == false.
bool block_ldlt(Sym& S, std::vector<double>& T, int c0, int kb, double small, bool action,
                std::vector<double>& backup, int& num_zero, double* dinv, std::vector<int>& perm) {
   auto t = [&](int i, int j) -> double& { return T[(size_t)j * kb + i]; };
   auto swap_local = [&](int p, int q) {
      if (p == q) return;
      for (int k = 0; k < kb; ++k) std::swap(t(k, p), t(k, q));
      for (int k = 0; k < kb; ++k) std::swap(t(p, k), t(q, k));
      S.swap(c0 + p, c0 + q);
      // backup is m x kb: rows p,q (positions) and columns p,q
      const int m = S.m;
      for (int k = 0; k < m; ++k) std::swap(backup[(size_t)p * m + k], backup[(size_t)q * m + k]);
      for (int k = 0; k < kb; ++k) std::swap(backup[(size_t)k * m + c0 + p], backup[(size_t)k * m + c0 + q]);
      std::swap(perm[c0 + p], perm[c0 + q]);
   };
   // the trailing block is kept symmetric so that swaps stay valid
   auto mirror = [&](int s0) {
      for (int j = s0; j < kb; ++j)
         for (int i = j + 1; i < kb; ++i) t(j, i) = t(i, j);
   };
   for (int s = 0; s < kb;) {
      double mu0 = -1.0, mu1 = -1.0;
      int i0 = s, j0 = s, i1 = s;
      for (int j = s; j < kb; ++j)
         for (int i = j; i < kb; ++i) {
            const double v = std::fabs(t(i, j));
            if (v > mu0) { mu0 = v; i0 = i; j0 = j; }
         }
      for (int i = s; i < kb; ++i) {
         const double v = std::fabs(t(i, i));
         if (v > mu1) { mu1 = v; i1 = i; }
      }
      if (mu0 < small) {
         // all remaining pivots are zero (1x1, d = 0), their L columns zero
         if (!action) return false;
         for (int k = s; k < kb; ++k) {
            dinv[2 * k] = 0.0;
            dinv[2 * k + 1] = 0.0;
            for (int i = k + 1; i < kb; ++i) t(i, k) = 0.0;
            ++num_zero;
         }
         break;
      }
      if (mu1 >= kAlphaBP * mu0) {
         swap_local(s, i1);
         const double dv = t(s, s);
         dinv[2 * s] = 1.0 / dv;
         dinv[2 * s + 1] = 0.0;
         for (int i = s + 1; i < kb; ++i) {
            const double li = t(i, s) * dinv[2 * s];
            for (int j = s + 1; j <= i; ++j) t(i, j) = fnma(li, t(j, s), t(i, j));
         }
         for (int i = s + 1; i < kb; ++i) t(i, s) *= dinv[2 * s];
         s += 1;
         mirror(s);
      } else {
         // 2x2 on (j0, i0), j0 < i0
         swap_local(s, j0);
         if (i0 == s) i0 = j0;
         swap_local(s + 1, i0);
         const double a = t(s, s), b = t(s + 1, s), c = t(s + 1, s + 1);
         const double det = std::fma(a, c, -(b * b));
         double* d = dinv + 2 * s;
         d[0] = c / det;
         d[1] = -b / det;
         d[2] = kInf;
         d[3] = a / det;
         for (int i = s + 2; i < kb; ++i) {
            double l[2] = {t(i, s), t(i, s + 1)};
            apply_dinv(d, true, small, l);
            for (int j = s + 2; j <= i; ++j)
               t(i, j) = fnma(l[1], t(j, s + 1), fnma(l[0], t(j, s), t(i, j)));
         }
         for (int i = s + 2; i < kb; ++i) {
            double l[2] = {t(i, s), t(i, s + 1)};
            apply_dinv(d, true, small, l);
            t(i, s) = l[0];
            t(i, s + 1) = l[1];
         }
         t(s + 1, s) = 0.0;  // 2x2 block: L11 is the identity there
         s += 2;
         mirror(s);
      }
   }
   return true;
}

inline bool is_two_first(const double* dinv, int k, int nk) {
   return k + 1 < nk && std::isinf(dinv[2 * k + 2]);
}

// Rank-1/2 elimination of the pivot at positions [p, p+w) (w = 1 or 2) of
// S with D^{-1} = d: computes the L column(s) for every row in `rows`, and
// subtracts L D L^T from every (r, s) with r, s in `rows` and s a fully-
// summed position (< n); the contribution block is left for the end.
void eliminate(Sym& S, int p, int w, const double* d, double small, const std::vector<int>& rows,
               int n) {
   const size_t R = rows.size();
   std::vector<double> l(R * 2), v(R * 2);
   for (size_t k = 0; k < R; ++k) {
      const int r = rows[k];
      v[2 * k] = S.at(r, p);
      v[2 * k + 1] = (w == 2) ? S.at(r, p + 1) : 0.0;
      double x[2] = {v[2 * k], v[2 * k + 1]};
      apply_dinv(d, w == 2, small, x);
      l[2 * k] = x[0];
      l[2 * k + 1] = x[1];
   }
   for (size_t kj = 0; kj < R; ++kj) {
      const int s = rows[kj];
      if (s >= n) continue;
      for (size_t ki = 0; ki < R; ++ki) {
         const int r = rows[ki];
         double x = fnma(l[2 * ki], v[2 * kj], S.at(r, s));
         if (w == 2) x = fnma(l[2 * ki + 1], v[2 * kj + 1], x);
         S.at(r, s) = x;
      }
   }
   for (size_t kj = 0; kj < R; ++kj)  // exact symmetry of the fully-summed part
      for (size_t ki = 0; ki < R; ++ki)
         if (rows[ki] < n && rows[ki] > rows[kj]) S.at(rows[kj], rows[ki]) = S.at(rows[ki], rows[kj]);
   for (size_t k = 0; k < R; ++k) {
      const int r = rows[k];
      S.at(r, p) = l[2 * k];
      S.at(p, r) = l[2 * k];
      if (w == 2) {
         S.at(r, p + 1) = l[2 * k + 1];
         S.at(p + 1, r) = l[2 * k + 1];
      }
   }
}

// Dense APP + TPP factorization of one front held in S (m x m), n fully-
// summed columns. perm[0..n) is permuted alongside. dinv gets 2*nelim
// entries. The contribution block (rows/cols >= n) receives the update of
// every eliminated pivot.
FrontResult factor_front(Sym& S, int n, int nb, double u, double small, bool action,
                         bool failed_tpp, std::vector<int>& perm, std::vector<double>& dinv_out) {
   const int m = S.m;
   FrontResult res;
   std::vector<char> elim(m, 0);
   std::vector<double> dpos(2 * (size_t)std::max(n, 1), 0.0);  // D^{-1} by position
   std::vector<int> order;  // eliminated positions in elimination order
   std::vector<int> nelim_blk;
   const int nblk = (n + nb - 1) / nb;
   for (int b = 0; b < nblk; ++b) {
      const int c0 = b * nb, kb = std::min(nb, n - c0);
      std::vector<double> backup((size_t)kb * m);
      for (int j = 0; j < kb; ++j)
         for (int i = 0; i < m; ++i) backup[(size_t)j * m + i] = S.at(i, c0 + j);
      std::vector<double> T((size_t)kb * kb);
      for (int j = 0; j < kb; ++j)
         for (int i = 0; i < kb; ++i) T[(size_t)j * kb + i] = S.at(c0 + i, c0 + j);
      std::vector<double> dloc(2 * (size_t)kb);
      int nz = 0;
      if (!block_ldlt(S, T, c0, kb, small, action, backup, nz, dloc.data(), perm)) {
         res.singular = true;
         return res;
      }
      // ApplyN/ApplyT on every uneliminated row outside the block
      std::vector<int> others;
      for (int r = 0; r < m; ++r)
         if (!elim[r] && (r < c0 || r >= c0 + kb)) others.push_back(r);
      const size_t R = others.size();
      std::vector<double> L(R * kb);
      int npass = kb;
      for (size_t k = 0; k < R; ++k) {
         const int r = others[k];
         double* x = &L[k * kb];
         for (int j = 0; j < kb; ++j) x[j] = S.at(r, c0 + j);
         // x := x L11^{-T} (unit lower L11 in T)
         for (int j = 0; j < kb; ++j)
            for (int i = 0; i < j; ++i) x[j] = fnma(x[i], T[(size_t)i * kb + j], x[j]);
         for (int j = 0; j < kb;) {
            const bool two = is_two_first(dloc.data(), j, kb);
            apply_dinv(&dloc[2 * j], two, small, &x[j]);
            j += two ? 2 : 1;
         }
         for (int j = 0; j < npass; ++j)
            if (std::fabs(x[j]) > 1.0 / u) { npass = j; break; }
      }
      if (npass > 0 && std::isfinite(dloc[2 * (npass - 1)]) && dloc[2 * (npass - 1) + 1] != 0.0)
         --npass;  // first half of a 2x2
      // passed columns: write L; failed columns: restore from the backup
      for (int j = 0; j < npass; ++j) {
         for (size_t k = 0; k < R; ++k) {
            S.at(others[k], c0 + j) = L[k * kb + j];
            S.at(c0 + j, others[k]) = L[k * kb + j];
         }
         for (int i = j + 1; i < kb; ++i) {
            S.at(c0 + i, c0 + j) = T[(size_t)j * kb + i];
            S.at(c0 + j, c0 + i) = T[(size_t)j * kb + i];
         }
      }
      for (int j = npass; j < kb; ++j)
         for (int i = 0; i < m; ++i) {
            if (elim[i]) continue;
            const bool in_block_passed = (i >= c0 && i < c0 + npass);
            if (in_block_passed) continue;  // L[f, p] of a failed row f stays L
            S.at(i, c0 + j) = backup[(size_t)j * m + i];
            S.at(c0 + j, i) = backup[(size_t)j * m + i];
         }
      // failed rows of passed columns inside the block hold L11 entries
      for (int j = 0; j < npass; ++j)
         for (int i = npass; i < kb; ++i) {
            S.at(c0 + i, c0 + j) = T[(size_t)j * kb + i];
            S.at(c0 + j, c0 + i) = T[(size_t)j * kb + i];
         }
      // update uneliminated fully-summed columns with the passed pivots
      std::vector<int> rows;
      for (int r = 0; r < m; ++r)
         if (!elim[r] && !(r >= c0 && r < c0 + npass)) rows.push_back(r);
      for (int j = 0; j < npass;) {
         const bool two = is_two_first(dloc.data(), j, kb);
         const int w = two ? 2 : 1;
         // L columns are already final: subtract l d^{-1}-scaled products
         // using v = L * D (the unscaled column) recovered from L.
         const int p = c0 + j;
         // D (not D^{-1}) of the pivot
         double Dm[4];
         if (two) {
            const double* dv = &dloc[2 * j];
            const double det = std::fma(dv[0], dv[3], -(dv[1] * dv[1]));
            Dm[0] = dv[3] / det; Dm[1] = -dv[1] / det; Dm[3] = dv[0] / det;
         } else {
            Dm[0] = (dloc[2 * j] == 0.0) ? 0.0 : 1.0 / dloc[2 * j];
         }
         for (int s : rows) {
            if (s >= n) continue;
            const double ls0 = S.at(s, p), ls1 = two ? S.at(s, p + 1) : 0.0;
            const double vs0 = two ? dot2(Dm[0], ls0, Dm[1], ls1) : Dm[0] * ls0;
            const double vs1 = two ? dot2(Dm[1], ls0, Dm[3], ls1) : 0.0;
            for (int r : rows) {
               double x = fnma(S.at(r, p), vs0, S.at(r, s));
               if (two) x = fnma(S.at(r, p + 1), vs1, x);
               S.at(r, s) = x;
            }
         }
         j += w;
      }
      // keep the fully-summed block symmetric
      for (int s : rows)
         if (s < n)
            for (int r : rows)
               if (r < n && r > s) S.at(s, r) = S.at(r, s);
      for (int j = 0; j < npass; ++j) {
         elim[c0 + j] = 1;
         order.push_back(c0 + j);
         dpos[2 * (c0 + j)] = dloc[2 * j];
         dpos[2 * (c0 + j) + 1] = dloc[2 * j + 1];
         if (std::isinf(dloc[2 * j])) ++res.num_two;
         if (dloc[2 * j] == 0.0 && dloc[2 * j + 1] == 0.0) ++res.zero;
      }
      nelim_blk.push_back(npass);
      res.nelim += npass;
   }
   res.nelim_first = res.nelim;
   // permute_failed: eliminated first (elimination order), then failed in
   // position (block) order
   std::vector<int> newpos;  // newpos[k] = old position placed at k
   newpos.reserve(m);
   for (int p : order) newpos.push_back(p);
   for (int p = 0; p < n; ++p)
      if (!elim[p]) newpos.push_back(p);
   for (int p = n; p < m; ++p) newpos.push_back(p);
   {
      Sym T2{m, std::vector<double>((size_t)m * m)};
      for (int j = 0; j < m; ++j)
         for (int i = 0; i < m; ++i) T2.at(i, j) = S.at(newpos[i], newpos[j]);
      S.a.swap(T2.a);
      std::vector<int> p2(n);
      std::vector<double> d2(2 * (size_t)std::max(n, 1), 0.0);
      for (int k = 0; k < n; ++k) {
         p2[k] = perm[newpos[k]];
         d2[2 * k] = dpos[2 * newpos[k]];
         d2[2 * k + 1] = dpos[2 * newpos[k] + 1];
      }
      perm.swap(p2);
      dpos.swap(d2);
   }
   int nelim = res.nelim;
   // second pass: TPP on the failed columns
   if (nelim < n && (m == n || failed_tpp)) {
      while (nelim < n) {
         bool found = false;
         for (int p = nelim; p < n && !found; ++p) {
            double amax = 0.0;
            int t = -1;
            for (int i = nelim; i < m; ++i) {
               if (i == p) continue;
               const double v = std::fabs(S.at(i, p));
               if (v > amax) { amax = v; t = i; }
            }
            const double app = std::fabs(S.at(p, p));
            std::vector<int> rows;
            if (std::max(app, amax) < small) {
               if (!action) { res.singular = true; return res; }
               S.swap(p, nelim);
               std::swap(perm[p], perm[nelim]);
               dpos[2 * nelim] = 0.0;
               dpos[2 * nelim + 1] = 0.0;
               for (int i = nelim + 1; i < m; ++i) { S.at(i, nelim) = 0.0; S.at(nelim, i) = 0.0; }
               ++res.zero;
               ++nelim;
               found = true;
               break;
            }
            if (app >= u * amax) {
               S.swap(p, nelim);
               std::swap(perm[p], perm[nelim]);
               double d[4] = {1.0 / S.at(nelim, nelim), 0.0, 0.0, 0.0};
               dpos[2 * nelim] = d[0];
               dpos[2 * nelim + 1] = 0.0;
               for (int i = nelim + 1; i < m; ++i) rows.push_back(i);
               eliminate(S, nelim, 1, d, small, rows, n);
               ++nelim;
               found = true;
               break;
            }
            if (t >= 0 && t < n) {
               const double a11 = S.at(p, p), a21 = S.at(t, p), a22 = S.at(t, t);
               const double det = std::fma(a11, a22, -(a21 * a21));
               if (det != 0.0) {
                  const double d11 = a22 / det, d21 = -a21 / det, d22 = a11 / det;
                  double pmax = 0.0, tmax = 0.0;
                  for (int i = nelim; i < m; ++i) {
                     if (i == p || i == t) continue;
                     pmax = std::max(pmax, std::fabs(S.at(i, p)));
                     tmax = std::max(tmax, std::fabs(S.at(i, t)));
                  }
                  const double g = std::max(dot2(std::fabs(d11), pmax, std::fabs(d21), tmax),
                                            dot2(std::fabs(d21), pmax, std::fabs(d22), tmax));
                  if (u * g <= 1.0) {
                     S.swap(p, nelim);
                     std::swap(perm[p], perm[nelim]);
                     if (t == nelim) t = p;
                     S.swap(t, nelim + 1);
                     std::swap(perm[t], perm[nelim + 1]);
                     const double b11 = S.at(nelim, nelim), b21 = S.at(nelim + 1, nelim),
                                  b22 = S.at(nelim + 1, nelim + 1);
                     const double dt = std::fma(b11, b22, -(b21 * b21));
                     double d[4] = {b22 / dt, -b21 / dt, kInf, b11 / dt};
                     for (int k = 0; k < 4; ++k) dpos[2 * nelim + k] = d[k];
                     for (int i = nelim + 2; i < m; ++i) rows.push_back(i);
                     eliminate(S, nelim, 2, d, small, rows, n);
                     S.at(nelim + 1, nelim) = 0.0;
                     S.at(nelim, nelim + 1) = 0.0;
                     ++res.num_two;
                     nelim += 2;
                     found = true;
                     break;
                  }
               }
            }
         }
         if (!found) break;
      }
   }
   res.nelim = nelim;
   // contribution block: CB -= L D L^T over all eliminated pivots
   const int q = m - n;
   if (q > 0 && nelim > 0) {
      std::vector<double> V((size_t)q * nelim);  // V = L D (rows n..m)
      for (int j = 0; j < nelim;) {
         const bool two = j + 1 < nelim && std::isinf(dpos[2 * j + 2]);
         if (two) {
            const double* dv = &dpos[2 * j];
            const double det = std::fma(dv[0], dv[3], -(dv[1] * dv[1]));
            const double D0 = dv[3] / det, D1 = -dv[1] / det, D3 = dv[0] / det;
            for (int i = 0; i < q; ++i) {
               const double l0 = S.at(n + i, j), l1 = S.at(n + i, j + 1);
               V[(size_t)j * q + i] = dot2(D0, l0, D1, l1);
               V[(size_t)(j + 1) * q + i] = dot2(D1, l0, D3, l1);
            }
            j += 2;
         } else {
            const double D0 = (dpos[2 * j] == 0.0) ? 0.0 : 1.0 / dpos[2 * j];
            for (int i = 0; i < q; ++i) V[(size_t)j * q + i] = D0 * S.at(n + i, j);
            j += 1;
         }
      }
      for (int c = 0; c < q; ++c)
         for (int r = c; r < q; ++r) {
            double s = 0.0;
            for (int k = 0; k < nelim; ++k) s = std::fma(V[(size_t)k * q + r], S.at(n + c, k), s);
            S.at(n + r, n + c) -= s;
            S.at(n + c, n + r) = S.at(n + r, n + c);
         }
   }
   dinv_out.assign(dpos.begin(), dpos.begin() + 2 * (size_t)nelim);
   return res;
}

struct IndefNode {
   int m = 0, n = 0, nelim = 0, ndelay_in = 0;
   std::vector<int> rows;     // global (1-based pivot-order) index of each front row
   std::vector<double> L;     // m x nelim, column-major (ld m); strictly lower meaningful
   std::vector<double> dinv;  // 2 * nelim
   std::vector<double> fail;  // m x (n - nelim): failed columns, rows [0, m)
   std::vector<double> cb;    // q x q, q = m - n, lower
};

struct IndefFactor {
   int n = 0, nnodes = 0;
   std::vector<IndefNode> nodes;
   int num_delay = 0, num_neg = 0, num_two = 0, num_zero = 0, maxfront = 0;
   int not_first_pass = 0, not_second_pass = 0;
   int flag = 0;
};

}  // namespace

extern "C" void* oracle_factor_indef(int n, int nnodes, const int* sptr, const int* sparent,
                                     const int64_t* rptr, const int* rlist, const int64_t* nptr,
                                     const int64_t* nlist, const double* aval, const double* scaling,
                                     int nb, double u, double small, int action, int failed_tpp) {
   IndefFactor* F = new IndefFactor;
   F->n = n;
   F->nnodes = nnodes;
   F->nodes.resize(nnodes);
   std::vector<int> par(nnodes + 1), fc(nnodes + 1), nc(nnodes + 1), ld(nnodes + 1);
   int mf;
   oracle_symbolic_tree(nnodes, sptr, sparent, rptr, par.data(), fc.data(), nc.data(), ld.data(), &mf);
   std::vector<int> map(n + 1, 0);
   for (int ni = 0; ni < nnodes; ++ni) {  // children before parents
      IndefNode& nd = F->nodes[ni];
      const int nrow = (int)(rptr[ni + 1] - rptr[ni]);
      const int ncol = sptr[ni + 1] - sptr[ni];
      const int* rl = rlist + rptr[ni] - 1;
      int ndin = 0;
      for (int c = fc[ni]; c != -1; c = nc[c]) ndin += F->nodes[c].n - F->nodes[c].nelim;
      const int m = nrow + ndin, nf = ncol + ndin;
      nd.m = m;
      nd.n = nf;
      nd.ndelay_in = ndin;
      F->maxfront = std::max(F->maxfront, m);
      Sym S{m, std::vector<double>((size_t)m * m, 0.0)};
      std::vector<int> perm(nf);
      for (int i = 0; i < ncol; ++i) perm[i] = rl[i];
      // init_a_block with the delay shift of rows >= ncol
      for (int64_t p = nptr[ni]; p < nptr[ni + 1]; ++p) {
         const int64_t src = nlist[2 * (p - 1)] - 1;
         const int64_t dest = nlist[2 * (p - 1) + 1] - 1;
         const int c = (int)(dest / nrow);
         const int r0 = (int)(dest % nrow);
         const int r = r0 >= ncol ? r0 + ndin : r0;
         // scaled entry: rscale * a * cscale by symbolic row-list entries
         // (init_a_block, src/kernels/assemble.hxx:166-177)
         const double v = scaling ? scaling[rl[r0] - 1] * aval[src] * scaling[rl[c] - 1] : aval[src];
         S.at(r, c) = v;
         S.at(c, r) = v;
      }
      for (int i = 0; i < ncol; ++i) map[rl[i]] = i;
      for (int i = ncol; i < nrow; ++i) map[rl[i]] = i + ndin;
      int delay_col = ncol;
      for (int c = fc[ni]; c != -1; c = nc[c]) {
         IndefNode& cn = F->nodes[c];
         const int cncol = sptr[c + 1] - sptr[c];
         const int cnrow = (int)(rptr[c + 1] - rptr[c]);
         const int* crl = rlist + rptr[c] - 1;
         const int nd_out = cn.n - cn.nelim;
         // assemble_delays
         for (int i = 0; i < nd_out; ++i) {
            perm[delay_col + i] = cn.rows[cn.nelim + i];
            for (int j = 0; j < nd_out - i; ++j) {
               const double v = cn.fail[(size_t)i * cn.m + cn.nelim + i + j];
               S.at(delay_col + i + j, delay_col + i) = v;
               S.at(delay_col + i, delay_col + i + j) = v;
            }
            for (int j = cncol; j < cnrow; ++j) {
               const int r = map[crl[j]];
               const double v = cn.fail[(size_t)i * cn.m + cn.ndelay_in + j];
               S.at(r, delay_col + i) = v;
               S.at(delay_col + i, r) = v;
            }
         }
         delay_col += nd_out;
         // contribution block (rows rlist_c[ncol_c:]), added
         const int q = cnrow - cncol;
         for (int j = 0; j < q; ++j) {
            const int pj = map[crl[cncol + j]];
            for (int i = j; i < q; ++i) {
               const int pi = map[crl[cncol + i]];
               const double v = cn.cb[(size_t)j * q + i];
               S.at(pi, pj) += v;
               if (pi != pj) S.at(pj, pi) += v;
            }
         }
         std::vector<double>().swap(cn.cb);
         std::vector<double>().swap(cn.fail);
      }
      std::vector<double> dinv;
      FrontResult r = factor_front(S, nf, nb, u, small, action != 0, failed_tpp != 0, perm, dinv);
      if (r.singular) {
         F->flag = -5;
         return F;
      }
      nd.nelim = r.nelim;
      nd.rows.resize(m);
      for (int i = 0; i < nf; ++i) nd.rows[i] = perm[i];
      for (int i = nf; i < m; ++i) nd.rows[i] = rl[ncol + (i - nf)];
      nd.L.assign((size_t)m * r.nelim, 0.0);
      for (int j = 0; j < r.nelim; ++j)
         for (int i = j + 1; i < m; ++i) nd.L[(size_t)j * m + i] = S.at(i, j);
      nd.dinv = dinv;
      const int nfail = nf - r.nelim;
      nd.fail.assign((size_t)m * nfail, 0.0);
      for (int j = 0; j < nfail; ++j)
         for (int i = 0; i < m; ++i) nd.fail[(size_t)j * m + i] = S.at(i, r.nelim + j);
      const int q = m - nf;
      nd.cb.assign((size_t)q * q, 0.0);
      for (int j = 0; j < q; ++j)
         for (int i = j; i < q; ++i) nd.cb[(size_t)j * q + i] = S.at(nf + i, nf + j);
      if (r.nelim_first < nf) F->not_first_pass += nf - r.nelim_first;
      if (failed_tpp || m == nf) F->not_second_pass += nf - r.nelim;
      F->num_delay += nf - r.nelim;
      // inertia from D^{-1} (src/NumericTree.hxx:150-179)
      for (int i = 0; i < r.nelim;) {
         const double a11 = dinv[2 * i], a21 = dinv[2 * i + 1];
         if (i + 1 == r.nelim || std::isfinite(dinv[2 * i + 2])) {
            if (a11 == 0.0) ++F->num_zero;
            if (a11 < 0.0) ++F->num_neg;
            ++i;
         } else {
            const double a22 = dinv[2 * i + 3];
            ++F->num_two;
            const double det = a11 * a22 - a21 * a21, tr = a11 + a22;
            if (det < 0) ++F->num_neg;
            else if (tr < 0) F->num_neg += 2;
            i += 2;
         }
      }
      if (par[ni] >= nnodes && nf - r.nelim > 0) {
         F->flag = -5;  // a root left columns it could not eliminate
         return F;
      }
   }
   if (F->num_zero > 0) F->flag = 7;
   return F;
}

extern "C" void oracle_indef_stats(const void* h, int* out) {
   const IndefFactor* F = static_cast<const IndefFactor*>(h);
   out[0] = F->flag;
   out[1] = F->num_delay;
   out[2] = F->num_neg;
   out[3] = F->num_two;
   out[4] = F->num_zero;
   out[5] = F->maxfront;
   out[6] = F->not_first_pass;
   out[7] = F->not_second_pass;
}

extern "C" void oracle_indef_node(const void* h, int node, int* dims, int* rows, double* L,
                                  double* dinv) {
   const IndefFactor* F = static_cast<const IndefFactor*>(h);
   const IndefNode& nd = F->nodes[node];
   dims[0] = nd.m;
   dims[1] = nd.n;
   dims[2] = nd.nelim;
   dims[3] = nd.ndelay_in;
   if (rows) std::memcpy(rows, nd.rows.data(), sizeof(int) * nd.m);
   if (L) std::memcpy(L, nd.L.data(), sizeof(double) * nd.L.size());
   if (dinv) std::memcpy(dinv, nd.dinv.data(), sizeof(double) * nd.dinv.size());
}

// Solves (NumericTree::solve_fwd / solve_diag_bwd_inner,
// src/NumericTree.hxx:408-532; ldlt_app_solve_fwd/diag/bwd,
// src/kernels/ldlt_app.cxx:53-115): gather by the node's row list, unit
// lower solve and update (forward), D^{-1} (diag), transposed (backward).
// job: 0 = full, 1 = fwd, 2 = diag, 3 = bwd, 4 = diag+bwd. x in pivot order.
extern "C" void oracle_solve_indef(const void* h, int job, int nrhs, double* x, int ldx) {
   const IndefFactor* F = static_cast<const IndefFactor*>(h);
   const bool fwd = job == 0 || job == 1, diag = job == 0 || job == 2 || job == 4,
              bwd = job == 0 || job == 3 || job == 4;
   std::vector<double> xl;
   for (int c = 0; c < nrhs; ++c) {
      double* xc = x + (size_t)c * ldx;
      if (fwd)
         for (int ni = 0; ni < F->nnodes; ++ni) {
            const IndefNode& nd = F->nodes[ni];
            xl.assign(nd.m, 0.0);
            for (int i = 0; i < nd.m; ++i) xl[i] = xc[nd.rows[i] - 1];
            for (int j = 0; j < nd.nelim; ++j)
               for (int i = j + 1; i < nd.m; ++i) xl[i] -= nd.L[(size_t)j * nd.m + i] * xl[j];
            for (int i = 0; i < nd.m; ++i) xc[nd.rows[i] - 1] = xl[i];
         }
      for (int ni = F->nnodes - 1; ni >= 0; --ni) {
         const IndefNode& nd = F->nodes[ni];
         if (!diag && !bwd) break;
         xl.assign(nd.m, 0.0);
         for (int i = 0; i < nd.m; ++i) xl[i] = xc[nd.rows[i] - 1];
         if (diag)
            for (int i = 0; i < nd.nelim;) {
               const double* d = &nd.dinv[2 * i];
               if (i + 1 < nd.nelim && std::isinf(d[2])) {
                  const double a = xl[i], b = xl[i + 1];
                  xl[i] = d[0] * a + d[1] * b;
                  xl[i + 1] = d[1] * a + d[3] * b;
                  i += 2;
               } else {
                  xl[i] *= d[0];
                  i += 1;
               }
            }
         if (bwd)
            for (int j = nd.nelim - 1; j >= 0; --j)
               for (int i = j + 1; i < nd.m; ++i) xl[j] -= nd.L[(size_t)j * nd.m + i] * xl[i];
         for (int i = 0; i < nd.nelim; ++i) xc[nd.rows[i] - 1] = xl[i];
      }
   }
}

extern "C" void oracle_indef_free(void* h) { delete static_cast<IndefFactor*>(h); }

// Dense front entry point for the closed-form tests: factor the m x m
// symmetric matrix a (full, column-major, ld m) with n fully-summed
// columns; returns nelim, writes perm (n), dinv (2 n), and overwrites a with
// the factored front (L in columns [0, nelim), failed columns, updated CB).
extern "C" int oracle_front_ldlt(int m, int n, double* a, int nb, double u, double small,
                                 int action, int failed_tpp, int* perm, double* dinv,
                                 int* nelim_first) {
   Sym S{m, std::vector<double>(a, a + (size_t)m * m)};
   std::vector<int> p(n);
   for (int i = 0; i < n; ++i) p[i] = i;
   std::vector<double> d;
   FrontResult r = factor_front(S, n, nb, u, small, action != 0, failed_tpp != 0, p, d);
   if (r.singular) return -5;
   std::memcpy(a, S.a.data(), sizeof(double) * (size_t)m * m);
   for (int i = 0; i < n; ++i) perm[i] = p[i];
   std::memset(dinv, 0, sizeof(double) * 2 * (size_t)n);
   std::memcpy(dinv, d.data(), sizeof(double) * d.size());
   if (nelim_first) *nelim_first = r.nelim_first;
   return r.nelim;
}
