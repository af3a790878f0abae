// Matching-based scalings and matching-based ordering (host code); the
// methods and their sources are stated in scaling.hpp.
#include "scaling.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>
#include <queue>
#include <utility>
#include <vector>

#include "analyse.hpp"

namespace sylver_amd {

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

// Bipartite graph of the full symmetric matrix: left vertex j = column j,
// right vertex i = row i, one edge per nonzero a_ij with the assignment cost
// c(j, i) = log cmax_j - log |a_ij| >= 0 (cmax_j = largest |a| of column j).
// Maximizing the product of the matched |a_ij| is minimizing the summed cost.
struct Graph {
   int n = 0;
   std::vector<int64_t> ptr;    // n + 1
   std::vector<int> idx;        // right vertex of each edge
   std::vector<double> cost;
   std::vector<double> absval;
   std::vector<double> lcmax;   // log cmax_j (-inf for an empty column)
};

Graph build_graph(int n, const int64_t* ptr, const int* row, const double* val) {
   Graph g;
   g.n = n;
   g.ptr.assign(n + 1, 0);
   for (int j = 0; j < n; ++j)
      for (int64_t p = ptr[j] - 1; p < ptr[j + 1] - 1; ++p) {
         const int i = row[p] - 1;
         if (val[p] == 0.0) continue;
         g.ptr[j + 1]++;
         if (i != j) g.ptr[i + 1]++;
      }
   for (int j = 0; j < n; ++j) g.ptr[j + 1] += g.ptr[j];
   g.idx.resize(g.ptr[n]);
   g.absval.resize(g.ptr[n]);
   std::vector<int64_t> pos(g.ptr.begin(), g.ptr.end() - 1);
   for (int j = 0; j < n; ++j)
      for (int64_t p = ptr[j] - 1; p < ptr[j + 1] - 1; ++p) {
         const int i = row[p] - 1;
         if (val[p] == 0.0) continue;
         const double a = std::fabs(val[p]);
         g.idx[pos[j]] = i;
         g.absval[pos[j]++] = a;
         if (i != j) {
            g.idx[pos[i]] = j;
            g.absval[pos[i]++] = a;
         }
      }
   g.lcmax.assign(n, -kInf);
   for (int j = 0; j < n; ++j) {
      double m = 0.0;
      for (int64_t p = g.ptr[j]; p < g.ptr[j + 1]; ++p) m = std::max(m, g.absval[p]);
      if (m > 0.0) g.lcmax[j] = std::log(m);
   }
   g.cost.resize(g.ptr[n]);
   for (int j = 0; j < n; ++j)
      for (int64_t p = g.ptr[j]; p < g.ptr[j + 1]; ++p) g.cost[p] = g.lcmax[j] - std::log(g.absval[p]);
   return g;
}

// Symmetric scaling from dual potentials (v on columns, u on rows) that
// satisfy v_j + u_i <= c(j, i): s_i = exp((u_i + v_i - log cmax_i) / 2)
// bounds every |s_i a_ij s_j| by 1. Variables in `bad` (unmatched) are
// scaled afterwards so that their largest entry towards the others is 1.
void duals_to_scaling(const Graph& g, const std::vector<double>& u, const std::vector<double>& v,
                      const std::vector<char>& bad, double* scaling) {
   const int n = g.n;
   for (int i = 0; i < n; ++i)
      scaling[i] = (bad[i] || !std::isfinite(g.lcmax[i])) ? 0.0 : std::exp(0.5 * (u[i] + v[i] - g.lcmax[i]));
   for (int i = 0; i < n; ++i) {
      if (scaling[i] > 0.0) continue;
      double m = 0.0;
      for (int64_t p = g.ptr[i]; p < g.ptr[i + 1]; ++p) {
         const int k = g.idx[p];
         if (k != i && scaling[k] > 0.0 && !bad[k]) m = std::max(m, g.absval[p] * scaling[k]);
      }
      scaling[i] = m > 0.0 ? 1.0 / m : 1.0;
   }
}

// Minimum-cost matching by shortest augmenting paths (Dijkstra on reduced
// costs c - v_j - u_i >= 0). match_l[j] = row matched to column j (-1 when
// none), u/v the final potentials. Returns the number of matched columns.
int shortest_path_matching(const Graph& g, std::vector<int>& match_l, std::vector<double>& u,
                           std::vector<double>& v) {
   const int n = g.n;
   match_l.assign(n, -1);
   std::vector<int> match_r(n, -1);
   v.assign(n, 0.0);
   u.assign(n, kInf);
   // initial potentials: v_j = min_i c(j, i) (= 0: the largest entry), then
   // u_i = min_j c(j, i) - v_j; greedy matching on tight edges
   for (int j = 0; j < n; ++j) {
      double m = kInf;
      for (int64_t p = g.ptr[j]; p < g.ptr[j + 1]; ++p) m = std::min(m, g.cost[p]);
      v[j] = std::isfinite(m) ? m : 0.0;
   }
   for (int j = 0; j < n; ++j)
      for (int64_t p = g.ptr[j]; p < g.ptr[j + 1]; ++p) {
         const int i = g.idx[p];
         u[i] = std::min(u[i], g.cost[p] - v[j]);
      }
   for (int i = 0; i < n; ++i)
      if (!std::isfinite(u[i])) u[i] = 0.0;
   int matched = 0;
   for (int j = 0; j < n; ++j)
      for (int64_t p = g.ptr[j]; p < g.ptr[j + 1]; ++p) {
         const int i = g.idx[p];
         if (match_r[i] < 0 && g.cost[p] - v[j] - u[i] <= 0.0) {
            match_l[j] = i;
            match_r[i] = j;
            ++matched;
            break;
         }
      }
   std::vector<double> dist(n, kInf);
   std::vector<int> pred(n, -1);
   std::vector<char> done(n, 0);
   std::vector<int> touched, popped, scanned;
   using Item = std::pair<double, int>;
   for (int j0 = 0; j0 < n; ++j0) {
      if (match_l[j0] >= 0 || g.ptr[j0] == g.ptr[j0 + 1]) continue;
      std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
      touched.clear();
      popped.clear();
      scanned.clear();
      auto relax = [&](int j, double lj) {
         for (int64_t p = g.ptr[j]; p < g.ptr[j + 1]; ++p) {
            const int k = g.idx[p];
            if (done[k]) continue;
            const double nd = lj + std::max(0.0, g.cost[p] - v[j] - u[k]);
            if (nd < dist[k]) {
               if (dist[k] == kInf) touched.push_back(k);
               dist[k] = nd;
               pred[k] = j;
               heap.push({nd, k});
            }
         }
      };
      scanned.push_back(j0);
      relax(j0, 0.0);
      int found = -1;
      double dlast = 0.0;
      while (!heap.empty()) {
         const auto [d, i] = heap.top();
         heap.pop();
         if (done[i] || d > dist[i]) continue;
         done[i] = 1;
         popped.push_back(i);
         if (match_r[i] < 0) {
            found = i;
            dlast = d;
            break;
         }
         const int j = match_r[i];
         scanned.push_back(j);
         relax(j, d);
      }
      if (found >= 0) {
         // potentials: scanned columns v_j += dlast - l(j) (l = 0 for the
         // start, d(its matched row) otherwise), popped rows u_i -= dlast - d(i)
         for (int j : scanned) v[j] += dlast - (j == j0 ? 0.0 : dist[match_l[j]]);
         for (int i : popped) u[i] -= dlast - dist[i];
         // augment along the predecessor chain
         int i = found;
         for (;;) {
            const int j = pred[i];
            const int next = match_l[j];
            match_l[j] = i;
            match_r[i] = j;
            if (j == j0) break;
            i = next;
         }
         ++matched;
      }
      for (int k : touched) {
         dist[k] = kInf;
         pred[k] = -1;
         done[k] = 0;
      }
   }
   return matched;
}

}  // namespace

int hungarian_scale_sym(int n, const int64_t* ptr, const int* row, const double* val, bool scale_if_singular,
                        double* scaling, int* match) {
   if (n <= 0) return 0;
   const Graph g = build_graph(n, ptr, row, val);
   std::vector<int> ml;
   std::vector<double> u, v;
   const int matched = shortest_path_matching(g, ml, u, v);
   const bool singular = matched < n;
   if (singular && !scale_if_singular) return -2;
   std::vector<char> bad(n, 0);
   std::vector<char> rmatched(n, 0);
   for (int j = 0; j < n; ++j)
      if (ml[j] >= 0) rmatched[ml[j]] = 1;
   for (int i = 0; i < n; ++i) bad[i] = (ml[i] < 0 || !rmatched[i]) ? 1 : 0;
   duals_to_scaling(g, u, v, bad, scaling);
   if (match)
      for (int j = 0; j < n; ++j) match[j] = ml[j];
   return singular ? 1 : 0;
}

int auction_scale_sym(int n, const int64_t* ptr, const int* row, const double* val, double* scaling,
                      int* match) {
   if (n <= 0) return 0;
   const Graph g = build_graph(n, ptr, row, val);
   // maximize sum b(j, i), b = -c, with prices p_i on the rows
   std::vector<double> price(n, 0.0);
   std::vector<int> owner(n, -1), assign(n, -1);
   const int64_t nnz = g.ptr[n];
   const double eps_final = 1e-4;
   int unmatched = 0;
   for (double eps = 0.5;; eps = std::max(eps / 5.0, eps_final)) {
      std::fill(owner.begin(), owner.end(), -1);
      std::fill(assign.begin(), assign.end(), -1);
      std::vector<int> queue;
      for (int j = n - 1; j >= 0; --j)
         if (g.ptr[j] < g.ptr[j + 1]) queue.push_back(j);
      int64_t bids = 0;
      const int64_t cap = 40 * nnz + 100 * (int64_t)n;
      while (!queue.empty() && bids < cap) {
         const int j = queue.back();
         queue.pop_back();
         double w1 = -kInf, w2 = -kInf;
         int i1 = -1;
         for (int64_t p = g.ptr[j]; p < g.ptr[j + 1]; ++p) {
            const double w = -g.cost[p] - price[g.idx[p]];
            if (w > w1) {
               w2 = w1;
               w1 = w;
               i1 = g.idx[p];
            } else if (w > w2) {
               w2 = w;
            }
         }
         // one candidate only: a bid that any competitor would have to beat by far
         const double incr = std::isfinite(w2) ? (w1 - w2) + eps : 16.0 + eps;
         price[i1] += incr;
         ++bids;
         if (owner[i1] >= 0) {
            assign[owner[i1]] = -1;
            queue.push_back(owner[i1]);
         }
         owner[i1] = j;
         assign[j] = i1;
      }
      unmatched = (int)queue.size();
      if (eps <= eps_final) break;
   }
   // duals: v_j = -max_i (b(j, i) - p_i) (the column's profit), u_i = -p_i;
   // then v_j + u_i <= c(j, i) on every edge (exactly dual feasible)
   std::vector<double> u(n), v(n, 0.0);
   for (int i = 0; i < n; ++i) u[i] = -price[i];
   for (int j = 0; j < n; ++j) {
      double best = -kInf;
      for (int64_t p = g.ptr[j]; p < g.ptr[j + 1]; ++p) best = std::max(best, -g.cost[p] - price[g.idx[p]]);
      v[j] = std::isfinite(best) ? -best : 0.0;
   }
   std::vector<char> bad(n, 0), rmatched(n, 0);
   for (int j = 0; j < n; ++j)
      if (assign[j] >= 0) rmatched[assign[j]] = 1;
   for (int i = 0; i < n; ++i) bad[i] = (assign[i] < 0 || !rmatched[i]) ? 1 : 0;
   duals_to_scaling(g, u, v, bad, scaling);
   if (match)
      for (int j = 0; j < n; ++j) match[j] = assign[j];
   return unmatched > 0 ? 1 : 0;
}

int match_order(int n, const int64_t* ptr, const int* row, const double* val, int* order, double* scaling) {
   if (n <= 0) return 0;
   std::vector<int> m(n);
   const int flag = hungarian_scale_sym(n, ptr, row, val, true, scaling, m.data());
   // cycles of the matching permutation cut into pairs (2x2 candidates); an
   // odd cycle leaves one single variable, unmatched variables stay single
   std::vector<int> comp(n, -1), first(n, -1), second(n, -1);
   int nc = 0;
   std::vector<char> seen(n, 0);
   for (int j = 0; j < n; ++j) {
      if (seen[j]) continue;
      if (m[j] < 0 || m[j] == j) {
         seen[j] = 1;
         first[nc] = j;
         comp[j] = nc++;
         continue;
      }
      std::vector<int> cyc;
      for (int k = j; !seen[k] && k >= 0; k = m[k]) {
         seen[k] = 1;
         cyc.push_back(k);
      }
      size_t t = 0;
      for (; t + 1 < cyc.size(); t += 2) {
         first[nc] = cyc[t];
         second[nc] = cyc[t + 1];
         comp[cyc[t]] = comp[cyc[t + 1]] = nc++;
      }
      if (t < cyc.size()) {
         first[nc] = cyc[t];
         comp[cyc[t]] = nc++;
      }
   }
   // compressed lower pattern (1-based), diagonal included
   std::vector<std::vector<int>> cols(nc);
   for (int c = 0; c < nc; ++c) cols[c].push_back(c);
   for (int j = 0; j < n; ++j)
      for (int64_t p = ptr[j] - 1; p < ptr[j + 1] - 1; ++p) {
         const int a = comp[row[p] - 1], b = comp[j];
         if (a != b) cols[std::min(a, b)].push_back(std::max(a, b));
      }
   std::vector<int64_t> cptr(nc + 1, 1);
   std::vector<int> crow;
   for (int c = 0; c < nc; ++c) {
      auto& v = cols[c];
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
      for (int r : v) crow.push_back(r + 1);
      cptr[c + 1] = (int64_t)crow.size() + 1;
      std::vector<int>().swap(v);
   }
   std::vector<int> corder(nc);
   nested_dissection(nc, cptr.data(), crow.data(), corder.data());
   std::vector<int> seq(nc);
   for (int c = 0; c < nc; ++c) seq[corder[c] - 1] = c;
   int pos = 0;
   for (int c : seq) {
      order[first[c]] = ++pos;
      if (second[c] >= 0) order[second[c]] = ++pos;
   }
   return flag == 1 ? 1 : 0;
}

}  // namespace sylver_amd
