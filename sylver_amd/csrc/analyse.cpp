// Symbolic analysis on the CPU (stays on the CPU by design: the device numeric
// phase consumes its outputs). The reference delegates ordering and
// supernode detection to SPRAL basic_analyse and METIS
// (src/spldlt_analyse_mod.F90:363,776-782), which are not part of the
// reference tree; this file is this library's own implementation with the
// same outputs and conventions. build_map and the symbolic tree follow the
// reference files cited at each function.
#include "analyse.hpp"
#include "scaling.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <stdexcept>

namespace sylver_amd {

namespace {

constexpr int kErrANOOR = -2;
constexpr int kErrAPtr = -3;
constexpr int kErrAAllOOR = -4;
constexpr int kErrOrder = -8;
constexpr int kErrVal = -9;
constexpr int kWarnIdxOOR = 1;
constexpr int kWarnDupIdx = 2;
constexpr int kWarnDupAndOOR = 3;
constexpr int kWarnMissingDiag = 4;
constexpr int kWarnMissDiagOORDup = 5;
constexpr int kWarnAnalSingular = 6;
// Fraction of explicit zeros accepted when merging a chain of supernodes.
double kRelaxFill = 0.02;

// Adjacency (both triangles, no diagonal) in 0-based CSR form.
struct Graph {
   int n = 0;
   std::vector<int64_t> xadj;
   std::vector<int> adj;
};

Graph make_graph(int n, const int64_t* ptr, const int* row) {
   Graph g;
   g.n = n;
   std::vector<int64_t> cnt(n + 1, 0);
   for (int j = 0; j < n; ++j)
      for (int64_t p = ptr[j] - 1; p < ptr[j + 1] - 1; ++p) {
         int i = row[p] - 1;
         if (i == j) continue;
         cnt[i + 1]++;
         cnt[j + 1]++;
      }
   for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
   g.xadj = cnt;
   g.adj.resize(cnt[n]);
   std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
   for (int j = 0; j < n; ++j)
      for (int64_t p = ptr[j] - 1; p < ptr[j + 1] - 1; ++p) {
         int i = row[p] - 1;
         if (i == j) continue;
         g.adj[pos[i]++] = j;
         g.adj[pos[j]++] = i;
      }
   return g;
}

// Elimination tree of the matrix whose lower pattern is given column-wise
// (0-based, rows > col) -- Liu's algorithm with path compression.
std::vector<int> etree_from_rows(int n, const std::vector<int64_t>& rptr,
                                 const std::vector<int>& rcols) {
   std::vector<int> parent(n, -1), ancestor(n, -1);
   for (int i = 0; i < n; ++i) {
      for (int64_t p = rptr[i]; p < rptr[i + 1]; ++p) {
         int r = rcols[p];
         while (r != -1 && r < i) {
            int next = ancestor[r];
            ancestor[r] = i;
            if (next == -1) {
               parent[r] = i;
               break;
            }
            r = next;
         }
      }
   }
   return parent;
}

// Postorder of a forest; children visited in ascending order.
std::vector<int> postorder(int n, const std::vector<int>& parent) {
   std::vector<int> head(n, -1), next(n, -1), post;
   post.reserve(n);
   for (int j = n - 1; j >= 0; --j) {
      if (parent[j] == -1) continue;
      next[j] = head[parent[j]];
      head[parent[j]] = j;
   }
   std::vector<int> stack;
   for (int root = 0; root < n; ++root) {
      if (parent[root] != -1) continue;
      stack.push_back(root);
      while (!stack.empty()) {
         int p = stack.back();
         int c = head[p];
         if (c == -1) {
            stack.pop_back();
            post.push_back(p);
         } else {
            head[p] = next[c];
            stack.push_back(c);
         }
      }
   }
   return post;
}

// Column counts of the Cholesky factor (row-subtree skeleton method of
// Gilbert, Ng and Peyton). Input: postordered matrix (post = identity), lower
// pattern by columns (rows > col). Returns |L(:,j)| including the diagonal.
std::vector<int> column_counts(int n, const std::vector<int>& parent,
                               const std::vector<int64_t>& cptr,
                               const std::vector<int>& crow) {
   std::vector<int> delta(n), first(n, -1), maxfirst(n, -1), prevleaf(n, -1),
       ancestor(n);
   for (int k = 0; k < n; ++k) {
      delta[k] = (first[k] == -1) ? 1 : 0;
      for (int j = k; j != -1 && first[j] == -1; j = parent[j]) first[j] = k;
   }
   for (int i = 0; i < n; ++i) ancestor[i] = i;
   for (int j = 0; j < n; ++j) {
      if (parent[j] != -1) delta[parent[j]]--;
      for (int64_t p = cptr[j]; p < cptr[j + 1]; ++p) {
         int i = crow[p];
         if (i <= j || first[j] <= maxfirst[i]) continue;
         maxfirst[i] = first[j];
         int jprev = prevleaf[i];
         prevleaf[i] = j;
         delta[j]++;
         if (jprev != -1) {
            int q = jprev;
            while (q != ancestor[q]) q = ancestor[q];
            for (int s = jprev; s != q;) {
               int sp = ancestor[s];
               ancestor[s] = q;
               s = sp;
            }
            delta[q]--;
         }
      }
      if (parent[j] != -1) ancestor[j] = parent[j];
   }
   for (int j = 0; j < n; ++j)
      if (parent[j] != -1) delta[parent[j]] += delta[j];
   return delta;
}

}  // namespace

int64_t node_flops(int64_t m, int64_t n) {
   // src/spldlt_analyse_mod.F90:1744-1764: sum_{jj=m-n+1}^{m} jj^2
   int64_t f = 0;
   for (int64_t jj = m - n + 1; jj <= m; ++jj) f += jj * jj;
   return f;
}

// ---------------------------------------------------------------------------
// Nested dissection for general graphs (the reference orders with METIS,
// src/spldlt_analyse_mod.F90:776-782; METIS is not available here).
// Recursive bisection by breadth-first level structures from a
// pseudo-peripheral vertex; the separator is the middle level, trimmed to
// the vertices that touch the far side. Small parts are ordered by minimum
// degree on the explicit elimination graph.
// ---------------------------------------------------------------------------
namespace {

struct NDState {
   const Graph* g;
   std::vector<int> mark;       // subgraph id of each vertex (current part)
   std::vector<int> level;
   std::vector<int> out;        // elimination sequence being built (reverse)
   int next_id = 1;
};

// BFS restricted to vertices with mark == id; returns vertices in BFS order
// and fills level; `last` receives a vertex of the deepest level.
int bfs(NDState& st, const std::vector<int>& verts, int start, int id,
        std::vector<int>& order, int& last) {
   const Graph& g = *st.g;
   order.clear();
   for (int v : verts) st.level[v] = -1;
   st.level[start] = 0;
   order.push_back(start);
   size_t head = 0;
   int depth = 0;
   while (head < order.size()) {
      int v = order[head++];
      depth = st.level[v];
      for (int64_t p = g.xadj[v]; p < g.xadj[v + 1]; ++p) {
         int w = g.adj[p];
         if (st.mark[w] != id || st.level[w] != -1) continue;
         st.level[w] = st.level[v] + 1;
         order.push_back(w);
      }
   }
   // pick a minimum-degree vertex in the last level
   int best = order.back(), bestdeg = 1 << 30;
   for (size_t k = order.size(); k-- > 0;) {
      int v = order[k];
      if (st.level[v] != depth) break;
      int deg = 0;
      for (int64_t p = g.xadj[v]; p < g.xadj[v + 1]; ++p)
         if (st.mark[g.adj[p]] == id) ++deg;
      if (deg < bestdeg) {
         bestdeg = deg;
         best = v;
      }
   }
   last = best;
   return depth;
}

void order_small(NDState& st, const std::vector<int>& verts, int id,
                 std::vector<int>& seq) {
   // Minimum degree on the explicit elimination graph of a small part.
   const Graph& g = *st.g;
   int k = (int)verts.size();
   std::vector<int> local(k);
   std::vector<std::vector<char>> adj(k, std::vector<char>(k, 0));
   // map global -> local through level[] scratch
   for (int a = 0; a < k; ++a) st.level[verts[a]] = a;
   for (int a = 0; a < k; ++a) {
      int v = verts[a];
      for (int64_t p = g.xadj[v]; p < g.xadj[v + 1]; ++p) {
         int w = g.adj[p];
         if (st.mark[w] != id) continue;
         adj[a][st.level[w]] = 1;
      }
   }
   std::vector<char> done(k, 0);
   for (int step = 0; step < k; ++step) {
      int best = -1, bestdeg = 1 << 30;
      for (int a = 0; a < k; ++a) {
         if (done[a]) continue;
         int deg = 0;
         for (int b = 0; b < k; ++b)
            if (!done[b] && adj[a][b]) ++deg;
         if (deg < bestdeg) {
            bestdeg = deg;
            best = a;
         }
      }
      done[best] = 1;
      seq.push_back(verts[best]);
      // clique the neighbours
      std::vector<int> nb;
      for (int b = 0; b < k; ++b)
         if (!done[b] && adj[best][b]) nb.push_back(b);
      for (int x : nb)
         for (int y : nb)
            if (x != y) adj[x][y] = 1;
   }
}

}  // namespace

void nested_dissection(int n, const int64_t* ptr, const int* row, int* order) {
   Graph g = make_graph(n, ptr, row);
   NDState st;
   st.g = &g;
   st.mark.assign(n, 0);
   st.level.assign(n, -1);
   // Each work item: a vertex set (all with the same mark id). Emit order:
   // part A, part B, separator -- built by pushing onto a stack of "emit"
   // records so that separators follow their parts.
   struct Item {
      std::vector<int> verts;
      bool emit;  // emit verts as-is (a separator)
   };
   std::vector<Item> stack;
   std::vector<int> seq;
   seq.reserve(n);
   {
      std::vector<int> all(n);
      std::iota(all.begin(), all.end(), 0);
      stack.push_back({std::move(all), false});
   }
   std::vector<int> bfsorder;
   const int kLeaf = 64;
   while (!stack.empty()) {
      Item it = std::move(stack.back());
      stack.pop_back();
      if (it.emit) {
         for (int v : it.verts) seq.push_back(v);
         continue;
      }
      int id = st.next_id++;
      for (int v : it.verts) st.mark[v] = id;
      if ((int)it.verts.size() <= kLeaf) {
         order_small(st, it.verts, id, seq);
         continue;
      }
      // connected components
      std::vector<std::vector<int>> comps;
      for (int v : it.verts) st.level[v] = -1;
      for (int v : it.verts) {
         if (st.level[v] != -1) continue;
         int last;
         bfs(st, std::vector<int>{v}, v, id, bfsorder, last);
         // bfs() reset only `v`; mark visited with level >= 0 already
         comps.push_back(bfsorder);
      }
      if (comps.size() > 1) {
         for (auto it2 = comps.rbegin(); it2 != comps.rend(); ++it2)
            stack.push_back({std::move(*it2), false});
         continue;
      }
      std::vector<int>& verts = comps[0];
      // pseudo-peripheral vertex
      int start = verts[0], last;
      int depth = bfs(st, verts, start, id, bfsorder, last);
      for (int iter = 0; iter < 4; ++iter) {
         int d2 = bfs(st, verts, last, id, bfsorder, last);
         if (d2 <= depth) break;
         depth = d2;
         start = bfsorder[0];
      }
      depth = bfs(st, verts, start, id, bfsorder, last);
      if (depth < 2) {
         order_small(st, verts, id, seq);
         continue;
      }
      // level sizes, pick the level splitting the vertices in half
      std::vector<int64_t> cnt(depth + 1, 0);
      for (int v : bfsorder) cnt[st.level[v]]++;
      int64_t half = (int64_t)bfsorder.size() / 2, acc = 0;
      int sl = 1;
      for (int l = 0; l <= depth; ++l) {
         if (acc + cnt[l] > half) {
            sl = l;
            break;
         }
         acc += cnt[l];
      }
      sl = std::max(1, std::min(sl, depth - 1));
      std::vector<int> A, B, S;
      for (int v : bfsorder) {
         int l = st.level[v];
         if (l < sl) {
            A.push_back(v);
         } else if (l > sl) {
            B.push_back(v);
         } else {
            bool touches = false;
            for (int64_t p = g.xadj[v]; p < g.xadj[v + 1]; ++p) {
               int w = g.adj[p];
               if (st.mark[w] == id && st.level[w] == sl + 1) {
                  touches = true;
                  break;
               }
            }
            (touches ? S : A).push_back(v);
         }
      }
      stack.push_back({std::move(S), true});
      if (!B.empty()) stack.push_back({std::move(B), false});
      if (!A.empty()) stack.push_back({std::move(A), false});
   }
   for (int k = 0; k < n; ++k) order[seq[k]] = k + 1;
}

// ---------------------------------------------------------------------------
// build_map (src/spldlt_analyse_mod.F90:130-232). For each node, first the
// entries reached through the transpose of A-lower (entries of A whose column
// index is above the pivot), then the entries of A-lower itself; entries
// mapping above the node's pivot are skipped. Positions are
// (col-sptr)*nrow + local row, 1-based.
// ---------------------------------------------------------------------------
void build_map(int n, const int64_t* ptr, const int* row, const int* perm,
               const int* invp, int nnodes, const int* sptr, const int64_t* rptr,
               const int* rlist, int64_t* nptr, int64_t* nlist) {
   int64_t nnz = ptr[n] - 1;
   std::vector<int> map(n + 1, 0);
   std::vector<int64_t> ptr2(n + 3, 0);
   std::vector<int> row2(nnz > 0 ? nnz : 1);
   std::vector<int64_t> origin(nnz > 0 ? nnz : 1);
   // transpose of A without the diagonal; ptr2 is 1-based like the reference
   for (int i = 1; i <= n; ++i)
      for (int64_t jj = ptr[i - 1]; jj < ptr[i]; ++jj) {
         int k = row[jj - 1];
         if (k == i) continue;
         ptr2[k + 1] += 1;  // ptr2(k+2) in 1-based storage
      }
   ptr2[0] = 1;
   ptr2[1] = 1;
   for (int i = 1; i <= n; ++i) ptr2[i + 1] += ptr2[i];
   for (int i = 1; i <= n; ++i)
      for (int64_t jj = ptr[i - 1]; jj < ptr[i]; ++jj) {
         int k = row[jj - 1];
         if (k == i) continue;
         int64_t at = ptr2[k];  // ptr2(k+1)
         row2[at - 1] = i;
         origin[at - 1] = jj;
         ptr2[k] = at + 1;
      }
   int64_t pp = 1;
   for (int node = 1; node <= nnodes; ++node) {
      int blkm = (int)(rptr[node] - rptr[node - 1]);
      nptr[node - 1] = pp;
      for (int64_t jj = rptr[node - 1]; jj < rptr[node]; ++jj)
         map[rlist[jj - 1]] = (int)(jj - rptr[node - 1] + 1);
      for (int j = sptr[node - 1]; j < sptr[node]; ++j) {
         int col = invp[j - 1];
         for (int64_t i = ptr2[col - 1]; i < ptr2[col]; ++i) {
            int k = std::abs(perm[row2[i - 1] - 1]);
            if (k < j) continue;
            nlist[2 * (pp - 1) + 1] = (int64_t)(j - sptr[node - 1]) * blkm + map[k];
            nlist[2 * (pp - 1) + 0] = origin[i - 1];
            ++pp;
         }
      }
      for (int j = sptr[node - 1]; j < sptr[node]; ++j) {
         int col = invp[j - 1];
         for (int64_t ii = ptr[col - 1]; ii < ptr[col]; ++ii) {
            int k = std::abs(perm[row[ii - 1] - 1]);
            if (k < j) continue;
            nlist[2 * (pp - 1) + 1] = (int64_t)(j - sptr[node - 1]) * blkm + map[k];
            nlist[2 * (pp - 1) + 0] = ii;
            ++pp;
         }
      }
   }
   nptr[nnodes] = pp;
}

// ---------------------------------------------------------------------------
// Symbolic tree (src/SymbolicTree.cxx:7-102): per-node sizes, parent
// (0-based, nnodes = virtual root), least_desc, child linked list built by
// prepending in node order, factor size.
// ---------------------------------------------------------------------------
void build_symbolic_tree(Analysis& a) {
   int nn = a.nnodes;
   a.fronts.assign(nn + 1, SymbolicFront());
   for (int ni = 0; ni <= nn; ++ni) a.fronts[ni].least_desc = ni;
   a.maxfront = 0;
   for (int ni = 0; ni < nn; ++ni) {
      SymbolicFront& f = a.fronts[ni];
      f.nrow = (int)(a.rptr[ni + 1] - a.rptr[ni]);
      f.ncol = a.sptr[ni + 1] - a.sptr[ni];
      f.rlist_off = a.rptr[ni] - 1;
      f.num_a = a.nptr[ni + 1] - a.nptr[ni];
      f.amap_off = a.nptr[ni] - 1;
      f.parent = a.sparent[ni] - 1;
      a.maxfront = std::max(a.maxfront, f.nrow);
      SymbolicFront& p = a.fronts[f.parent];
      p.least_desc = std::min(p.least_desc, f.least_desc);
   }
   for (int ni = 0; ni < nn; ++ni) {
      SymbolicFront& p = a.fronts[std::min(a.fronts[ni].parent, nn)];
      a.fronts[ni].next_child = p.first_child;
      p.first_child = ni;
   }
   a.nfactor = 0;
   for (int ni = 0; ni < nn; ++ni)
      a.nfactor += (int64_t)a.fronts[ni].nrow * a.fronts[ni].ncol;
   a.fronts[nn].parent = -1;
}

// ---------------------------------------------------------------------------
// Full analysis.
// ---------------------------------------------------------------------------
std::vector<double> clean_values(const Analysis& a, const double* val) {
   const int64_t nz = a.n > 0 ? a.ptr[a.n] - 1 : 0;
   if (!a.has_conv_map) return std::vector<double>(val, val + nz);
   std::vector<double> clean(nz, 0.0);
   for (size_t k = 0; k < a.conv_src.size(); ++k) clean[a.conv_dst[k]] += val[a.conv_src[k]];
   return clean;
}

int analyse(int n, int* order, const int64_t* ptr, const int* row, const double* val, bool check,
            int ordering, int nemin, Analysis& a) {
   a = Analysis();
   a.n = n;
   a.nemin = nemin < 1 ? 1 : nemin;
   if (n < 0) return kErrANOOR;
   if (n > 0 && (ptr == nullptr || row == nullptr)) return kErrAPtr;
   int flag = 0;
   // ---- input check and cleaning (lower CSC, 1-based) ----
   if (n > 0) {
      if (ptr[0] < 1 || ptr[n] < ptr[0]) return kErrAPtr;
      for (int j = 0; j < n; ++j)
         if (ptr[j + 1] < ptr[j]) return kErrAPtr;
   }
   int64_t unnz = n > 0 ? ptr[n] - 1 : 0;
   a.user_nnz = unnz;
   {
      // Gather valid entries, flagging out-of-range (incl. upper triangle)
      // and duplicates; duplicates are summed via the conversion map.
      std::vector<std::vector<std::pair<int, int64_t>>> cols(n);
      int64_t noor = 0, ndup = 0, nvalid = 0;
      for (int j = 0; j < n; ++j)
         for (int64_t p = ptr[j] - 1; p < ptr[j + 1] - 1; ++p) {
            int i = row[p];
            if (check && (i < j + 1 || i > n)) {
               ++noor;
               continue;
            }
            if (!check && (i < 1 || i > n)) return kErrAPtr;
            cols[j].push_back({i, p});
            ++nvalid;
         }
      if (n > 0 && nvalid == 0 && unnz > 0 && check) return kErrAAllOOR;
      a.ptr.assign(n + 1, 1);
      a.row.clear();
      std::vector<int64_t> csrc, cdst;
      bool changed = noor > 0;
      int nmissing = 0;
      for (int j = 0; j < n; ++j) {
         auto& c = cols[j];
         if (check)
            std::stable_sort(c.begin(), c.end(),
                             [](auto& x, auto& y) { return x.first < y.first; });
         bool hasdiag = false;
         for (size_t k = 0; k < c.size(); ++k) {
            if (c[k].first == j + 1) hasdiag = true;
            if (check && k > 0 && c[k].first == c[k - 1].first) {
               ++ndup;
               changed = true;
               csrc.push_back(c[k].second);
               cdst.push_back((int64_t)a.row.size() - 1);
               continue;
            }
            csrc.push_back(c[k].second);
            cdst.push_back((int64_t)a.row.size());
            a.row.push_back(c[k].first);
         }
         if (!hasdiag) ++nmissing;
         a.ptr[j + 1] = (int64_t)a.row.size() + 1;
      }
      if (check) {
         // cleaning may reorder entries inside a column
         for (size_t k = 0; k < csrc.size(); ++k)
            if (csrc[k] != cdst[k]) changed = true;
      }
      a.has_conv_map = changed;
      if (changed) {
         a.conv_src = std::move(csrc);
         a.conv_dst = std::move(cdst);
      }
      a.matrix_outrange = (int)noor;
      a.matrix_dup = (int)ndup;
      a.matrix_missing_diag = nmissing;
      if (check) {
         bool oor = noor > 0, dup = ndup > 0, miss = nmissing > 0;
         if (miss && (oor || dup))
            flag = kWarnMissDiagOORDup;
         else if (miss)
            flag = kWarnMissingDiag;
         else if (oor && dup)
            flag = kWarnDupAndOOR;
         else if (oor)
            flag = kWarnIdxOOR;
         else if (dup)
            flag = kWarnDupIdx;
      }
   }
   const int64_t* cptr = a.ptr.data();
   const int* crow = a.row.data();
   int64_t nnz = n > 0 ? cptr[n] - 1 : 0;

   // ---- ordering ----
   std::vector<int> ord(n);
   if (ordering == 0) {
      if (order == nullptr) return kErrOrder;
      std::vector<char> seen(n, 0);
      for (int i = 0; i < n; ++i) {
         int o = order[i];
         if (o < 1 || o > n || seen[o - 1]) return kErrOrder;
         seen[o - 1] = 1;
         ord[i] = o;
      }
   } else if (ordering == 2) {
      // matching-based ordering (src/spldlt_analyse_mod.F90:789-820); the
      // matching's scaling is kept for spldlt_factorize with scaling = 3
      if (n > 0 && val == nullptr) return kErrVal;
      if (n > 0) {
         const std::vector<double> cv = clean_values(a, val);
         a.scaling.assign(n, 1.0);
         if (match_order(n, cptr, crow, cv.data(), ord.data(), a.scaling.data()) == 1 && flag == 0)
            flag = kWarnAnalSingular;
      }
   } else {
      if (n > 0) nested_dissection(n, cptr, crow, ord.data());
   }

   // ---- pattern of P A P^T (lower, 0-based) in pivot numbering ----
   // For each pivot column pj, rows pi > pj.
   std::vector<int64_t> lptr(n + 1, 0);
   std::vector<int> lrow;
   {
      std::vector<int64_t> cnt(n + 1, 0);
      for (int j = 0; j < n; ++j)
         for (int64_t p = cptr[j] - 1; p < cptr[j + 1] - 1; ++p) {
            int i = crow[p] - 1;
            if (i == j) continue;
            int pi = ord[i] - 1, pj = ord[j] - 1;
            cnt[std::min(pi, pj) + 1]++;
         }
      for (int k = 0; k < n; ++k) cnt[k + 1] += cnt[k];
      lptr = cnt;
      lrow.resize(cnt[n]);
      std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
      for (int j = 0; j < n; ++j)
         for (int64_t p = cptr[j] - 1; p < cptr[j + 1] - 1; ++p) {
            int i = crow[p] - 1;
            if (i == j) continue;
            int pi = ord[i] - 1, pj = ord[j] - 1;
            lrow[pos[std::min(pi, pj)]++] = std::max(pi, pj);
         }
   }
   // row-wise view (for each row, the columns left of the diagonal)
   auto row_view = [n](const std::vector<int64_t>& cp, const std::vector<int>& cr,
                       std::vector<int64_t>& rp, std::vector<int>& rc) {
      rp.assign(n + 1, 0);
      for (int64_t p = 0; p < (int64_t)cr.size(); ++p) rp[cr[p] + 1]++;
      for (int k = 0; k < n; ++k) rp[k + 1] += rp[k];
      rc.resize(cr.size());
      std::vector<int64_t> pos(rp.begin(), rp.end() - 1);
      for (int j = 0; j < n; ++j)
         for (int64_t p = cp[j]; p < cp[j + 1]; ++p) rc[pos[cr[p]]++] = j;
   };
   std::vector<int64_t> rp;
   std::vector<int> rc;
   row_view(lptr, lrow, rp, rc);
   std::vector<int> parent = etree_from_rows(n, rp, rc);
   std::vector<int> post = postorder(n, parent);
   // relabel by postorder
   std::vector<int> pinv(n);  // old pivot -> postordered pivot
   for (int k = 0; k < n; ++k) pinv[post[k]] = k;
   std::vector<int> par2(n);
   for (int k = 0; k < n; ++k) {
      int p = parent[post[k]];
      par2[k] = (p == -1) ? -1 : pinv[p];
   }
   std::vector<int64_t> lptr2(n + 1, 0);
   std::vector<int> lrow2(lrow.size());
   {
      for (int k = 0; k < n; ++k) {
         int j = post[k];
         lptr2[k + 1] = lptr2[k] + (lptr[j + 1] - lptr[j]);
      }
      for (int k = 0; k < n; ++k) {
         int j = post[k];
         int64_t at = lptr2[k];
         for (int64_t p = lptr[j]; p < lptr[j + 1]; ++p) lrow2[at++] = pinv[lrow[p]];
      }
   }
   std::vector<int> cc = column_counts(n, par2, lptr2, lrow2);

   // ---- supernodes: fundamental chains plus nemin amalgamation ----
   // Processed in postorder; the supernode holding column j has root j when j
   // is reached and is merged into the supernode of its parent column if the
   // merge adds no fill to a single-column parent, or if both have fewer than
   // nemin pivots.
   std::vector<int> sn_of(n);           // supernode id (= its root column) of each column
   std::vector<int> nelim(n, 1);        // pivots per supernode id
   std::vector<int> members_head(n), members_tail(n), members_next(n, -1);
   std::vector<int> nchild_col(n, 0);
   std::vector<double> zcount(n, 0.0);  // explicit zeros accumulated by relaxed merges
   for (int j = 0; j < n; ++j) {
      sn_of[j] = j;
      members_head[j] = j;
      members_tail[j] = j;
   }
   for (int j = 0; j < n; ++j)
      if (par2[j] != -1) nchild_col[par2[j]]++;
   std::vector<int> rep(n);  // union-find representative (root column)
   for (int j = 0; j < n; ++j) rep[j] = j;
   auto find = [&rep](int x) {
      int r = x;
      while (rep[r] != r) r = rep[r];
      while (rep[x] != r) {
         int nx = rep[x];
         rep[x] = r;
         x = nx;
      }
      return r;
   };
   // Lower-trapezoid entries of a front with nc pivots and nr rows.
   auto trap = [](double nc, double nr) { return nc * (nc + 1) / 2 + nc * (nr - nc); };
   for (int j = 0; j < n; ++j) {
      int p = par2[j];
      if (p == -1) continue;
      int s = find(j);      // == j
      int ps = find(p);     // == p (p not yet processed as a child)
      bool merge;
      if (nelim[ps] == 1 && nchild_col[p] == 1 && cc[j] - 1 == cc[p]) {
         merge = true;  // fundamental: no fill
      } else if (nelim[ps] < a.nemin && nelim[s] < a.nemin) {
         merge = true;  // amalgamation of small nodes
      } else {
         // Relaxed merge: accept when the explicit zeros the merge adds
         // (counted cumulatively) are a small fraction of the merged front.
         double ns = nelim[s], np = nelim[ps];
         double ms = ns + cc[j] - 1, mp = np + cc[p] - 1;
         double merged = trap(ns + np, ns + mp);
         double zeros = merged - trap(ns, ms) - trap(np, mp) + zcount[s] + zcount[ps];
         merge = zeros <= kRelaxFill * merged;
         if (merge) zcount[ps] = zeros - zcount[s];
      }
      if (!merge) continue;
      // members: s's columns first, then ps's
      members_next[members_tail[s]] = members_head[ps];
      members_head[ps] = members_head[s];
      nelim[ps] += nelim[s];
      zcount[ps] += zcount[s];
      rep[s] = ps;
   }
   // supernode tree: ids are root columns
   std::vector<int> sroots;
   for (int j = 0; j < n; ++j)
      if (find(j) == j) sroots.push_back(j);
   int nn = (int)sroots.size();
   std::vector<int> sid(n, -1);
   for (int k = 0; k < nn; ++k) sid[sroots[k]] = k;
   std::vector<int> spar(nn, -1);
   for (int k = 0; k < nn; ++k) {
      int p = par2[sroots[k]];
      spar[k] = (p == -1) ? -1 : sid[find(p)];
   }
   std::vector<int> spost = postorder(nn, spar);
   std::vector<int> sinv(nn);
   for (int k = 0; k < nn; ++k) sinv[spost[k]] = k;
   // new pivot numbering: supernodes in postorder, member columns in order
   std::vector<int> newcol(n);  // postordered column -> final pivot (0-based)
   a.nnodes = nn;
   a.sptr.assign(nn + 1, 1);
   a.sparent.assign(nn, nn + 1);
   {
      int next = 0;
      for (int k = 0; k < nn; ++k) {
         int root = sroots[spost[k]];
         a.sptr[k] = next + 1;
         std::vector<int> cols;
         for (int c = members_head[root]; c != -1; c = members_next[c]) {
            cols.push_back(c);
            if (c == root) break;
         }
         std::sort(cols.begin(), cols.end());
         for (int c : cols) newcol[c] = next++;
         int sp = spar[spost[k]];
         a.sparent[k] = (sp == -1) ? nn + 1 : sinv[sp] + 1;
      }
      a.sptr[nn] = next + 1;
   }
   // final order: variable v -> pivot newcol[pinv[ord[v]-1]]
   a.order.resize(n);
   a.invp.resize(n);
   for (int v = 0; v < n; ++v) {
      int fp = newcol[pinv[ord[v] - 1]];
      a.order[v] = fp + 1;
      a.invp[fp] = v + 1;
   }
   if (order != nullptr)
      for (int v = 0; v < n; ++v) order[v] = a.order[v];

   // ---- row lists: own pivots, then the union of A rows and children rows ----
   {
      // A-lower pattern in final pivot numbering, by column (rows > col)
      std::vector<int64_t> fptr(n + 1, 0);
      std::vector<int> frow;
      {
         std::vector<int64_t> cnt(n + 1, 0);
         for (int j = 0; j < n; ++j)
            for (int64_t p = cptr[j] - 1; p < cptr[j + 1] - 1; ++p) {
               int i = crow[p] - 1;
               if (i == j) continue;
               int pi = a.order[i] - 1, pj = a.order[j] - 1;
               cnt[std::min(pi, pj) + 1]++;
            }
         for (int k = 0; k < n; ++k) cnt[k + 1] += cnt[k];
         fptr = cnt;
         frow.resize(cnt[n]);
         std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
         for (int j = 0; j < n; ++j)
            for (int64_t p = cptr[j] - 1; p < cptr[j + 1] - 1; ++p) {
               int i = crow[p] - 1;
               if (i == j) continue;
               int pi = a.order[i] - 1, pj = a.order[j] - 1;
               frow[pos[std::min(pi, pj)]++] = std::max(pi, pj);
            }
      }
      std::vector<std::vector<int>> below(nn);  // rows below the node's pivots
      std::vector<int> marker(n, -1);
      std::vector<int> chead(nn, -1), cnext(nn, -1);
      for (int k = nn - 1; k >= 0; --k) {
         int p = a.sparent[k] - 1;
         if (p >= nn) continue;
         cnext[k] = chead[p];
         chead[p] = k;
      }
      a.rptr.assign(nn + 1, 1);
      a.rlist.clear();
      for (int k = 0; k < nn; ++k) {
         int f = a.sptr[k] - 1, l = a.sptr[k + 1] - 1;  // [f, l)
         std::vector<int> rows;
         for (int c = f; c < l; ++c)
            for (int64_t p = fptr[c]; p < fptr[c + 1]; ++p) {
               int r = frow[p];
               if (r >= l && marker[r] != k) {
                  marker[r] = k;
                  rows.push_back(r);
               }
            }
         for (int c = chead[k]; c != -1; c = cnext[c]) {
            for (int r : below[c])
               if (r >= l && marker[r] != k) {
                  marker[r] = k;
                  rows.push_back(r);
               }
            std::vector<int>().swap(below[c]);
         }
         std::sort(rows.begin(), rows.end());
         for (int c = f; c < l; ++c) a.rlist.push_back(c + 1);
         for (int r : rows) a.rlist.push_back(r + 1);
         a.rptr[k + 1] = (int64_t)a.rlist.size() + 1;
         below[k] = std::move(rows);
      }
   }

   // ---- build_map and symbolic tree ----
   a.nptr.assign(nn + 1, 1);
   a.nlist.assign(2 * (size_t)std::max<int64_t>(nnz, 1), 0);
   build_map(n, cptr, crow, a.order.data(), a.invp.data(), nn, a.sptr.data(),
             a.rptr.data(), a.rlist.data(), a.nptr.data(), a.nlist.data());
   a.nlist.resize(2 * (size_t)(a.nptr[nn] - 1));
   build_symbolic_tree(a);

   // ---- statistics ----
   a.num_factor = 0;
   a.num_flops = 0;
   std::vector<int> depth(nn + 1, 0);
   a.maxdepth = 0;
   for (int k = nn - 1; k >= 0; --k) {
      int p = a.sparent[k] - 1;
      depth[k] = (p >= nn) ? 1 : depth[p] + 1;
      a.maxdepth = std::max(a.maxdepth, depth[k]);
   }
   for (int k = 0; k < nn; ++k) {
      int64_t m = a.rptr[k + 1] - a.rptr[k];
      int64_t c = a.sptr[k + 1] - a.sptr[k];
      a.num_factor += c * (c + 1) / 2 + c * (m - c);
      a.num_flops += node_flops(m, c);
   }
   // structural rank: variables with no entry at all are structurally singular
   {
      std::vector<char> touched(n, 0);
      for (int j = 0; j < n; ++j)
         for (int64_t p = cptr[j] - 1; p < cptr[j + 1] - 1; ++p) {
            touched[j] = 1;
            touched[crow[p] - 1] = 1;
         }
      int rank = 0;
      for (int j = 0; j < n; ++j) rank += touched[j];
      if (rank < n && flag >= 0 && flag != kWarnMissDiagOORDup) flag = kWarnAnalSingular;
   }
   return flag;
}

}  // namespace sylver_amd
