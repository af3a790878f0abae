// Device numeric phase of the indefinite path: APP LDL^T with delayed pivots
// on one GPU.
//
// The reference factors fronts in postorder with StarPU tasks
// (src/NumericTree.hxx:60-180, factor_front_indef_app + factor_front_indef_failed,
// src/factor_indef.hxx:1338-1930, src/factor_failed.hxx:29-157). Delayed
// columns change the size of every ancestor front, so the level schedule is
// driven from the host one level at a time: the host learns the number of
// delayed columns of a level (one small read-back) before it lays out the
// fronts of the next level. Per level:
//   zero -> scatter A (rows shifted by the delays) -> delayed columns of the
//   children -> extend-add (one launch per child rank) -> panel
//   factorization (one workgroup per front, k_app_front) -> read back nelim
//   -> L D into a workspace -> MFMA GEMM update of the contribution blocks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "kernels.hpp"
#include "kernels_indef.hpp"
#include "numeric.hpp"
#include "numeric_impl.hpp"

namespace sylver_amd {

struct IndefState {
   int nnodes = 0;
   std::vector<IndefDesc> h;
   std::vector<std::vector<int>> levels;
   std::vector<int> height;
   IndefDesc* d_nodes = nullptr;
   double* arena = nullptr;
   int64_t arena_cap = 0;  // workspace of one level's fronts
   int* d_vars = nullptr;
   int64_t vars_cap = 0;
   int* d_pres = nullptr;
   int64_t pres_cap = 0;
   double* d_dinv = nullptr;
   int64_t dinv_cap = 0;
   double* d_ws = nullptr;
   int64_t ws_cap = 0;
   int* d_list = nullptr;
   int64_t list_cap = 0;
   GemmTask* d_tasks = nullptr;
   int64_t tasks_cap = 0;
   int* d_status = nullptr;
   int* d_out = nullptr;
   int64_t out_cap = 0;
   double* d_xw = nullptr;
   int64_t xw_cap = 0;
   hipEvent_t ev[2] = {nullptr, nullptr};  // look-ahead of the multi-workgroup panel
   double* lstore = nullptr;  // L of every front (m x nelim, ld m)
   int64_t lstore_cap = 0;
   double* cbs = nullptr;     // trailing blocks of up to kCbBig entries, until the parent assembled them
   int64_t cbs_cap = 0;
   double* d_spart = nullptr;  // block sums of the blocked backward solve
   int64_t spart_cap = 0;
   // own allocations of large trailing blocks, reused (no free and
   // re-allocation of tens of GB per factorization): free list and the
   // capacity of each node's block
   std::vector<std::pair<int64_t, double*>> cb_pool;
   std::vector<int64_t> cb_cap;

   double* cb_acquire(int64_t need, int64_t& cap) {
      int best = -1;
      for (int i = 0; i < (int)cb_pool.size(); ++i)
         if (cb_pool[i].first >= need && cb_pool[i].first <= 2 * need &&
             (best < 0 || cb_pool[i].first < cb_pool[best].first))
            best = i;
      if (best >= 0) {
         double* p = cb_pool[best].second;
         cap = cb_pool[best].first;
         cb_pool.erase(cb_pool.begin() + best);
         return p;
      }
      cap = need + need / 8;
      return dalloc<double>(cap);
   }
   void cb_release(double*& p, int64_t cap) {
      if (p) cb_pool.push_back({cap, p});
      p = nullptr;
   }
   int64_t x_rows = 0;  // sum of 2 m over nodes
   // statistics of the last factorization
   int num_delay = 0, num_neg = 0, num_two = 0, num_zero = 0, maxfront = 0;
   int not_first_pass = 0, not_second_pass = 0;
   int64_t num_factor = 0;
   double flops = 0;

   ~IndefState() {
      for (void* p : {(void*)d_nodes, (void*)arena, (void*)d_vars, (void*)d_pres, (void*)d_dinv,
                      (void*)d_ws, (void*)d_list, (void*)d_tasks, (void*)d_status, (void*)d_out,
                      (void*)d_xw})
         dfree(p);
      for (hipEvent_t e : ev)
         if (e) (void)hipEventDestroy(e);
      dfree(lstore);
      dfree(cbs);
      dfree(d_spart);
      for (IndefDesc& d : h)
         if (d.cb_big) dfree(d.cb_big);
      for (auto& pc : cb_pool) dfree(pc.second);
   }
};

void destroy_indef(IndefState* s) { delete s; }

namespace {

inline unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kThreads - 1) / kThreads); }
// Grid of a grid-stride kernel over n items: a dispatch covers at most 2^32
// work-items (the level arena of a large tree holds more entries than that).
inline unsigned grid_stride(int64_t n) { return (unsigned)std::min<int64_t>(grid_of(n), int64_t(1) << 16); }

// Grow a device buffer to at least `need` elements, keeping `keep` elements.
template <typename T>
void grow(T*& p, int64_t& cap, int64_t need, int64_t keep, hipStream_t s) {
   if (need <= cap) return;
   int64_t nc = std::max<int64_t>(need, cap + cap / 2);
   T* q = dalloc<T>(nc);
   if (p && keep > 0) HIPCHECK(hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
   if (p) {
      HIPCHECK(hipStreamSynchronize(s));
      dfree(p);
   }
   p = q;
   cap = nc;
}

int64_t ws_elems(int m, int n) { return app_ws_elems(m, n); }

// Trailing blocks of at least this many entries get their own allocation,
// freed once the parent has assembled them.
constexpr int64_t kCbBig = int64_t(1) << 22;

// Multi-workgroup APP panel of one large front (k_lp_* launches, see
// kernels_indef.hpp): per block of kAppNB columns diag -> permutation ->
// apply + threshold -> finish -> update, then permute_failed and the second
// pass. Everything stays on the device; the host only enqueues.
void large_panel(IndefState& S, int node, int slot, double u, double small, int action, int failed_tpp,
                 hipStream_t s, hipStream_t s2) {
   const IndefDesc& d = S.h[node];
   const int m = d.m, n = d.n;
   auto grid_cap = [](int64_t items, int64_t cap) {
      return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cap, (items + kThreads - 1) / kThreads));
   };
   const int nt = (n + 63) / 64;
   k_lp_start<<<(unsigned)std::min<int64_t>((int64_t)nt * (nt + 1) / 2 + 1, 2048), kThreads, 0, s>>>(
       node, S.d_nodes, S.arena, S.d_dinv, S.d_ws);
   // Look-ahead: after block k's finish, the update of block k+1's diagonal
   // block runs first and block k+1's pivoting follows on the main stream,
   // while the rest of block k's update runs on the second stream; block
   // k+1's permutation waits for it. D^{-1} of the block is double-buffered.
   for (hipEvent_t& e : S.ev)
      if (!e) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
   k_lp_diag<<<1, kThreads, 0, s>>>(node, 0, 0, S.d_nodes, S.arena, S.d_ws, small, action);
   for (int c0 = 0, k = 0; c0 < n; c0 += kAppNB, ++k) {
      const int kb = std::min(kAppNB, n - c0), par = k & 1;
      const bool next = c0 + kb < n;
      if (k > 0) HIPCHECK(hipStreamWaitEvent(s, S.ev[1], 0));
      k_lp_perm<<<1 + grid_cap((int64_t)m + n, 1024), kThreads, 0, s>>>(node, c0, S.d_nodes, S.arena, S.d_vars,
                                                                       S.d_pres, S.d_ws);
      k_lp_apply<<<grid_cap((int64_t)m - kb, 1024), kThreads, 0, s>>>(node, c0, par, S.d_nodes, S.arena, S.d_ws,
                                                                       u, small);
      k_lp_finish<<<1, kThreads, 0, s>>>(node, c0, par, S.d_nodes, S.arena, S.d_dinv, S.d_ws);
      HIPCHECK(hipEventRecord(S.ev[0], s));
      if (next) {
         k_lp_update<<<3, kThreads, 0, s>>>(node, c0, par, 1, S.d_nodes, S.arena, S.d_ws);
         k_lp_diag<<<1, kThreads, 0, s>>>(node, c0 + kb, par ^ 1, S.d_nodes, S.arena, S.d_ws, small, action);
      }
      // tiles of the largest possible index space (every earlier column failed)
      const int64_t tr = (m - c0 + kLpTile - 1) / kLpTile + (c0 + kLpTile - 1) / kLpTile;
      const int64_t tiles = tr * (tr + 1) / 2;
      HIPCHECK(hipStreamWaitEvent(s2, S.ev[0], 0));
      k_lp_update<<<(unsigned)std::min<int64_t>(tiles, 1024), kThreads, 0, s2>>>(node, c0, par, next ? 2 : 0,
                                                                                  S.d_nodes, S.arena, S.d_ws);
      HIPCHECK(hipEventRecord(S.ev[1], s2));
   }
   HIPCHECK(hipStreamWaitEvent(s, S.ev[1], 0));
   // the second pass needs the first pass's outcome on the host (one read-back)
   int st[16];
   const int64_t st_off = d.ws_off + (int64_t)m * (kAppNB + n + 5) + 8 + kAppNB * kAppNB + 2 * kAppNB;
   HIPCHECK(hipMemcpyAsync(st, reinterpret_cast<int*>(S.d_ws + st_off), sizeof(st), hipMemcpyDeviceToHost, s));
   HIPCHECK(hipStreamSynchronize(s));
   const int nelim1 = st[kStNelim];
   const bool failed = !st[kStBad] && nelim1 < n;
   const bool allowed = failed && (m == n || failed_tpp);
   if (failed) {
      k_lp_np<<<1, kThreads, 0, s>>>(node, S.d_nodes, S.d_ws);
      const int64_t mn = (int64_t)m * n;
      k_lp_pgather<<<grid_cap(mn, 4096), kThreads, 0, s>>>(node, S.d_nodes, S.arena, S.d_vars, S.d_pres,
                                                           S.d_dinv, S.d_ws);
      k_lp_pscatter<<<grid_cap(mn, 4096), kThreads, 0, s>>>(node, S.d_nodes, S.arena, S.d_vars, S.d_pres,
                                                            S.d_dinv, S.d_ws);
   }
   k_tpp_start<<<1, 64, 0, s>>>(node, S.d_nodes, S.d_ws, allowed ? 1 : 0);
   if (allowed) {
      // statistics of the failed columns, then one pivot per step (at most
      // one step per column; steps after the last pivot return at once)
      const int nf = n - nelim1;
      k_tpp_update<<<nf, kThreads, 0, s>>>(node, S.d_nodes, S.arena, S.d_ws);
      for (int k = 0; k < nf; ++k) {
         k_tpp_decide<<<1, kThreads, 0, s>>>(node, S.d_nodes, S.arena, S.d_vars, S.d_pres, S.d_dinv, S.d_ws, u,
                                             small, action);
         k_tpp_update<<<std::max(1, nf - k - 1), kThreads, 0, s>>>(node, S.d_nodes, S.arena, S.d_ws);
      }
   }
   k_lp_tail<<<1, 64, 0, s>>>(node, slot, S.d_nodes, S.d_ws, S.d_status, S.d_out);
}

}  // namespace

void factor_indef(const Analysis& a, const DevicePlan* P, DeviceFactors* F, const sylver_options_t& opts,
                  FactorResult& res) {
   hipStream_t s = F->stream;
   if (!F->ind) F->ind = new IndefState();
   IndefState& S = *F->ind;
   const int nn = a.nnodes;
   S.nnodes = nn;
   for (int i = 0; i < (int)S.h.size(); ++i)  // left from an earlier call that stopped early
      if (S.h[i].cb_big) S.cb_release(S.h[i].cb_big, S.cb_cap[i]);
   S.h.assign(nn, IndefDesc{});
   S.cb_cap.assign(nn, 0);
   S.height.assign(nn, 0);
   for (int i = 0; i < nn; ++i) {
      const SymbolicFront& f = a.fronts[i];
      IndefDesc& d = S.h[i];
      d.ncol = f.ncol;
      d.nrow = f.nrow;
      d.parent = f.parent;
      d.first_child = f.first_child;
      d.next_child = f.next_child;
      d.rlist_off = f.rlist_off;
      d.amap_off = f.amap_off;
      d.num_a = f.num_a;
      d.cmap_off = P->hnodes[i].cmap_off;
      if (f.parent < nn) S.height[f.parent] = std::max(S.height[f.parent], S.height[i] + 1);
   }
   int nlev = 0;
   for (int i = 0; i < nn; ++i) nlev = std::max(nlev, S.height[i] + 1);
   S.levels.assign(nlev, {});
   for (int i = 0; i < nn; ++i) S.levels[S.height[i]].push_back(i);
   if (!S.d_nodes) S.d_nodes = dalloc<IndefDesc>(std::max(nn, 1));
   if (!S.d_status) S.d_status = dalloc<int>(std::max(nn, 1));
   HIPCHECK(hipMemsetAsync(S.d_status, 0, std::max(nn, 1) * sizeof(int), s));
   // initial capacities from the symbolic sizes; grown when delays need more
   {
      int64_t vr = 0, pr = 0, lf = 0;
      for (int i = 0; i < nn; ++i) {
         const int m = S.h[i].nrow;
         vr += m;
         pr += S.h[i].ncol;
         lf += (int64_t)m * S.h[i].ncol;
      }
      grow(S.lstore, S.lstore_cap, lf + lf / 4 + 64, 0, s);
      grow(S.d_vars, S.vars_cap, vr + vr / 4 + 64, 0, s);
      grow(S.d_pres, S.pres_cap, pr + pr / 4 + 64, 0, s);
      grow(S.d_dinv, S.dinv_cap, 2 * (pr + pr / 4 + 64), 0, s);
   }
   const double u = opts.u, small = opts.small;
   const int action = opts.action ? 1 : 0;
   // FP32 tier: contribution-block updates on FP32 MFMA (operands rounded to
   // FP32, FP64 fronts); pivot tests, D and its inverse stay FP64
   const bool low = F->precision != SYLVER_AMD_PREC_FP64;
   // fronts on the multi-workgroup panel: at least panel_min fully-summed
   // columns; in auto mode also fronts of 2^18 entries or more (many
   // rows, where one workgroup's pass over all rows per pivot is slow)
   const int panel_opt = ext_options().indef_panel_min;
   const int panel_min = panel_opt > 0 ? panel_opt : 256;
   auto is_large = [&](const IndefDesc& d) {
      return d.n >= panel_min || (panel_opt <= 0 && (int64_t)d.m * d.n >= (int64_t(1) << 18));
   };
   const int failed_tpp = (opts.failed_pivot_method == 1 || opts.pivot_method == 3) ? 1 : 0;
   int64_t vars_used = 0, pres_used = 0, dinv_used = 0, x_used = 0, lf_used = 0, cb_used = 0;
   S.num_factor = 0;
   S.maxfront = 0;
   S.flops = 0;
   for (int L = 0; L < nlev; ++L) {
      const std::vector<int>& lv = S.levels[L];
      // sizes of this level's fronts from the children's delays
      int64_t lv_arena = 0, lv_ws = 0;
      const int64_t vars_before = vars_used, pres_before = pres_used, dinv_before = dinv_used;
      for (int i : lv) {
         IndefDesc& d = S.h[i];
         int ndin = 0, dcol = d.ncol;
         for (int c = d.first_child; c != -1; c = S.h[c].next_child) {
            IndefDesc& cd = S.h[c];
            cd.delay_col0 = dcol;
            dcol += cd.n - cd.nelim;
            ndin += cd.n - cd.nelim;
         }
         d.ndelay_in = ndin;
         d.m = d.nrow + ndin;
         d.n = d.ncol + ndin;
         d.ld = (d.m + 7) / 8 * 8;
         d.off = lv_arena;  // the level's fronts share one workspace
         lv_arena += (int64_t)d.ld * d.m;
         d.var_off = vars_used;
         vars_used += d.m;
         d.pre_off = pres_used;
         pres_used += d.n;
         d.d_off = dinv_used;
         dinv_used += 2 * (int64_t)d.n;
         d.ws_off = lv_ws;
         lv_ws += ws_elems(d.m, d.n);
         d.x_off = x_used;
         x_used += 2 * (int64_t)d.m;
         d.nelim = 0;
         d.nelim_first = 0;
         S.maxfront = std::max(S.maxfront, d.m);
      }
      grow(S.arena, S.arena_cap, lv_arena, 0, s);
      grow(S.d_vars, S.vars_cap, vars_used, vars_before, s);
      grow(S.d_pres, S.pres_cap, pres_used, pres_before, s);
      grow(S.d_dinv, S.dinv_cap, dinv_used, dinv_before, s);
      grow(S.d_ws, S.ws_cap, lv_ws, 0, s);
      // lists: level nodes, children with delays, children by rank
      std::vector<int> lists(lv.begin(), lv.end());
      const int64_t o_delay = (int64_t)lists.size();
      for (int i : lv)
         for (int c = S.h[i].first_child; c != -1; c = S.h[c].next_child)
            if (S.h[c].n > S.h[c].nelim) lists.push_back(c);
      const int n_delay = (int)(lists.size() - o_delay);
      std::vector<std::pair<int64_t, int>> ranks;  // (offset, count)
      for (int r = 0;; ++r) {
         const int64_t o = (int64_t)lists.size();
         for (int i : lv) {
            int c = S.h[i].first_child;
            for (int k = 0; k < r && c != -1; ++k) c = S.h[c].next_child;
            if (c != -1 && S.h[c].m > S.h[c].n) lists.push_back(c);
         }
         const int cnt = (int)(lists.size() - o);
         bool any = false;
         for (int i : lv) {
            int c = S.h[i].first_child;
            for (int k = 0; k < r && c != -1; ++k) c = S.h[c].next_child;
            if (c != -1) any = true;
         }
         if (!any) break;
         if (cnt > 0) ranks.push_back({o, cnt});
      }
      grow(S.d_list, S.list_cap, (int64_t)lists.size(), 0, s);
      grow(S.d_out, S.out_cap, 2 * (int64_t)lv.size(), 0, s);
      HIPCHECK(hipMemcpyAsync(S.d_list, lists.data(), lists.size() * sizeof(int), hipMemcpyHostToDevice, s));
      HIPCHECK(hipMemcpyAsync(S.d_nodes, S.h.data(), nn * sizeof(IndefDesc), hipMemcpyHostToDevice, s));
      const int nl = (int)lv.size();
      // zero the level's fronts (contiguous)
      k_fill<<<grid_stride(lv_arena), kThreads, 0, s>>>(lv_arena, 0.0, S.arena);
      k_ind_init<<<nl, kThreads, 0, s>>>(S.d_list, S.d_nodes, P->d_nlist, P->d_rlist, F->d_aval,
                                         F->has_scaling ? F->d_scaling : nullptr, S.arena, S.d_vars, S.d_pres);
      if (n_delay > 0) {
         int maxd = 1;
         for (int k = 0; k < n_delay; ++k) {
            const IndefDesc& c = S.h[lists[o_delay + k]];
            maxd = std::max(maxd, c.n - c.nelim);
         }
         dim3 g((unsigned)std::min(maxd, 512), (unsigned)n_delay);
         k_ind_delays<<<g, kThreads, 0, s>>>(S.d_list + o_delay, S.d_nodes, P->d_cmap, S.arena, S.d_vars,
                                             S.cbs);
      }
      for (auto& rk : ranks) {
         int maxq = 1;
         for (int k = 0; k < rk.second; ++k) {
            const IndefDesc& c = S.h[lists[rk.first + k]];
            maxq = std::max(maxq, c.m - c.n);
         }
         dim3 g((unsigned)std::min(std::max(maxq / 8, 64), 1024), rk.second);
         k_ind_ea<<<g, kThreads, 0, s>>>(S.d_list + rk.first, S.d_nodes, P->d_cmap, S.arena, S.cbs);
      }
      // panels: one workgroup per front for the small ones (batched), the
      // multi-workgroup sequence for each large one; out slots small first
      std::vector<int> slot(nl);
      int nsmall = 0;
      for (int k = 0; k < nl; ++k)
         if (!is_large(S.h[lv[k]])) slot[k] = nsmall++;
      {
         int j = nsmall;
         for (int k = 0; k < nl; ++k)
            if (is_large(S.h[lv[k]])) slot[k] = j++;
      }
      if (nsmall > 0) {
         std::vector<int> sl;
         for (int k = 0; k < nl; ++k)
            if (!is_large(S.h[lv[k]])) sl.push_back(lv[k]);
         const int64_t o_small = (int64_t)lists.size();
         grow(S.d_list, S.list_cap, o_small + nsmall, o_small, s);
         HIPCHECK(hipMemcpyAsync(S.d_list + o_small, sl.data(), sl.size() * sizeof(int), hipMemcpyHostToDevice, s));
         k_app_front<<<nsmall, kThreads, 0, s>>>(S.d_list + o_small, S.d_nodes, S.arena, S.d_vars, S.d_pres,
                                                 S.d_dinv, S.d_ws, u, small, action, failed_tpp, S.d_status,
                                                 S.d_out);
      }
      for (int k = 0; k < nl; ++k) {
         if (!is_large(S.h[lv[k]])) continue;
         large_panel(S, lv[k], slot[k], u, small, action, failed_tpp, s, F->stream2);
      }
      std::vector<int> outs(2 * nl), out(2 * nl);
      HIPCHECK(hipMemcpyAsync(outs.data(), S.d_out, outs.size() * sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
      for (int k = 0; k < nl; ++k) {
         out[2 * k] = outs[2 * slot[k]];
         out[2 * k + 1] = outs[2 * slot[k] + 1];
      }
      // the children's own trailing blocks are assembled (the read-back synchronized)
      for (int i : lv)
         for (int c = S.h[i].first_child; c != -1; c = S.h[c].next_child)
            if (S.h[c].cb_big) S.cb_release(S.h[c].cb_big, S.cb_cap[c]);
      // contribution blocks: L D into the workspace, then the MFMA update
      std::vector<int> cb_nodes;
      std::vector<GemmTask> tasks;
      for (int k = 0; k < nl; ++k) {
         IndefDesc& d = S.h[lv[k]];
         d.nelim = out[2 * k];
         d.nelim_first = out[2 * k + 1];
         S.num_factor += (int64_t)d.m * d.nelim;
         for (int64_t r = d.m - d.nelim + 1; r <= d.m; ++r) S.flops += (double)r * r;
         const int q = d.m - d.n;
         if (q <= 0 || d.nelim <= 0) continue;
         cb_nodes.push_back(lv[k]);
         const int nt = (q + 127) / 128;
         for (int tj = 0; tj < nt; ++tj)
            for (int ti = tj; ti < nt; ++ti) {
               GemmTask t{};
               if (low) {  // FP32 operands in the workspace (float offsets)
                  const int ldq = ind32_ldq(q);
                  const int64_t base = ind32_base(d.ws_off);
                  t.a_off = base + (int64_t)ldq * d.nelim + (int64_t)ti * 128;
                  t.b_off = base + (int64_t)tj * 128;
                  t.lda = ldq;
                  t.ldb = ldq;
               } else {
                  t.a_off = d.off + d.n + (int64_t)ti * 128;
                  t.b_off = d.ws_off + (int64_t)tj * 128;
                  t.lda = d.ld;
                  t.ldb = q;
               }
               t.c_off = d.off + (d.n + (int64_t)ti * 128) + (d.n + (int64_t)tj * 128) * d.ld;
               t.ldc = d.ld;
               t.M = std::min(128, q - ti * 128);
               t.N = std::min(128, q - tj * 128);
               t.K = d.nelim;
               tasks.push_back(t);
            }
      }
      if (!cb_nodes.empty()) {
         while (tasks.size() % 8) {
            GemmTask t{};
            t.M = 0;
            tasks.push_back(t);
         }
         grow(S.d_list, S.list_cap, (int64_t)cb_nodes.size(), 0, s);
         grow(S.d_tasks, S.tasks_cap, (int64_t)tasks.size(), 0, s);
         HIPCHECK(hipMemcpyAsync(S.d_list, cb_nodes.data(), cb_nodes.size() * sizeof(int), hipMemcpyHostToDevice, s));
         HIPCHECK(hipMemcpyAsync(S.d_tasks, tasks.data(), tasks.size() * sizeof(GemmTask), hipMemcpyHostToDevice, s));
         HIPCHECK(hipMemcpyAsync(S.d_nodes, S.h.data(), nn * sizeof(IndefDesc), hipMemcpyHostToDevice, s));
         dim3 g(64, (unsigned)cb_nodes.size());
         if (low) {
            k_ind_ld32<<<g, kThreads, 0, s>>>(S.d_list, S.d_nodes, S.arena, S.d_dinv, S.d_ws);
            k_update_cb32<<<(unsigned)tasks.size(), 2 * kThreads, 0, s>>>(S.d_tasks, reinterpret_cast<const float*>(S.d_ws),
                                                                      S.arena);
         } else {
            k_ind_ld<<<g, kThreads, 0, s>>>(S.d_list, S.d_nodes, S.arena, S.d_dinv, S.d_ws);
            k_update<double, 2, 128, 128, 2 * kThreads><<<(unsigned)tasks.size(), 2 * kThreads, 0, s>>>(S.d_tasks, S.arena,
                                                                                                       S.d_ws);
         }
         HIPCHECK(hipStreamSynchronize(s));  // the task and list buffers are rewritten next level
      }
      // L into the factor store, trailing blocks into the contribution store
      const int64_t lf_before = lf_used, cb_before = cb_used;
      for (int i : lv) {
         IndefDesc& d = S.h[i];
         d.lf_off = lf_used;
         lf_used += (int64_t)d.m * d.nelim;
         d.cb_off = 0;
         d.cb_big = nullptr;
         const int64_t t = d.m - d.nelim;
         if (d.parent < nn && t > 0) {
            if (t * t >= kCbBig) {
               d.cb_big = S.cb_acquire(t * t, S.cb_cap[i]);
            } else {
               d.cb_off = cb_used;
               cb_used += t * t;
            }
         }
      }
      grow(S.lstore, S.lstore_cap, std::max<int64_t>(lf_used, 1), lf_before, s);
      grow(S.cbs, S.cbs_cap, std::max<int64_t>(cb_used, 1), cb_before, s);
      grow(S.d_list, S.list_cap, nl, 0, s);
      HIPCHECK(hipMemcpyAsync(S.d_list, lv.data(), nl * sizeof(int), hipMemcpyHostToDevice, s));
      HIPCHECK(hipMemcpyAsync(S.d_nodes, S.h.data(), nn * sizeof(IndefDesc), hipMemcpyHostToDevice, s));
      {
         int64_t big = 1;
         for (int i : lv) big = std::max<int64_t>(big, (int64_t)S.h[i].m * S.h[i].m);
         dim3 g((unsigned)std::min<int64_t>((big + kThreads - 1) / kThreads, 2048), (unsigned)nl);
         k_ind_store<<<g, kThreads, 0, s>>>(S.d_list, S.d_nodes, S.arena, S.lstore, S.cbs, nn);
      }
      // a launch the runtime refused (grid or resources) must not pass silently
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipStreamSynchronize(s));  // list and node buffers are rewritten next level
   }
   S.x_rows = x_used;
   HIPCHECK(hipMemcpyAsync(S.d_nodes, S.h.data(), nn * sizeof(IndefDesc), hipMemcpyHostToDevice, s));
   // flags and statistics (src/NumericTree.hxx:150-179, src/factor_failed.hxx:55-150)
   std::vector<int> status(std::max(nn, 1));
   std::vector<double> dv(std::max<int64_t>(dinv_used, 1));
   HIPCHECK(hipMemcpyAsync(status.data(), S.d_status, nn * sizeof(int), hipMemcpyDeviceToHost, s));
   if (dinv_used > 0)
      HIPCHECK(hipMemcpyAsync(dv.data(), S.d_dinv, dinv_used * sizeof(double), hipMemcpyDeviceToHost, s));
   HIPCHECK(hipStreamSynchronize(s));
   S.num_delay = S.num_neg = S.num_two = S.num_zero = 0;
   S.not_first_pass = S.not_second_pass = 0;
   bool singular = false;
   for (int i = 0; i < nn; ++i) {
      const IndefDesc& d = S.h[i];
      if (status[i] != 0) singular = true;
      S.num_delay += d.n - d.nelim;
      if (d.nelim_first < d.n) S.not_first_pass += d.n - d.nelim_first;
      if (failed_tpp || d.m == d.n) S.not_second_pass += d.n - d.nelim;
      if (d.parent >= nn && d.nelim < d.n) singular = true;
      const double* dd = dv.data() + d.d_off;
      for (int k = 0; k < d.nelim;) {
         const double a11 = dd[2 * k], a21 = dd[2 * k + 1];
         if (k + 1 == d.nelim || std::isfinite(dd[2 * k + 2])) {
            if (a11 == 0.0) ++S.num_zero;
            if (a11 < 0.0) ++S.num_neg;
            ++k;
         } else {
            const double a22 = dd[2 * k + 3];
            ++S.num_two;
            const double det = a11 * a22 - a21 * a21, tr = a11 + a22;
            if (det < 0)
               ++S.num_neg;
            else if (tr < 0)
               S.num_neg += 2;
            k += 2;
         }
      }
   }
   res.num_delay = S.num_delay;
   res.num_neg = S.num_neg;
   res.num_two = S.num_two;
   res.num_zero = S.num_zero;
   res.not_first_pass = S.not_first_pass;
   res.not_second_pass = S.not_second_pass;
   res.matrix_rank = a.n - S.num_zero;
   res.maxfront = S.maxfront;
   res.num_factor = S.num_factor;
   res.num_flops = S.flops;
   if (singular)
      res.flag = SYLVER_ERROR_SINGULAR;
   else if (S.num_zero > 0)
      res.flag = SYLVER_WARNING_FACT_SINGULAR;
}

// Solves with the indefinite factors: job 0 full, 1 fwd, 2 diag, 3 bwd,
// 4 diag+bwd. b and sol are n x nrhs in pivot order on the device.
void solve_indef(const DevicePlan* P, DeviceFactors* F, int job, const double* b, double* sol, int nrhs,
                 hipStream_t s) {
   IndefState& S = *F->ind;
   const int n = P->n;
   const int64_t nn = (int64_t)n * nrhs;
   grow(S.d_xw, S.xw_cap, std::max<int64_t>(S.x_rows * nrhs, 1), 0, s);
   int64_t lcap = 0;
   for (auto& lv : S.levels) lcap += (int64_t)lv.size();
   grow(S.d_list, S.list_cap, lcap, 0, s);
   std::vector<int> all;
   std::vector<int64_t> loff;
   for (auto& lv : S.levels) {
      loff.push_back((int64_t)all.size());
      all.insert(all.end(), lv.begin(), lv.end());
   }
   HIPCHECK(hipMemcpyAsync(S.d_list, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice, s));
   const bool fwd = job == 0 || job == 1, diag = job == 0 || job == 2 || job == 4,
              bwd = job == 0 || job == 3 || job == 4;
   int maxm = 1;
   for (const IndefDesc& d : S.h) maxm = std::max(maxm, d.m);
   grow(S.d_spart, S.spart_cap, (int64_t)((maxm + kThreads - 1) / kThreads + 1) * nrhs * 64, 0, s);
   // fronts with many eliminated columns or a large L: blocked elimination
   auto blocked = [&](int L, bool forward) {
      for (int i : S.levels[L]) {
         const IndefDesc& d = S.h[i];
         if (!isv_blocked(d.m, d.nelim)) continue;
         if (forward) {
            for (int j0 = 0; j0 < d.nelim; j0 += 64) {
               const int jb = std::min(64, d.nelim - j0);
               k_isv_fdiag<<<1, kThreads, 0, s>>>(i, j0, S.d_nodes, S.lstore, S.d_xw, nrhs);
               if (d.m > j0 + jb)
                  k_isv_fupd<<<grid_of(d.m - j0 - jb), kThreads, 0, s>>>(i, j0, S.d_nodes, S.lstore, S.d_xw, nrhs);
            }
         } else {
            for (int j0 = (d.nelim - 1) / 64 * 64; j0 >= 0; j0 -= 64) {
               const int jb = std::min(64, d.nelim - j0);
               const int rows = d.m - j0 - jb, nch = rows > 0 ? (int)grid_of(rows) : 0;
               if (nch > 0)
                  k_isv_bdot<<<nch, kThreads, 0, s>>>(i, j0, S.d_nodes, S.lstore, S.d_xw, S.d_spart, nrhs);
               k_isv_bdiag<<<1, kThreads, 0, s>>>(i, j0, nch, S.d_nodes, S.lstore, S.d_xw, S.d_spart, nrhs);
            }
         }
         k_isv_out<<<grid_of((int64_t)d.nelim * nrhs), kThreads, 0, s>>>(i, S.d_nodes, S.d_vars, S.d_xw, sol, n,
                                                                        nrhs);
      }
   };
   if (fwd) {
      for (size_t L = 0; L < S.levels.size(); ++L) {
         k_ind_fwd<<<(unsigned)S.levels[L].size(), kThreads, 0, s>>>(
             S.d_list + loff[L], S.d_nodes, S.lstore, S.d_vars, S.d_pres, P->d_cmap, b, n, S.d_xw, sol, nrhs);
         blocked((int)L, true);
      }
   } else {
      k_copy<<<grid_stride(nn), kThreads, 0, s>>>(nn, b, sol);
   }
   if (diag)
      k_ind_diag<<<std::max(S.nnodes, 1), kThreads, 0, s>>>(S.d_nodes, S.nnodes, S.d_vars, S.d_dinv, sol, n,
                                                            nrhs);
   if (bwd)
      for (int L = (int)S.levels.size() - 1; L >= 0; --L) {
         k_ind_bwd<<<(unsigned)S.levels[L].size(), kThreads, 0, s>>>(S.d_list + loff[L], S.d_nodes, S.lstore,
                                                                     S.d_vars, S.d_xw, sol, n, nrhs);
         blocked(L, false);
      }
}

int indef_stats(const DeviceFactors* F, int* out) {
   if (!F || !F->ind) return SYLVER_ERROR_CALL_SEQUENCE;
   const IndefState& S = *F->ind;
   out[0] = S.num_delay;
   out[1] = S.num_neg;
   out[2] = S.num_two;
   out[3] = S.num_zero;
   out[4] = S.maxfront;
   out[5] = S.not_first_pass;
   out[6] = S.not_second_pass;
   out[7] = 0;
   return 0;
}

// Read-back of one node of the indefinite factors: dims {m, n, nelim,
// ndelay_in}, the front rows' variables, L (m x nelim, ld m) and D^{-1}.
int indef_node(const DeviceFactors* F, int node, int* dims, int* rows, double* L, double* dinv) {
   if (!F || !F->ind || node < 0 || node >= F->ind->nnodes) return SYLVER_ERROR_CALL_SEQUENCE;
   const IndefState& S = *F->ind;
   const IndefDesc& d = S.h[node];
   try {
      if (dims) {
         dims[0] = d.m;
         dims[1] = d.n;
         dims[2] = d.nelim;
         dims[3] = d.ndelay_in;
      }
      if (rows && d.m > 0) HIPCHECK(hipMemcpy(rows, S.d_vars + d.var_off, d.m * sizeof(int), hipMemcpyDeviceToHost));
      if (L && d.nelim > 0)
         HIPCHECK(hipMemcpy2D(L, (size_t)d.m * sizeof(double), S.lstore + d.lf_off, (size_t)d.m * sizeof(double),
                              (size_t)d.m * sizeof(double), d.nelim, hipMemcpyDeviceToHost));
      if (dinv && d.nelim > 0)
         HIPCHECK(hipMemcpy(dinv, S.d_dinv + d.d_off, 2 * d.nelim * sizeof(double), hipMemcpyDeviceToHost));
      return 0;
   } catch (const HipError&) {
      return SYLVER_ERROR_CUDA_UNKNOWN;
   }
}

}  // namespace sylver_amd
