// Device numeric phase: the multifrontal factorization as a device-resident
// level schedule on one GPU, its solves and FP64 refinement.
//
// The reference walks the assembly tree in postorder on the CPU and submits
// per-tile StarPU tasks (src/NumericTreePosdef.hxx:133-353). Here the tree is
// cut into levels by height (leaves are level 0); every level runs as a fixed
// sequence of batched launches over all its fronts:
//   zero -> scatter A -> extend-add (one launch per child rank)
//   -> one-workgroup factorization of small fronts
//   -> per block column of large fronts: diagonal POTRF+TRTRI, TRSM as GEMM,
//      MFMA Schur update of the trailing block (contribution block included).
// The launch list and all task descriptors are built once per pattern and
// live in HBM; the list is captured in a hipGraph and replayed per factorize.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "kernels.hpp"
#include "numeric.hpp"
#include "numeric_impl.hpp"

namespace sylver_amd {


sylver_amd_options_t& ext_options() {
   static sylver_amd_options_t o = [] {
      sylver_amd_options_t d;
      std::memset(&d, 0, sizeof(d));
      d.precision = SYLVER_AMD_PREC_FP64;
      d.bf16_min_ncol = 1024;
      d.refine = SYLVER_AMD_REFINE_AUTO;
      d.refine_tol = 1e-15;
      d.refine_max_iter = 30;
      d.use_graph = 1;
      d.small_assembly = 0;
      d.strip_trsm = 1;
      d.level_assembly = 1;
      return d;
   }();
   return o;
}

// Plan-shaping switches of the extension options for a factorization in
// precision prec (strip_trsm 1, the default: the strip TRSM for every
// precision -- measured at 80^3: FP64 1.4 ms, FP32 1.0 ms faster than the
// TRSM by the inverse; 2: FP64 factors only).
PlanKnobs plan_knobs(int prec) {
   const sylver_amd_options_t& o = ext_options();
   const bool strip = o.strip_trsm == 1 || (o.strip_trsm == 2 && prec == SYLVER_AMD_PREC_FP64);
   const int es = prec == SYLVER_AMD_PREC_FP64 ? 8 : 4;  // bytes per factor entry
   return PlanKnobs{o.small_assembly != 0, strip, o.level_assembly != 0, kAsmLdsBytes / es};
}

int device_count() {
   int c = 0;
   if (hipGetDeviceCount(&c) != hipSuccess) return 0;
   return c;
}

// ---------------------------------------------------------------------------
// Plan: pattern-dependent device state.
// ---------------------------------------------------------------------------
void destroy_factors(DeviceFactors* f) { delete f; }

void device_sync() {
   if (device_count() > 0) (void)hipDeviceSynchronize();
}

namespace {

// Block columns of at most this many large fronts (one level's) take the
// strip TRSM; beyond it (batched fronts) each front's chain is short next to
// the launch's width.
constexpr int kStepMaxFronts = 64;

// The plan's host side: layout, maps, task lists and launch programs. No
// HIP call (it also runs, for its checks, on a host without a device).
DevicePlan* build_plan_host(const Analysis& a, int small_max, int wide_ncol, int num_cus, const PlanKnobs& k) {
   std::unique_ptr<DevicePlan> P(new DevicePlan());
   P->n = a.n;
   P->nnodes = a.nnodes;
   P->small_max = small_max;
   P->wide_ncol = wide_ncol;
   // fronts at fixed caller offsets (the batched entry points) hold their
   // values already: nothing to assemble
   P->fuse_small = k.fuse_small && a.fixed_off.empty();
   P->strip_trsm = k.strip_trsm;
   P->level_asm = k.level_asm;
   P->asm_elems = k.asm_elems;
   P->variant_gen = a.variant_gen;
   const int nn = a.nnodes;
   // one-workgroup kernel or tiled path for front i
   const bool have_variant = (int)a.variant.size() == nn;
   auto one_wg = [&](int i) {
      if (have_variant && a.variant[i] >= 0) return a.variant[i] == 1;
      return a.fronts[i].nrow <= small_max;
   };
   P->hnodes.resize(nn);
   std::vector<int> height(nn + 1, 0);
   for (int i = 0; i < nn; ++i) {
      const SymbolicFront& f = a.fronts[i];
      NodeDesc& d = P->hnodes[i];
      d.m = f.nrow;
      d.n = f.ncol;
      d.parent = f.parent;
      d.ld = (int)align_up(f.nrow, 8);  // 8-aligned columns for 16-byte loads
      d.rlist_off = f.rlist_off;
      d.amap_off = f.amap_off;
      d.num_a = f.num_a;
      d.first_child = f.first_child;
      d.next_child = f.next_child;
      if (f.parent < nn) height[f.parent] = std::max(height[f.parent], height[i] + 1);
      P->num_flops += node_flops(f.nrow, f.ncol);
   }
   if (P->level_asm) {
      // each front's A-map sorted by position (the column runs of L_ASM)
      P->hasort.resize(2 * std::max<int64_t>((int64_t)a.nlist.size() / 2, 1));
      std::vector<int64_t> idx;
      for (int i = 0; i < nn; ++i) {
         const SymbolicFront& f = a.fronts[i];
         const int64_t* src = a.nlist.data() + 2 * f.amap_off;
         idx.resize(f.num_a);
         for (int64_t p = 0; p < f.num_a; ++p) idx[p] = p;
         std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) { return src[2 * x + 1] < src[2 * y + 1]; });
         for (int64_t p = 0; p < f.num_a; ++p) {
            P->hasort[2 * (f.amap_off + p)] = src[2 * idx[p]];
            P->hasort[2 * (f.amap_off + p) + 1] = src[2 * idx[p] + 1];
         }
      }
   }
   int nlev = 0;
   for (int i = 0; i < nn; ++i) nlev = std::max(nlev, height[i] + 1);
   P->nlevels = nlev;
   P->levels.assign(nlev, {});
   for (int i = 0; i < nn; ++i) P->levels[height[i]].push_back(i);
   // arena, dinv and solve offsets: level by level
   int64_t off = 0, doff = 0, soff = 0;
   for (int L = 0; L < nlev; ++L)
      for (int i : P->levels[L]) {
         NodeDesc& d = P->hnodes[i];
         d.off = off;
         off += align_up((int64_t)d.m * d.ld, 64);
         d.dinv_off = doff;
         if (d.n > 0) {
            int64_t nblk = (d.n + kNB - 1) / kNB;
            int64_t last = d.n - (nblk - 1) * kNB;
            doff += align_up((nblk - 1) * kNB * kNB + last * last, 64);
         }
         d.solve_off = soff;
         soff += d.m;
      }
   P->arena_elems = off;
   if ((int)a.fixed_off.size() == nn && (int)a.fixed_ld.size() == nn) {
      int64_t end = 0;
      for (int i = 0; i < nn; ++i) {
         NodeDesc& d = P->hnodes[i];
         d.off = a.fixed_off[i];
         d.ld = a.fixed_ld[i];
         end = std::max(end, d.off + (int64_t)d.ld * d.m);
      }
      P->arena_elems = end;
   }
   P->dinv_elems = doff;
   P->solve_rows = soff;
   // child maps: position of each CB row in the parent front
   // (reference src/NumericTreePosdef.hxx:219-238)
   P->hcmap_off.assign(nn + 1, 0);
   for (int i = 0; i < nn; ++i) P->hcmap_off[i + 1] = P->hcmap_off[i] + (P->hnodes[i].m - P->hnodes[i].n);
   P->hcmap.assign(P->hcmap_off[nn], 0);
   {
      std::vector<int> map(a.n + 1, 0);
      for (int p = 0; p < nn; ++p) {
         const SymbolicFront& fp = a.fronts[p];
         if (fp.first_child == -1) continue;
         const int* rl = &a.rlist[fp.rlist_off];
         for (int i = 0; i < fp.ncol; ++i) map[rl[i]] = i;
         for (int i = fp.ncol; i < fp.nrow; ++i) map[rl[i]] = i;  // no delays on this path
         for (int c = fp.first_child; c != -1; c = a.fronts[c].next_child) {
            const SymbolicFront& fc = a.fronts[c];
            const int* rc = &a.rlist[fc.rlist_off];
            int q = fc.nrow - fc.ncol;
            for (int k = 0; k < q; ++k) P->hcmap[P->hcmap_off[c] + k] = map[rc[fc.ncol + k]];
         }
      }
   }
   for (int i = 0; i < nn; ++i) P->hnodes[i].cmap_off = P->hcmap_off[i];

   // ---- factor program ----
   std::vector<DiagTask> step_dinv;  // diagonal blocks factored by L_DFACT (inverses formed at the end)
   auto push = [](std::vector<Launch>& prog, int kind, int64_t off, int64_t count) {
      if (count > 0) {
         Launch l;
         l.kind = kind;
         l.count = (int)count;
         l.off = off;
         prog.push_back(l);
      }
   };
   int side_last = -1;  // last event recorded on the side stream, not yet waited for
   // an assembly run's original entries (the front's sorted A-map pairs
   // whose position falls in its columns) and its child column ranges
   auto asm_ranges = [&](AsmTask& t) {
      const NodeDesc& d = P->hnodes[t.node];
      const int64_t* am = P->hasort.data() + 2 * d.amap_off;
      auto first_at = [&](int64_t pos) {  // first pair with 1-based position >= pos
         int64_t lo = 0, hi = d.num_a;
         while (lo < hi) {
            const int64_t mid = (lo + hi) / 2;
            if (am[2 * mid + 1] < pos)
               lo = mid + 1;
            else
               hi = mid;
         }
         return lo;
      };
      t.p0 = first_at((int64_t)t.c0 * d.m + 1);
      t.p1 = first_at((int64_t)t.c1 * d.m + 1);
      t.g0 = (int)P->asmch.size();
      for (int c = d.first_child; c != -1; c = P->hnodes[c].next_child) {
         const NodeDesc& cd = P->hnodes[c];
         const int q = cd.m - cd.n;
         const int* cm = P->hcmap.data() + P->hcmap_off[c];
         const int j0 = (int)(std::lower_bound(cm, cm + q, t.c0) - cm);
         const int j1 = (int)(std::lower_bound(cm, cm + q, t.c1) - cm);
         if (j0 < j1) P->asmch.push_back(int4{c, j0, j1, 0});
      }
      t.gn = (int)P->asmch.size() - t.g0;
   };
   for (int L = 0; L < nlev; ++L) {
      const std::vector<int>& lv = P->levels[L];
      // one-workgroup fronts assembled inside their factor launch
      auto fused = [&](int i) { return P->fuse_small && P->hnodes[i].n > 0 && one_wg(i); };
      {
         int64_t f0 = P->nlistv.size();
         for (int i : lv)
            if (fused(i)) P->nlistv.push_back(i);
         push(P->fprog, L_SMALL_ASM, f0, P->nlistv.size() - f0);
      }
      if (P->level_asm) {
         // one launch: each task owns a run of one front's columns (zero, A
         // entries, children) of at most asm_elems lower-triangle entries,
         // assembled in LDS (pad = 1); a longer single column works on the
         // front directly (pad = 0)
         int64_t a0 = P->asmt.size();
         for (int i : lv) {
            if (fused(i)) continue;
            const NodeDesc& d = P->hnodes[i];
            for (int c = 0; c < d.m;) {
               int c1 = c;
               int64_t area = 0;
               while (c1 < d.m && (c1 == c || area + (d.m - c1) <= P->asm_elems)) area += d.m - c1++;
               AsmTask t{};
               t.node = i;
               t.c0 = c;
               t.c1 = c1;
               t.staged = area <= P->asm_elems ? 1 : 0;
               asm_ranges(t);
               P->asmt.push_back(t);
               c = c1;
            }
         }
         push(P->fprog, L_ASM, a0, P->asmt.size() - a0);
      } else {
         // zero
         int64_t z0 = P->zero.size();
         for (int i : lv) {
            if (fused(i)) continue;
            const NodeDesc& d = P->hnodes[i];
            int chunk = std::max(4, std::min(256, 16384 / std::max(d.m, 1)));
            for (int c = 0; c < d.m; c += chunk) P->zero.push_back(ColTask{i, c, std::min(d.m, c + chunk), 0});
         }
         push(P->fprog, L_ZERO, z0, P->zero.size() - z0);
         // scatter A
         int64_t i0 = P->init.size();
         for (int i : lv) {
            if (fused(i)) continue;
            const NodeDesc& d = P->hnodes[i];
            for (int64_t p = 0; p < d.num_a; p += 4096)
               P->init.push_back(InitTask{i, 0, p, std::min<int64_t>(d.num_a, p + 4096)});
         }
         push(P->fprog, L_INIT, i0, P->init.size() - i0);
         // extend-add by child rank
         for (int rank = 0;; ++rank) {
            int64_t e0 = P->ea.size();
            bool any = false;
            for (int i : lv) {
               int c = P->hnodes[i].first_child;
               for (int r = 0; r < rank && c != -1; ++r) c = P->hnodes[c].next_child;
               if (c == -1) continue;
               any = true;
               if (fused(i)) continue;
               const NodeDesc& dc = P->hnodes[c];
               int q = dc.m - dc.n;
               if (q <= 0) continue;
               int chunk = std::max(4, std::min(q, 16384 / std::max(q, 1)));
               for (int j = 0; j < q; j += chunk) P->ea.push_back(EaTask{c, j, std::min(q, j + chunk), 0});
            }
            if (!any) break;
            push(P->fprog, L_EA, e0, P->ea.size() - e0);
         }
      }
      // small fronts
      int64_t s0 = P->nlistv.size();
      std::vector<int> large;
      for (int i : lv) {
         const NodeDesc& d = P->hnodes[i];
         if (d.n == 0 || fused(i)) continue;
         if (one_wg(i))
            P->nlistv.push_back(i);
         else
            large.push_back(i);
      }
      push(P->fprog, L_SMALL, s0, P->nlistv.size() - s0);
      // large fronts, block column by block column
      int maxn = 0;
      for (int i : large) maxn = std::max(maxn, P->hnodes[i].n);
      for (int k0 = 0; k0 < maxn; k0 += kNB) {
         const bool first_of_pair = ((k0 / kNB) % 2) == 0;
         int nfr = 0;
         int64_t nwg = 0;
         for (int i : large) {
            const NodeDesc& d = P->hnodes[i];
            if (k0 >= d.n) continue;
            ++nfr;
            nwg += 1 + (d.m - k0 - std::min(kNB, d.n - k0) + 63) / 64;
         }
         // a latency-bound chain (few fronts, the strips resident on the
         // device at once, beside the far updates): the diagonal blocks
         // factored alone, then the panel TRSM by substitution in strips; else
         // the TRSM by the inverse (at 80^3, strips up to half the CUs against
         // up to all of them: FP64 factor 0.15 ms faster, FP32 the same)
         const bool step = P->strip_trsm && nfr <= kStepMaxFronts && nwg - nfr <= std::max(num_cus / 2, 1);
         if (step) {
            int64_t d0 = P->diag.size();
            for (int i : large) {
               const NodeDesc& d = P->hnodes[i];
               if (k0 >= d.n) continue;
               const DiagTask t{i, k0, std::min(kNB, d.n - k0), 0};
               P->diag.push_back(t);
               step_dinv.push_back(t);  // its inverse: with the small fronts' at the end
            }
            push(P->fprog, L_DFACT, d0, P->diag.size() - d0);
            int64_t s0 = P->strip.size();
            for (int i : large) {
               const NodeDesc& d = P->hnodes[i];
               if (k0 >= d.n) continue;
               const int kb = std::min(kNB, d.n - k0);
               for (int r = k0 + kb; r < d.m; r += 64) P->strip.push_back(StripTask{i, k0, kb, r});
            }
            push(P->fprog, L_STRIP, s0, P->strip.size() - s0);
         } else {
            int64_t d0 = P->diag.size();
            for (int i : large) {
               const NodeDesc& d = P->hnodes[i];
               if (k0 >= d.n) continue;
               P->diag.push_back(DiagTask{i, k0, std::min(kNB, d.n - k0), 0});
            }
            push(P->fprog, L_DIAG, d0, P->diag.size() - d0);
            int64_t t0 = P->gemm.size();
            for (int i : large) {
               const NodeDesc& d = P->hnodes[i];
               if (k0 >= d.n) continue;
               int kb = std::min(kNB, d.n - k0);
               for (int r = k0 + kb; r < d.m; r += 64) {  // 64-row tiles (k_update<T, 1, 64, 128>)
                  GemmTask g{};
                  g.a_off = d.off + r + (int64_t)k0 * d.ld;
                  g.b_off = d.dinv_off + (int64_t)(k0 / kNB) * kNB * kNB;
                  g.c_off = g.a_off;
                  g.lda = d.ld;
                  g.ldb = kb;
                  g.ldc = d.ld;
                  g.M = std::min(64, d.m - r);
                  g.N = kb;
                  g.K = kb;
                  P->gemm.push_back(g);
               }
            }
            while ((P->gemm.size() - t0) % 8 != 0 && P->gemm.size() > (size_t)t0) P->gemm.push_back(GemmTask{});
            push(P->fprog, L_TRSM, t0, P->gemm.size() - t0);
         }
         // Schur update. Block columns are paired: after the first of a pair
         // only the next panel's columns are updated (narrow update); after the
         // second, the trailing block receives both panels at once (K = 256),
         // halving the passes over the contribution block. That update is
         // split for look-ahead: the columns of the next pair (needed by the
         // next diagonal factorization) on the factor stream, the columns
         // beyond them on the side stream, where they overlap the next pair's
         // diagonal/TRSM chain.
         // lower-triangular TS x TS tiles of columns [c0, c1) (rows from c0 down)
         auto tiles = [&](const NodeDesc& d, int kstart, int K, int c0, int c1, int TS) {
            const int c0a = c0 & ~7;
            const int R = d.m - c0a;
            const int ntr = (R + TS - 1) / TS, ntc = (c1 - c0a + TS - 1) / TS;
            for (int tj = 0; tj < ntc; ++tj)
               for (int ti = tj; ti < ntr; ++ti) {
                  GemmTask g{};
                  g.a_off = d.off + c0a + ti * TS + (int64_t)kstart * d.ld;
                  g.b_off = d.off + c0a + tj * TS + (int64_t)kstart * d.ld;
                  g.c_off = d.off + (c0a + ti * TS) + (int64_t)(c0a + tj * TS) * d.ld;
                  g.lda = g.ldb = g.ldc = d.ld;
                  g.M = std::min(TS, R - ti * TS);
                  g.N = std::min(TS, c1 - (c0a + tj * TS));
                  g.K = K;
                  g.pad0 = (tj == 0) ? c0 - c0a : 0;
                  P->gemm.push_back(g);
               }
         };
         auto pad8 = [&](int64_t u0) {
            while ((P->gemm.size() - u0) % 8 != 0 && P->gemm.size() > (size_t)u0) P->gemm.push_back(GemmTask{});
         };
         // recorded on the factor stream once this block's panel is final (the
         // DIAG/TRSM launches just pushed are the last factor-stream entries)
         int ev_trsm = P->nevents++;
         P->fprog.back().rec_ev = ev_trsm;
         for (int part = 0; part < 2; ++part) {  // 0: next pair (factor stream), 1: far (side)
            for (int wide = 0; wide < 2; ++wide) {
               int64_t u0 = P->gemm.size();
               for (int i : large) {
                  const NodeDesc& d = P->hnodes[i];
                  if (k0 >= d.n) continue;
                  if ((d.n > wide_ncol) != (wide == 1)) continue;
                  const int kb = std::min(kNB, d.n - k0);
                  if (first_of_pair && k0 + kNB < d.n) {
                     if (part == 1) continue;
                     // the next panel's columns: 64 x 64 tiles (critical path)
                     const int k1 = k0 + kNB, kb1 = std::min(kNB, d.n - k1);
                     for (int cj = 0; cj < kb1; cj += 64)
                        for (int r = k1 + cj; r < d.m; r += 64) {
                           GemmTask g{};
                           g.a_off = d.off + r + (int64_t)k0 * d.ld;
                           g.b_off = d.off + k1 + cj + (int64_t)k0 * d.ld;
                           g.c_off = d.off + r + (int64_t)(k1 + cj) * d.ld;
                           g.lda = g.ldb = g.ldc = d.ld;
                           g.M = std::min(64, d.m - r);
                           g.N = std::min(64, kb1 - cj);
                           g.K = kb;
                           g.pad0 = 0;
                           P->gemm.push_back(g);
                        }
                     continue;
                  }
                  const int kstart = first_of_pair ? k0 : k0 - kNB;
                  const int K = k0 + kb - kstart;
                  const int r0 = k0 + kb;
                  if (d.m <= r0) continue;
                  const int nx = std::min(d.n, r0 + 2 * kNB);  // end of the next pair's columns
                  if (part == 0 && nx > r0) tiles(d, kstart, K, r0, nx, 64);  // critical path
                  if (part == 1 && d.m > nx) tiles(d, kstart, K, nx, d.m, 128);
               }
               pad8(u0);
               const int64_t cnt = P->gemm.size() - u0;
               if (cnt == 0) continue;
               push(P->fprog, wide ? L_UPDATE_WIDE : L_UPDATE, u0, cnt);
               Launch& L = P->fprog.back();
               L.tile = part == 0 ? 64 : 128;
               if (part == 0) {
                  // the next pair's columns were last touched by the side
                  // stream (the far update of the pair before); the narrow
                  // update after a pair's first block writes only that
                  // pair's second block, which the far update never touches
                  if (side_last >= 0 && !first_of_pair) {
                     L.wait_ev = side_last;
                     side_last = -1;
                  }
               } else {
                  L.stream = 1;
                  if (ev_trsm >= 0) {
                     L.wait_ev = ev_trsm;
                     ev_trsm = -1;
                  }
                  side_last = P->nevents++;
                  L.rec_ev = side_last;
               }
            }
         }
      }
      // join the side stream before the next level's assembly
      if (side_last >= 0) {
         Launch j;
         j.kind = L_SYNC;
         j.count = 0;
         j.off = 0;
         j.wait_ev = side_last;
         P->fprog.push_back(j);
         side_last = -1;
      }
   }
   // inverted diagonal blocks of the small fronts and of the fused block
   // steps (the separate DIAG launches write their own), for the solves;
   // fronts at caller offsets (the batched entry points) are factored, never
   // solved, so they go without
   if (a.fixed_off.empty()) {
      int64_t d0 = P->diag.size();
      P->diag.insert(P->diag.end(), step_dinv.begin(), step_dinv.end());
      for (int L = 0; L < nlev; ++L)
         for (int i : P->levels[L]) {
            const NodeDesc& d = P->hnodes[i];
            if (d.n == 0 || !one_wg(i)) continue;
            for (int k0 = 0; k0 < d.n; k0 += kNB)
               P->diag.push_back(DiagTask{i, k0, std::min(kNB, d.n - k0), 0});
         }
      push(P->fprog, L_DINV, d0, P->diag.size() - d0);
   }
   // algorithmic work of every factor launch (for the roofline figures)
   for (Launch& L : P->fprog) {
      for (int64_t t = L.off; t < L.off + L.count; ++t) {
         switch (L.kind) {
            case L_ZERO: {
               const ColTask& c = P->zero[t];
               const NodeDesc& d = P->hnodes[c.node];
               for (int col = c.c0; col < c.c1; ++col) L.words += d.m - col;
               break;
            }
            case L_INIT: {
               const InitTask& it = P->init[t];
               L.words += (double)(it.p1 - it.p0);
               L.bytes += 24.0 * (double)(it.p1 - it.p0);
               break;
            }
            case L_EA: {
               const EaTask& e = P->ea[t];
               const NodeDesc& d = P->hnodes[e.child];
               const double q = d.m - d.n;
               double cnt = 0;
               for (int j = e.j0; j < e.j1; ++j) cnt += q - j;
               L.words += 3.0 * cnt;  // read child, read+write parent
               L.bytes += 4.0 * cnt;  // row map
               break;
            }
            case L_ASM: {
               const AsmTask& c = P->asmt[t];
               const NodeDesc& d = P->hnodes[c.node];
               // a run staged in LDS writes its columns once and reads each
               // child entry once; a direct run zero-fills its column, then
               // reads the child and reads and writes the parent per entry
               const double per_child = c.staged != 0 ? 1.0 : 3.0;
               for (int col = c.c0; col < c.c1; ++col) L.words += d.m - col;
               for (int g = c.g0; g < c.g0 + c.gn; ++g) {
                  const int4 cr = P->asmch[g];
                  const int q = P->hnodes[cr.x].m - P->hnodes[cr.x].n;
                  for (int j = cr.y; j < cr.z; ++j) {
                     L.words += per_child * (q - j);
                     L.bytes += 4.0 * (q - j);  // row map
                  }
               }
               // original entries: the position pair and the FP64 value read
               // (a direct run also writes each into the front)
               L.bytes += 24.0 * (double)(c.p1 - c.p0);
               if (c.staged == 0) L.words += (double)(c.p1 - c.p0);
               break;
            }
            case L_SMALL: {
               const NodeDesc& d = P->hnodes[P->nlistv[t]];
               L.flops += (double)node_flops(d.m, d.n);
               L.words += (double)d.m * (d.m + 1);  // lower triangle read + written once
               break;
            }
            case L_SMALL_ASM: {
               const NodeDesc& d = P->hnodes[P->nlistv[t]];
               L.flops += (double)node_flops(d.m, d.n);
               // zero fill, then the factorization's read and write of the lower triangle
               L.words += 1.5 * (double)d.m * (d.m + 1);
               L.words += (double)d.num_a;
               L.bytes += 24.0 * (double)d.num_a;
               for (int c = d.first_child; c != -1; c = P->hnodes[c].next_child) {
                  const double q = P->hnodes[c].m - P->hnodes[c].n;
                  L.words += 3.0 * q * (q + 1) / 2;  // read child, read+write parent
                  L.bytes += 4.0 * q * (q + 1) / 2;  // row map
               }
               break;
            }
            case L_DIAG:
            case L_DINV: {
               const double kb = P->diag[t].kb;
               L.flops += (L.kind == L_DIAG ? 2.0 : 1.0) * kb * kb * kb / 3.0;
               L.words += 2.0 * kb * kb;
               break;
            }
            case L_DFACT: {
               const double kb = P->diag[t].kb;
               L.flops += kb * kb * kb / 3.0;
               L.words += 1.5 * kb * kb;
               break;
            }
            case L_STRIP: {
               const StripTask& c = P->strip[t];
               const double kb = c.kb;
               const double rows = std::min(64, P->hnodes[c.node].m - c.r0);
               L.flops += rows * kb * kb;
               L.words += 2.0 * rows * kb + 0.5 * kb * kb;
               break;
            }
            case L_TRSM: {
               const GemmTask& g = P->gemm[t];
               if (g.M <= 0) break;
               L.flops += (double)g.M * g.K * g.K;
               L.words += 2.0 * g.M * g.K + 0.5 * g.K * g.K;
               break;
            }
            case L_UPDATE:
            case L_UPDATE_WIDE: {
               const GemmTask& g = P->gemm[t];
               if (g.M <= 0) break;
               const int rows_skip = g.pad0;  // aligned-down rows of the first tile
               if (g.a_off == g.b_off) {
                  const double mm = g.M - rows_skip;
                  L.flops += mm * (mm + 1) * g.K;
               } else {
                  L.flops += 2.0 * g.M * (g.N - rows_skip) * g.K;
               }
               L.words += 2.0 * g.M * g.N + (double)(g.M + g.N) * g.K;
               break;
            }
            default:
               break;
         }
      }
   }

   // ---- solve programs ----
   // Forward: per level, gather (row chunks), then per block column a
   // diagonal multiply and the update of the rows below. Backward: per level
   // from the top, gather, the contribution-block transposed update, then per
   // block column from the last a diagonal multiply and the transposed update
   // of the earlier rows.
   constexpr int kRowChunk = 256;
   constexpr int kColChunk = 64;
   constexpr int kCbRowChunk = kNB;
   // fronts solved by one workgroup per sweep: one diagonal block, at most
   // 256 rows
   auto small_solve = [](const NodeDesc& d) { return d.n <= kNB && d.m <= 2 * kNB; };
   constexpr int kUpdRows = 64;   // rows per forward update task: one 16-row tile per wave
   // a small-front solve launch records its largest m and n (its LDS size)
   auto small_sizes = [&](std::vector<Launch>& prog, int kind, int64_t s0) {
      if (prog.empty() || prog.back().kind != kind || prog.back().off != s0) return;
      Launch& l = prog.back();
      for (int64_t k = s0; k < s0 + l.count; ++k) {
         l.smax_m = std::max(l.smax_m, P->hnodes[P->nlistv[k]].m);
         l.smax_n = std::max(l.smax_n, P->hnodes[P->nlistv[k]].n);
      }
   };
   // a step launch whose tasks all belong to one front gets it by value
   auto mark_single = [&](std::vector<Launch>& prog, int kind, int64_t t0) {
      if (prog.empty() || prog.back().kind != kind || prog.back().off != t0) return;
      Launch& l = prog.back();
      const int node = P->stask[t0].node;
      for (int64_t t = t0; t < t0 + l.count; ++t)
         if (P->stask[t].node != node) return;
      l.node1 = node;
   };
   // Level-wide chained solves (k_chain) unless SYLVER_AMD_SOLVE_STEPS=1
   // asks for one launch per 128-column step. flag0[i]: the first flag of
   // front i in the chained sweeps (-1: the front takes the step launches;
   // so do fronts whose factor columns span 2^31 bytes or more in FP64, as
   // the chained kernel addresses them with 32-bit offsets).
   const char* steps_env = std::getenv("SYLVER_AMD_SOLVE_STEPS");
   const bool chain = !(steps_env && steps_env[0] == '1');
   std::vector<int> flag0(P->hnodes.size(), -1);
   P->chain_flags = 0;
   if (chain)
      for (size_t i = 0; i < P->hnodes.size(); ++i) {
         const NodeDesc& d = P->hnodes[i];
         if (d.n == 0 || small_solve(d) || (int64_t)d.n * d.ld * 8 >= (int64_t(1) << 31)) continue;
         flag0[i] = (int)P->chain_flags;
         P->chain_flags += (d.n + kNB - 1) / kNB;
      }
   if (chain) {
      push(P->fwdprog, L_CLEAR, 0, P->chain_flags);
      push(P->bwdprog, L_CLEAR, 0, P->chain_flags);
   }
   // rows of the children's contribution blocks that land in [r, r1) of
   // front i, in child list order, appended to `out` as {child, i0, i1}
   auto child_ranges = [&](int i, int r, int r1, std::vector<int4>& out, int split) {
      for (int c = P->hnodes[i].first_child; c != -1; c = P->hnodes[c].next_child) {
         const NodeDesc& cd = P->hnodes[c];
         const int q = cd.m - cd.n;
         const int* cm = P->hcmap.data() + cd.cmap_off;
         const int i0 = (int)(std::lower_bound(cm, cm + q, r) - cm);
         const int i1 = (int)(std::lower_bound(cm, cm + q, r1) - cm);
         for (int a = i0; a < i1; a += split) out.push_back(int4{c, a, std::min(i1, a + split), 0});
      }
   };
   // A level's one-workgroup fronts and its larger fronts are independent
   // (no front of a level is an ancestor of another): the small-front launch
   // runs on the side stream beside the rest of the level, forked after the
   // level before and joined before the level after.
   auto overlap_small = [&](std::vector<Launch>& prog, size_t b0) {
      if (b0 == 0 || prog[b0 - 1].stream != 0) return;
      size_t si = prog.size();
      int others = 0;
      for (size_t k = b0; k < prog.size(); ++k) {
         if (prog[k].kind == L_FSMALL || prog[k].kind == L_BSMALL)
            si = k;
         else
            ++others;
      }
      if (si == prog.size() || others == 0) return;
      if (prog[b0 - 1].rec_ev < 0) prog[b0 - 1].rec_ev = P->nevents++;
      prog[si].stream = 1;
      prog[si].wait_ev = prog[b0 - 1].rec_ev;
      prog[si].rec_ev = P->nevents++;
      Launch j;
      j.kind = L_SYNC;
      j.count = 0;
      j.off = 0;
      j.wait_ev = prog[si].rec_ev;
      prog.push_back(j);
   };
   for (int L = 0; L < nlev; ++L) {
      const std::vector<int>& lv = P->levels[L];
      const size_t lb0 = P->fwdprog.size();
      int64_t g0 = P->stask.size();
      for (int i : lv) {
         const NodeDesc& d = P->hnodes[i];
         // small fronts and chained fronts gather in their own kernels
         if (small_solve(d) || flag0[i] >= 0) continue;
         for (int r = 0; r < d.m; r += kRowChunk) {
            const int r1 = std::min(d.m, r + kRowChunk);
            const int gbeg = (int)P->gch.size();
            child_ranges(i, r, r1, P->gch, kThreads);
            P->stask.push_back(SolveTask{i, gbeg, (int)P->gch.size() - gbeg, r, r1, 0});
         }
      }
      push(P->fwdprog, L_FGATHER, g0, P->stask.size() - g0);
      {
         const int64_t s0 = P->nlistv.size();
         for (int i : lv)
            if (small_solve(P->hnodes[i])) P->nlistv.push_back(i);
         push(P->fwdprog, L_FSMALL, s0, P->nlistv.size() - s0);
         small_sizes(P->fwdprog, L_FSMALL, s0);
      }
      if (chain) {
         // one launch per level: the diagonal blocks (a chain per front,
         // block order), then the contribution blocks, which wait only on
         // diagonal blocks of lower index (so the launch makes progress
         // whatever fits on the device; measured at 80^3 against a launch of
         // their own for levels wider than the device: FP32 refined solve
         // 0.56 ms, FP64 solve 0.45 ms faster)
         const int64_t q0 = P->ctask.size();
         for (int i : lv) {
            const NodeDesc& d = P->hnodes[i];
            if (flag0[i] < 0) continue;
            const int nb = (d.n + kNB - 1) / kNB;
            for (int j = 0; j < nb; ++j) {
               const int r0 = j * kNB, r1 = std::min(d.n, (j + 1) * kNB), gb = (int)P->cgch.size();
               child_ranges(i, r0, r1, P->cgch, 1 << 30);
               P->ctask.push_back(ChainTask{i, j, nb, r0, r1, flag0[i], 0, 0, gb, (int)P->cgch.size() - gb});
            }
         }
         for (int i : lv) {
            const NodeDesc& d = P->hnodes[i];
            if (flag0[i] < 0) continue;
            const int nb = (d.n + kNB - 1) / kNB;
            for (int r = d.n; r < d.m; r += kNB)
            {
               const int r1 = std::min(d.m, r + kNB), gb = (int)P->cgch.size();
               child_ranges(i, r, r1, P->cgch, 1 << 30);
               P->ctask.push_back(ChainTask{i, nb, nb, r, r1, flag0[i], 0, 0, gb, (int)P->cgch.size() - gb});
            }
         }
         push(P->fwdprog, L_FCHAIN, q0, P->ctask.size() - q0);
      }
      // fronts outside the chained launch: one launch per diagonal block
      int maxblk = 0;
      for (int i : lv)
         if (flag0[i] < 0 && !small_solve(P->hnodes[i])) maxblk = std::max(maxblk, (P->hnodes[i].n + kNB - 1) / kNB);
      // one launch per diagonal block: the diagonal solve (write task) and
      // the update of the rows below, every task forming y = inv(L_kk) x_k
      for (int sblk = 0; sblk < maxblk; ++sblk) {
         int64_t t0 = P->stask.size();
         for (int i : lv) {
            const NodeDesc& d = P->hnodes[i];
            const int k0 = sblk * kNB;
            if (k0 >= d.n) continue;
            const int kb = std::min(kNB, d.n - k0);
            if (small_solve(d) || flag0[i] >= 0) continue;  // small-front or chained launch
            P->stask.push_back(SolveTask{i, k0, kb, 0, 0, 1});
            for (int r = k0 + kb; r < d.m; r += kUpdRows)
               P->stask.push_back(SolveTask{i, k0, kb, r, std::min(d.m, r + kUpdRows), 0});
         }
         push(P->fwdprog, L_FDIAG, t0, P->stask.size() - t0);
         mark_single(P->fwdprog, L_FDIAG, t0);
      }
      overlap_small(P->fwdprog, lb0);
   }
   for (int L = nlev - 1; L >= 0; --L) {
      const std::vector<int>& lv = P->levels[L];
      const size_t lb0 = P->bwdprog.size();
      int64_t g0 = P->stask.size();
      for (int i : lv) {
         const NodeDesc& d = P->hnodes[i];
         if (small_solve(d) || flag0[i] >= 0) continue;  // these read the solution directly
         for (int r = 0; r < d.m; r += kRowChunk)
            P->stask.push_back(SolveTask{i, 0, 0, r, std::min(d.m, r + kRowChunk), 0});
      }
      push(P->bwdprog, L_BGATHER, g0, P->stask.size() - g0);
      {
         const int64_t s0 = P->nlistv.size();
         for (int i : lv)
            if (small_solve(P->hnodes[i])) P->nlistv.push_back(i);
         push(P->bwdprog, L_BSMALL, s0, P->nlistv.size() - s0);
         small_sizes(P->bwdprog, L_BSMALL, s0);
      }
      // contribution-block part: columns [0, n) against rows [n, m), split
      // into row chunks (partial sums per slot) and reduced in slot order
      // (chained fronts sum their slots at the start of their chained
      // launch; the others take a reduction launch)
      int64_t c0 = P->stask.size();
      std::vector<SolveTask> red;
      std::vector<int> cb0(lv.size(), 0), cbn(lv.size(), 0);
      int slots = 0;
      for (size_t li = 0; li < lv.size(); ++li) {
         const int i = lv[li];
         const NodeDesc& d = P->hnodes[i];
         if (d.m - d.n <= 0 || d.n == 0 || small_solve(d)) continue;
         cb0[li] = slots;
         cbn[li] = (d.m - d.n + kCbRowChunk - 1) / kCbRowChunk;
         for (int j = 0; j < d.n; j += kColChunk) {
            const int s0 = slots;
            for (int r = d.n; r < d.m; r += kCbRowChunk)
               P->stask.push_back(SolveTask{i, r, std::min(d.m, r + kCbRowChunk), j,
                                            std::min(d.n, j + kColChunk), slots++});
            if (flag0[i] < 0) red.push_back(SolveTask{i, s0, slots - s0, j, std::min(d.n, j + kColChunk), 0});
         }
      }
      P->cb_slots = std::max<int64_t>(P->cb_slots, slots);
      push(P->bwdprog, L_BUPD, c0, P->stask.size() - c0);
      int64_t r0 = P->stask.size();
      P->stask.insert(P->stask.end(), red.begin(), red.end());
      push(P->bwdprog, L_BRED, r0, P->stask.size() - r0);
      if (chain) {
         // one launch for the level: per front its diagonal blocks from the last
         const int64_t q0 = P->ctask.size();
         for (size_t li = 0; li < lv.size(); ++li) {
            const int i = lv[li];
            const NodeDesc& d = P->hnodes[i];
            if (flag0[i] < 0) continue;
            const int nb = (d.n + kNB - 1) / kNB;
            for (int j = nb - 1; j >= 0; --j)
               P->ctask.push_back(ChainTask{i, j, nb, j * kNB, std::min(d.n, (j + 1) * kNB), flag0[i], cb0[li],
                                            cbn[li], 0, 0});
         }
         push(P->bwdprog, L_BCHAIN, q0, P->ctask.size() - q0);
      }
      int maxblk = 0;
      for (int i : lv)
         if (flag0[i] < 0 && !small_solve(P->hnodes[i])) maxblk = std::max(maxblk, (P->hnodes[i].n + kNB - 1) / kNB);
      // one launch per diagonal block (from the last): the transposed
      // diagonal solve (write task) and the update of the earlier columns by
      // the block's rows, every task forming x_k' = inv(L_kk)^T x_k
      for (int sblk = maxblk - 1; sblk >= 0; --sblk) {
         int64_t t0 = P->stask.size();
         for (int i : lv) {
            const NodeDesc& d = P->hnodes[i];
            const int k0 = sblk * kNB;
            if (k0 >= d.n) continue;
            const int kb = std::min(kNB, d.n - k0);
            if (small_solve(d) || flag0[i] >= 0) continue;
            P->stask.push_back(SolveTask{i, k0, kb, 0, 0, 1});
            for (int j = 0; j < k0; j += kColChunk)
               P->stask.push_back(SolveTask{i, k0, kb, j, std::min(k0, j + kColChunk), 0});
         }
         push(P->bwdprog, L_BDIAG, t0, P->stask.size() - t0);
         mark_single(P->bwdprog, L_BDIAG, t0);
      }
      overlap_small(P->bwdprog, lb0);
   }

   // Every flag a chained launch waits for must be set by a workgroup of
   // lower index in the same launch (the launch's freedom from deadlock).
   for (const std::vector<Launch>* prog : {&P->fwdprog, &P->bwdprog})
      for (const Launch& L : *prog) {
         if (L.kind != L_FCHAIN && L.kind != L_BCHAIN) continue;
         const bool bwd = L.kind == L_BCHAIN;
         std::vector<int64_t> setter((size_t)P->chain_flags, -1);
         for (int64_t t = L.off; t < L.off + L.count; ++t) {
            const ChainTask& c = P->ctask[t];
            const int lo = bwd ? c.blk + 1 : 0, hi = bwd ? c.nb : std::min(c.blk, c.nb);
            for (int k = lo; k < hi; ++k)
               if (setter[(size_t)(c.flag0 + k)] < 0)
                  throw std::runtime_error("chained solve plan: a wait precedes its flag");
            if (c.blk < c.nb) {
               if (setter[(size_t)(c.flag0 + c.blk)] >= 0)
                  throw std::runtime_error("chained solve plan: a flag is set twice");
               setter[(size_t)(c.flag0 + c.blk)] = t;
            }
         }
      }

   // ---- CSR of the full matrix in pivot order ----
   // (the internal seam passes no matrix pattern: values are then addressed
   // through nlist only and no residual can be formed)
   if (a.ptr.empty()) {
      int64_t mx = 0;
      for (size_t k = 0; k + 1 < a.nlist.size(); k += 2) mx = std::max<int64_t>(mx, a.nlist[k]);
      P->clean_nnz = mx;
      P->csr_nnz = 0;
   } else {
      int n = a.n;
      const int64_t* cptr = a.ptr.data();
      const int* crow = a.row.data();
      P->clean_nnz = n > 0 ? cptr[n] - 1 : 0;
      std::vector<int64_t> rp(n + 1, 0);
      for (int j = 0; j < n; ++j)
         for (int64_t p = cptr[j] - 1; p < cptr[j + 1] - 1; ++p) {
            int i = crow[p] - 1;
            rp[a.order[i]]++;
            if (i != j) rp[a.order[j]]++;
         }
      for (int i = 0; i < n; ++i) rp[i + 1] += rp[i];
      std::vector<int> col(rp[n]);
      std::vector<int64_t> src(rp[n]);
      std::vector<int64_t> pos(rp.begin(), rp.end() - 1);
      for (int j = 0; j < n; ++j)
         for (int64_t p = cptr[j] - 1; p < cptr[j + 1] - 1; ++p) {
            int i = crow[p] - 1;
            int pi = a.order[i] - 1, pj = a.order[j] - 1;
            col[pos[pi]] = pj;
            src[pos[pi]++] = p;
            if (i != j) {
               col[pos[pj]] = pi;
               src[pos[pj]++] = p;
            }
         }
      P->csr_nnz = rp[n];
      P->hcsr_ptr = std::move(rp);
      P->hcsr_col = std::move(col);
      P->hcsr_src = std::move(src);
   }
   return P.release();
}

// Device copies of the analysis outputs and of the plan's task lists.
void upload_plan(DevicePlan* P, const Analysis& a) {
   if (P->csr_nnz > 0) {
      P->d_csr_ptr = dupload(P->hcsr_ptr);
      P->d_csr_col = dupload(P->hcsr_col);
      P->d_csr_src = dupload(P->hcsr_src);
      std::vector<int64_t>().swap(P->hcsr_ptr);
      std::vector<int>().swap(P->hcsr_col);
      std::vector<int64_t>().swap(P->hcsr_src);
   }
   P->d_sptr = dupload(a.sptr);
   P->d_sparent = dupload(a.sparent);
   P->d_rptr = dupload(a.rptr);
   P->d_rlist = dupload(a.rlist);
   P->d_nptr = dupload(a.nptr);
   P->d_nlist = dupload(a.nlist);
   P->d_cmap = dupload(P->hcmap);
   P->d_cmap_off = dupload(P->hcmap_off);
   P->d_nodes = dupload(P->hnodes);
   P->d_zero = dupload(P->zero);
   P->d_strip = dupload(P->strip);
   P->d_asmt = dupload(P->asmt);
   P->d_asmch = dupload(P->asmch);
   if (P->level_asm) P->d_asort = dupload(P->hasort);
   P->d_init = dupload(P->init);
   P->d_ea = dupload(P->ea);
   P->d_nlistv = dupload(P->nlistv);
   P->d_gch = dupload(P->gch);
   P->d_diag = dupload(P->diag);
   P->d_gemm = dupload(P->gemm);
   P->d_stask = dupload(P->stask);
   P->d_ctask = dupload(P->ctask);
   P->d_cgch = dupload(P->cgch);
   if ((int)a.invp.size() == a.n && a.n > 0) P->d_invp = dupload(a.invp);
}

// Checks every index a launch of the plan will use against the buffer it
// addresses (arena, inverse blocks, task lists, maps, exchange slots),
// without a device. Returns the number of violations; the first is
// described in *why.
int64_t check_plan_indices(const DevicePlan* P, const Analysis& a, std::string* why) {
   int64_t bad = 0;
   auto fail = [&](const std::string& m) {
      if (bad++ == 0 && why) *why = m;
   };
   const int nn = P->nnodes;
   const int64_t AE = P->arena_elems, DE = P->dinv_elems;
   auto in_arena = [&](int64_t off, int rows, int cols, int ld) {
      if (rows <= 0 || cols <= 0) return true;
      return off >= 0 && off + (int64_t)(rows - 1) + (int64_t)(cols - 1) * ld < AE;
   };
   for (int i = 0; i < nn; ++i) {
      const NodeDesc& d = P->hnodes[i];
      if (d.m < d.n || d.n < 0 || d.ld < d.m) fail("node " + std::to_string(i) + ": shape");
      if (!in_arena(d.off, d.m, d.m, d.ld)) fail("node " + std::to_string(i) + ": front outside the arena");
      if (d.parent < 0 || d.parent > nn) fail("node " + std::to_string(i) + ": parent");
      if (d.solve_off < 0 || d.solve_off + d.m > P->solve_rows) fail("node " + std::to_string(i) + ": solve rows");
      if (d.cmap_off < 0 || d.cmap_off + (d.m - d.n) > (int64_t)P->hcmap.size())
         fail("node " + std::to_string(i) + ": child map");
      if (d.n > 0) {
         const int64_t nblk = (d.n + kNB - 1) / kNB, last = d.n - (nblk - 1) * kNB;
         if (d.dinv_off < 0 || d.dinv_off + (nblk - 1) * kNB * kNB + last * last > DE)
            fail("node " + std::to_string(i) + ": inverse blocks");
      }
      if (d.parent < nn) {
         const NodeDesc& p = P->hnodes[d.parent];
         for (int k = 0; k < d.m - d.n; ++k) {
            const int c = P->hcmap[d.cmap_off + k];
            if (c < 0 || c >= p.m) fail("node " + std::to_string(i) + ": child map entry");
         }
      }
      if (d.num_a > 0 && (d.amap_off < 0 || 2 * (d.amap_off + d.num_a) > (int64_t)a.nlist.size()))
         fail("node " + std::to_string(i) + ": A map");
   }
   auto check_gemm = [&](const GemmTask& g, bool trsm) {
      if (g.M <= 0) return;
      if (!in_arena(g.a_off, g.M, g.K, g.lda)) fail("gemm: A outside the arena");
      if (trsm) {
         if (g.b_off < 0 || g.b_off + (int64_t)g.K * g.ldb > DE) fail("trsm: inverse block outside");
      } else if (!in_arena(g.b_off, g.N, g.K, g.ldb)) {
         fail("gemm: B outside the arena");
      }
      if (!in_arena(g.c_off, g.M, g.N, g.ldc)) fail("gemm: C outside the arena");
   };
   auto check_launches = [&](const std::vector<Launch>& prog, bool factor) {
      for (size_t li = 0; li < prog.size(); ++li) {
         const Launch& L = prog[li];
         if (L.wait_ev >= P->nevents || L.rec_ev >= P->nevents) fail("launch: event index");
         if (L.count < 0 || L.off < 0) fail("launch: range");
         const int64_t e = L.off + L.count;
         switch (L.kind) {
            case L_ZERO:
               if (e > (int64_t)P->zero.size()) fail("zero: range");
               else
                  for (int64_t t = L.off; t < e; ++t) {
                     const ColTask& c = P->zero[t];
                     if (c.node < 0 || c.node >= nn || c.c0 < 0 || c.c1 > P->hnodes[c.node].m || c.c0 >= c.c1)
                        fail("zero task");
                  }
               break;
            case L_INIT:
               if (e > (int64_t)P->init.size()) fail("init: range");
               else
                  for (int64_t t = L.off; t < e; ++t) {
                     const InitTask& c = P->init[t];
                     if (c.node < 0 || c.node >= nn || c.p0 < 0 || c.p1 > P->hnodes[c.node].num_a) fail("init task");
                  }
               break;
            case L_EA:
               if (e > (int64_t)P->ea.size()) fail("extend-add: range");
               else
                  for (int64_t t = L.off; t < e; ++t) {
                     const EaTask& c = P->ea[t];
                     if (c.child < 0 || c.child >= nn || P->hnodes[c.child].parent >= nn || c.j0 < 0 ||
                         c.j1 > P->hnodes[c.child].m - P->hnodes[c.child].n)
                        fail("extend-add task");
                  }
               break;
            case L_SMALL:
            case L_FSMALL:
            case L_BSMALL:
               if (e > (int64_t)P->nlistv.size()) fail("node list: range");
               else
                  for (int64_t t = L.off; t < e; ++t)
                     if (P->nlistv[t] < 0 || P->nlistv[t] >= nn) fail("node list entry");
               break;
            case L_SMALL_ASM:
               if (e > (int64_t)P->nlistv.size()) fail("node list: range");
               else
                  for (int64_t t = L.off; t < e; ++t) {
                     const int i = P->nlistv[t];
                     if (i < 0 || i >= nn) {
                        fail("node list entry");
                        continue;
                     }
                     const NodeDesc& d = P->hnodes[i];
                     if (d.amap_off < 0 || d.num_a < 0) fail("assembled front: A map");
                     int guard = 0;
                     for (int c = d.first_child; c != -1 && guard <= nn; c = P->hnodes[c].next_child, ++guard) {
                        if (c < 0 || c >= nn || P->hnodes[c].parent != i) {
                           fail("assembled front: child");
                           break;
                        }
                        const int64_t q = P->hnodes[c].m - P->hnodes[c].n;
                        if (P->hnodes[c].cmap_off < 0 || P->hnodes[c].cmap_off + q > (int64_t)P->hcmap.size())
                           fail("assembled front: child map");
                     }
                     if (guard > nn) fail("assembled front: child list cycle");
                  }
               break;
            case L_ASM:
               if (e > (int64_t)P->asmt.size()) {
                  fail("assembly: range");
               } else {
                  // the tasks of a front cover its columns once, in order
                  std::map<int, int> next_col;
                  std::vector<char> seen(nn, 0);
                  for (int64_t t = L.off; t < e; ++t) {
                     const AsmTask& c = P->asmt[t];
                     if (c.node < 0 || c.node >= nn) {
                        fail("assembly task: node");
                        continue;
                     }
                     const NodeDesc& d = P->hnodes[c.node];
                     if (c.c0 < 0 || c.c0 >= c.c1 || c.c1 > d.m) fail("assembly task: columns");
                     if (c.staged != 0) {  // assembled in LDS: the run's trapezoid must fit
                        int64_t area = 0;
                        for (int col = c.c0; col < c.c1 && col < d.m; ++col) area += d.m - col;
                        if (area > P->asm_elems) fail("assembly task: run larger than its LDS");
                     } else if (c.c1 != c.c0 + 1) {
                        fail("assembly task: a direct run is one column");
                     }
                     if (next_col[c.node] != c.c0) fail("assembly task: column runs");
                     next_col[c.node] = c.c1;
                     if (d.amap_off < 0 || (d.amap_off + d.num_a) * 2 > (int64_t)P->hasort.size()) {
                        fail("assembly task: A-map");
                     } else if (c.p0 < 0 || c.p0 > c.p1 || c.p1 > d.num_a) {
                        fail("assembly task: original-entry range");
                     } else {
                        // exactly the entries of the run's columns, each in its lower trapezoid
                        const int64_t* am = P->hasort.data() + 2 * d.amap_off;
                        const int64_t lo = (int64_t)c.c0 * d.m, hi = (int64_t)c.c1 * d.m;
                        if ((c.p0 > 0 && am[2 * (c.p0 - 1) + 1] - 1 >= lo) ||
                            (c.p1 < d.num_a && am[2 * c.p1 + 1] - 1 < hi))
                           fail("assembly task: original-entry range not exact");
                        for (int64_t p = c.p0; p < c.p1; ++p) {
                           const int64_t pos = am[2 * p + 1] - 1;
                           const int64_t col = pos / d.m, row = pos - col * d.m;
                           if (pos < lo || pos >= hi || row < col) {
                              fail("assembly task: original entry outside the run");
                              break;
                           }
                        }
                     }
                     // the child ranges: each child of the front with columns in
                     // the run, in list order, its columns landing in [c0, c1)
                     if (c.g0 < 0 || c.gn < 0 || (int64_t)c.g0 + c.gn > (int64_t)P->asmch.size()) {
                        fail("assembly task: child ranges");
                     } else {
                        int g = c.g0;
                        for (int ch = d.first_child, guard = 0; ch >= 0 && ch < nn && guard <= nn;
                             ch = P->hnodes[ch].next_child, ++guard) {
                           const NodeDesc& dc = P->hnodes[ch];
                           const int q = dc.m - dc.n;
                           if (dc.cmap_off < 0 || dc.cmap_off + q > (int64_t)P->hcmap.size()) break;
                           const int* cm = P->hcmap.data() + dc.cmap_off;
                           const int j0 = (int)(std::lower_bound(cm, cm + q, c.c0) - cm);
                           const int j1 = (int)(std::lower_bound(cm, cm + q, c.c1) - cm);
                           if (j0 >= j1) continue;
                           if (g >= c.g0 + c.gn || P->asmch[g].x != ch || P->asmch[g].y != j0 ||
                               P->asmch[g].z != j1) {
                              fail("assembly task: child range");
                              break;
                           }
                           ++g;
                        }
                        if (g != c.g0 + c.gn) fail("assembly task: child range count");
                     }
                     if (seen[c.node]) continue;  // the children once per front
                     seen[c.node] = 1;
                     int guard = 0;
                     for (int ch = d.first_child; ch != -1 && guard <= nn; ch = P->hnodes[ch].next_child, ++guard) {
                        if (ch < 0 || ch >= nn || P->hnodes[ch].parent != c.node) {
                           fail("assembly task: child");
                           break;
                        }
                        const NodeDesc& dc = P->hnodes[ch];
                        const int64_t q = dc.m - dc.n;
                        if (dc.cmap_off < 0 || dc.cmap_off + q > (int64_t)P->hcmap.size()) {
                           fail("assembly task: child map");
                           break;
                        }
                        for (int64_t k = 0; k < q; ++k) {
                           const int v = P->hcmap[dc.cmap_off + k];
                           if (v < 0 || v >= d.m || (k > 0 && v <= P->hcmap[dc.cmap_off + k - 1])) {
                              fail("assembly task: child map not increasing within the front");
                              break;
                           }
                        }
                     }
                     if (guard > nn) fail("assembly task: child list cycle");
                  }
                  for (const auto& kv : next_col)
                     if (kv.second != P->hnodes[kv.first].m) fail("assembly task: front not covered");
               }
               break;
            case L_DIAG:
            case L_DFACT:
            case L_DINV:
               if (e > (int64_t)P->diag.size()) fail("diag: range");
               else
                  for (int64_t t = L.off; t < e; ++t) {
                     const DiagTask& c = P->diag[t];
                     if (c.node < 0 || c.node >= nn || c.k0 < 0 || c.kb < 1 || c.kb > kNB ||
                         c.k0 + c.kb > P->hnodes[c.node].n)
                        fail("diag task");
                  }
               break;
            case L_STRIP:
               if (e > (int64_t)P->strip.size()) {
                  fail("strip: range");
               } else {
                  // the strip's diagonal block is factored by the launch before
                  const Launch* prev = li > 0 ? &prog[li - 1] : nullptr;
                  for (int64_t t = L.off; t < e; ++t) {
                     const StripTask& c = P->strip[t];
                     if (c.node < 0 || c.node >= nn) {
                        fail("strip task: node");
                        continue;
                     }
                     const NodeDesc& d = P->hnodes[c.node];
                     if (c.k0 < 0 || c.kb < 1 || c.kb > kNB || c.k0 + c.kb > d.n) fail("strip task: block");
                     if (c.r0 < c.k0 + c.kb || c.r0 >= d.m) fail("strip task: rows");
                     if ((int64_t)c.kb * d.ld * 8 > 0xFFFFFFFFll) fail("strip task: panel over 4 GiB");
                     bool dfact = false;
                     if (prev && prev->kind == L_DFACT && prev->off + prev->count <= (int64_t)P->diag.size())
                        for (int64_t q = prev->off; q < prev->off + prev->count; ++q)
                           dfact |= P->diag[q].node == c.node && P->diag[q].k0 == c.k0 && P->diag[q].kb == c.kb;
                     if (!dfact) fail("strip task: its diagonal block is not factored by the launch before");
                  }
               }
               break;
            case L_TRSM:
            case L_UPDATE:
            case L_UPDATE_WIDE:
               if (e > (int64_t)P->gemm.size()) fail("gemm: range");
               else
                  for (int64_t t = L.off; t < e; ++t) check_gemm(P->gemm[t], L.kind == L_TRSM);
               break;
            case L_FGATHER:
            case L_FDIAG:
            case L_BGATHER:
            case L_BDIAG:
            case L_BUPD:
            case L_BRED:
               if (e > (int64_t)P->stask.size()) fail("solve task: range");
               else
                  for (int64_t t = L.off; t < e; ++t) {
                     const SolveTask& c = P->stask[t];
                     if (c.node < 0 || c.node >= nn) {
                        fail("solve task node");
                        continue;
                     }
                     const NodeDesc& d = P->hnodes[c.node];
                     if (L.kind == L_FGATHER && (c.r0 < 0 || c.r1 > d.m || c.k0 < 0 || c.k0 + c.kb > (int)P->gch.size()))
                        fail("forward gather task");
                     if ((L.kind == L_FDIAG || L.kind == L_BDIAG) &&
                         (c.k0 < 0 || c.kb < 1 || c.k0 + c.kb > d.n || c.r0 < 0 || c.r1 > d.m))
                        fail("diagonal solve task");
                     if (L.kind == L_BGATHER && (c.r0 < 0 || c.r1 > d.m)) fail("backward gather task");
                     if (L.kind == L_BUPD && (c.k0 < d.n || c.kb > d.m || c.r0 < 0 || c.r1 > d.n || c.write < 0 ||
                                              c.write >= P->cb_slots))
                        fail("contribution-block solve task");
                     if (L.kind == L_BRED && (c.k0 < 0 || c.k0 + c.kb > P->cb_slots || c.r1 > d.n))
                        fail("reduction task");
                  }
               if (L.node1 >= nn) fail("solve launch node");
               break;
            case L_FCHAIN:
            case L_BCHAIN:
               if (e > (int64_t)P->ctask.size()) fail("chain: range");
               else
                  for (int64_t t = L.off; t < e; ++t) {
                     const ChainTask& c = P->ctask[t];
                     if (c.node < 0 || c.node >= nn) {
                        fail("chain task node");
                        continue;
                     }
                     const NodeDesc& d = P->hnodes[c.node];
                     if (c.blk < 0 || c.blk > c.nb || c.nb != (d.n + kNB - 1) / kNB || c.r0 < 0 || c.r1 > d.m ||
                         c.r1 - c.r0 > kNB || c.flag0 < 0 || c.flag0 + c.nb > P->chain_flags ||
                         c.g0 < 0 || c.g0 + c.gn > (int64_t)P->cgch.size() || c.cb0 < 0 || c.cb0 + c.cbn * ((d.n + 63) / 64) > P->cb_slots)
                        fail("chain task");
                  }
               break;
            case L_CLEAR:
            case L_SYNC:
               break;
            default:
               fail("launch kind");
         }
         if (L.stream != 0 && L.stream != 1) fail("launch stream");
      }
      // every side-stream launch records an event that a later launch of the
      // main stream waits for (the program ends joined)
      std::vector<int> pending;
      for (const Launch& L : prog) {
         if (L.wait_ev >= 0) pending.erase(std::remove(pending.begin(), pending.end(), L.wait_ev), pending.end());
         if (L.stream == 1) {
            if (L.rec_ev < 0) fail("side-stream launch without an event");
            pending.push_back(L.rec_ev);
         }
      }
      if (!factor && !pending.empty()) fail("side stream not joined");
   };
   check_launches(P->fprog, true);
   check_launches(P->fwdprog, false);
   check_launches(P->bwdprog, false);
   for (const auto* gv : {&P->gch, &P->cgch})
      for (const int4& g : *gv) {
         if (g.x < 0 || g.x >= nn) {
            fail("gather range: child");
            continue;
         }
         const NodeDesc& c = P->hnodes[g.x];
         if (g.y < 0 || g.y > g.z || g.z > c.m - c.n) fail("gather range: rows");
      }
   return bad;
}

DevicePlan* build_plan(const Analysis& a, int small_max, int wide_ncol, const PlanKnobs& k) {
   int dev = 0, num_cus = 0;
   HIPCHECK(hipGetDevice(&dev));
   HIPCHECK(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, dev));
   std::unique_ptr<DevicePlan> P(build_plan_host(a, small_max, wide_ncol, num_cus, k));
   upload_plan(P.get(), a);
   return P.release();
}

inline unsigned grid1(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

template <typename T>
void run_factor_program(const DevicePlan* P, DeviceFactors* F, hipStream_t s,
                        std::vector<hipEvent_t>* events = nullptr, bool compute_only = false) {
   const bool bf16_wide = std::is_same<T, float>::value && F->precision == SYLVER_AMD_PREC_MIXED;
   T* arena = static_cast<T*>(F->arena);
   T* dinv = static_cast<T*>(F->dinv);
   T* dinvT = static_cast<T*>(F->dinvT);
   size_t li = 0;
   const hipStream_t s_main = s;
   for (const Launch& L : P->fprog) {
      dim3 g(L.count), b(kThreads);
      // per-launch timing runs the program on one stream, so that each
      // kernel is timed alone
      s = (L.stream && !events) ? F->stream2 : s_main;
      if (L.wait_ev >= 0) HIPCHECK(hipStreamWaitEvent(s, F->xev[L.wait_ev], 0));
      if (events) HIPCHECK(hipEventRecord((*events)[2 * li], s));
      // compute_only: fronts already hold their values (front timing)
      const int kind =
          (compute_only && (L.kind == L_ZERO || L.kind == L_INIT || L.kind == L_EA || L.kind == L_ASM)) ? L_SYNC : L.kind;
      switch (kind) {
         case L_ZERO:
            k_zero<T><<<g, b, 0, s>>>(P->d_zero + L.off, P->d_nodes, arena);
            break;
         case L_INIT:
            k_init_a<T><<<g, b, 0, s>>>(P->d_init + L.off, P->d_nodes, P->d_nlist, P->d_rlist,
                                        F->d_aval, F->has_scaling ? F->d_scaling : nullptr, arena);
            break;
         case L_EA:
            k_extend_add<T><<<g, b, 0, s>>>(P->d_ea + L.off, P->d_nodes, P->d_cmap, arena);
            break;
         case L_ASM:
            k_assemble<T><<<g, b, P->asm_elems * sizeof(T), s>>>(P->d_asmt + L.off, P->d_nodes, P->d_asort, P->d_rlist,
                                                             F->d_aval, F->has_scaling ? F->d_scaling : nullptr,
                                                             P->d_cmap, P->d_asmch, arena);
            break;
         case L_SMALL:
            k_small_front<T><<<g, b, 0, s>>>(P->d_nlistv + L.off, P->d_nodes, arena, F->d_info);
            break;
         case L_SMALL_ASM:
            k_small_asm<T><<<g, b, 0, s>>>(P->d_nlistv + L.off, P->d_nodes, P->d_nlist, P->d_rlist, F->d_aval,
                                           F->has_scaling ? F->d_scaling : nullptr, P->d_cmap, arena, F->d_info);
            break;
         case L_DIAG:
            k_diag<T><<<g, b, 0, s>>>(P->d_diag + L.off, P->d_nodes, arena, dinv, dinvT, F->d_info, 1);
            break;
         case L_TRSM:
            // FP64: eight-wave workgroups (measured faster: 80^3 factor 0.3 ms,
            // the batched small fronts 1.1 ms; FP32 no better)
            if constexpr (sizeof(T) == 8)
               k_update<T, 1, 64, 128, 2 * kThreads><<<g, dim3(2 * kThreads), 0, s>>>(P->d_gemm + L.off, arena, dinv);
            else
               k_update<T, 1, 64, 128><<<g, b, 0, s>>>(P->d_gemm + L.off, arena, dinv);
            break;
         case L_UPDATE:
         case L_UPDATE_WIDE:
            if (L.kind == L_UPDATE_WIDE && bf16_wide)
               k_update_bf16<<<g, b, 0, s>>>(P->d_gemm + L.off, reinterpret_cast<float*>(arena));
            else if (L.tile == 64)
               k_update<T, 0, 64, 64, 2 * kThreads><<<g, dim3(2 * kThreads), 0, s>>>(P->d_gemm + L.off, arena, nullptr);
            else if constexpr (sizeof(T) == 4)
               k_update_far_f32<<<g, dim3(2 * kThreads), 0, s>>>(P->d_gemm + L.off, arena);
            else
               k_update<T, 0, 128, 128, 2 * kThreads><<<g, dim3(2 * kThreads), 0, s>>>(P->d_gemm + L.off, arena, nullptr);
            break;
         case L_DINV:
            k_diag<T><<<g, b, 0, s>>>(P->d_diag + L.off, P->d_nodes, arena, dinv, dinvT, F->d_info, 0);
            break;
         case L_DFACT:
            if constexpr (sizeof(T) == 4)
               k_diag_factor_f32<<<g, b, 0, s>>>(P->d_diag + L.off, P->d_nodes, arena, F->d_info);
            else
               k_diag_factor<T><<<g, b, 0, s>>>(P->d_diag + L.off, P->d_nodes, arena, F->d_info);
            break;
         case L_STRIP:
            k_strip<T><<<g, b, 0, s>>>(P->d_strip + L.off, P->d_nodes, arena);
            break;
         default:  // L_SYNC
            break;
      }
      if (events) HIPCHECK(hipEventRecord((*events)[2 * li + 1], s));
      if (L.rec_ev >= 0) HIPCHECK(hipEventRecord(F->xev[L.rec_ev], s));
      ++li;
   }
   HIPCHECK(hipGetLastError());
}

// dynamic LDS of a small-front solve launch: x (largest m) and y (largest n)
// rows of 16 doubles, each rounded up to 16 rows
size_t small_lds(const Launch& L) {
   return (size_t)(((L.smax_m + 15) & ~15) + ((L.smax_n + 15) & ~15)) * 16 * sizeof(double);
}

template <typename T>
void run_solve_program(const DevicePlan* P, DeviceFactors* F, const std::vector<Launch>& prog,
                       const double* b, double* xsol, int nrhs, hipStream_t s_main) {
   const T* arena = static_cast<const T*>(F->arena);
   const T* dinv = static_cast<const T*>(F->dinv);
   const T* dinvT = static_cast<const T*>(F->dinvT);
   const int n = P->n;
   for (const Launch& L : prog) {
      dim3 g(L.count), bl(kThreads);
      const SolveTask* tk = P->d_stask + L.off;
      const hipStream_t s = L.stream ? F->stream2 : s_main;
      if (L.wait_ev >= 0) HIPCHECK(hipStreamWaitEvent(s, F->xev[L.wait_ev], 0));
      switch (L.kind) {
         case L_FGATHER:
            k_fwd_gather<<<g, bl, 0, s>>>(tk, P->d_nodes, P->d_gch, P->d_rlist, P->d_cmap, b, n, F->d_xloc, nrhs);
            break;
         case L_FDIAG:
            if (L.node1 >= 0) {
               const SolveTask& t = P->stask[L.off];
               const NodeDesc& nd = P->hnodes[L.node1];
               k_solve_step1<T, false><<<g, bl, 0, s>>>(nd, L.node1, t.k0, t.kb, t.k0 + t.kb, nd.m, 64, arena, dinv,
                                                        P->d_rlist, F->d_xloc, xsol, n, nrhs);
            } else {
               k_solve_step<T, false><<<g, bl, 0, s>>>(tk, P->d_nodes, arena, dinv, P->d_rlist, F->d_xloc, xsol,
                                                       n, nrhs);
            }
            break;
         case L_BGATHER:
            k_bwd_gather<<<g, bl, 0, s>>>(tk, P->d_nodes, P->d_rlist, xsol, n, F->d_xloc, nrhs);
            break;
         case L_BDIAG:
            if (L.node1 >= 0) {
               const SolveTask& t = P->stask[L.off];
               const NodeDesc& nd = P->hnodes[L.node1];
               k_solve_step1<T, true><<<g, bl, 0, s>>>(nd, L.node1, t.k0, t.kb, 0, t.k0, 64, arena, dinvT,
                                                       P->d_rlist, F->d_xloc, xsol, n, nrhs);
            } else {
               k_solve_step<T, true><<<g, bl, 0, s>>>(tk, P->d_nodes, arena, dinvT, P->d_rlist, F->d_xloc, xsol,
                                                      n, nrhs);
            }
            break;
         case L_BUPD:
            if (L.count > 0)
               k_bwd_cbp<T><<<g, bl, 0, s>>>(tk, P->d_nodes, arena, xsol, n, P->d_rlist, F->d_part_cb, nrhs);
            break;
         case L_FSMALL:
            k_solve_small<T, false><<<g, bl, small_lds(L), s>>>(P->d_nlistv + L.off, P->d_nodes, arena, dinv,
                                                                P->d_rlist, F->d_xloc, xsol, n, nrhs, b, n,
                                                                P->d_cmap, (L.smax_m + 15) & ~15);
            break;
         case L_BSMALL:
            k_solve_small<T, true><<<g, bl, small_lds(L), s>>>(P->d_nlistv + L.off, P->d_nodes, arena, dinvT,
                                                               P->d_rlist, F->d_xloc, xsol, n, nrhs, nullptr, 0,
                                                               P->d_cmap, (L.smax_m + 15) & ~15);
            break;
         case L_BRED:
            if (L.count > 0) k_bwd_cbr<<<g, bl, 0, s>>>(tk, P->d_nodes, F->d_part_cb, F->d_xloc, nrhs);
            break;
         case L_CLEAR: {
            const int64_t ny = P->chain_flags * ((nrhs + 15) / 16) * kNB * 16;
            k_fill_y<<<dim3((unsigned)((ny + kThreads - 1) / kThreads)), bl, 0, s>>>(
                reinterpret_cast<unsigned long long*>(F->d_ybuf), ny);
            break;
         }
         case L_FCHAIN:
            k_chain<T, false><<<g, dim3(kChainThreads), 0, s>>>(P->d_ctask + L.off, P->d_nodes, arena, dinv,
                                                                P->d_rlist, F->d_xloc, xsol, n, nrhs, F->d_ybuf,
                                                                nullptr, F->d_chain, P->d_cgch, P->d_cmap, b, n);
            break;
         case L_BCHAIN:
            k_chain<T, true><<<g, dim3(kChainThreads), 0, s>>>(P->d_ctask + L.off, P->d_nodes, arena, dinvT,
                                                               P->d_rlist, F->d_xloc, xsol, n, nrhs, F->d_ybuf,
                                                               F->d_part_cb, F->d_chain, P->d_cgch, P->d_cmap,
                                                               nullptr, 0);
            break;
         default:  // L_SYNC
            break;
      }
      if (L.rec_ev >= 0) HIPCHECK(hipEventRecord(F->xev[L.rec_ev], s));
   }
   HIPCHECK(hipGetLastError());
}

constexpr int kMaxParts = 1024;  // workgroups per chunk of the refinement kernels

void ensure_solve_ws(const DevicePlan* P, DeviceFactors* F, int nrhs) {
   if (F->ws_nrhs >= nrhs) return;
   F->clear_solve_graphs();  // captured with the old buffers
   for (void* p : {(void*)F->d_xloc, (void*)F->d_b, (void*)F->d_x, (void*)F->d_r, (void*)F->d_sol,
                   (void*)F->d_norms, (void*)F->d_part_cb})
      dfree(p);
   F->d_xloc = dalloc<double>((size_t)P->solve_rows * nrhs);
   F->d_part_cb = dalloc<double>((size_t)std::max<int64_t>(P->cb_slots, 1) * 64 * nrhs);
   if (!F->d_chain) {
      F->d_chain = dalloc<int>(1);
      HIPCHECK(hipMemset(F->d_chain, 0, sizeof(int)));
   }
   dfree(F->d_ybuf);
   F->d_ybuf = dalloc<double>((size_t)std::max<int64_t>(P->chain_flags, 1) * ((nrhs + 15) / 16) * kNB * 16);
   F->d_b = dalloc<double>((size_t)P->n * nrhs);
   F->d_x = dalloc<double>((size_t)P->n * nrhs);
   F->d_r = dalloc<double>((size_t)P->n * nrhs);
   F->d_sol = dalloc<double>((size_t)P->n * nrhs);
   dfree(F->d_xin);
   F->d_xin = dalloc<double>((size_t)P->n * nrhs);
   dfree(F->d_sc);
   F->d_sc = dalloc<double>((size_t)P->n * nrhs);
   F->d_norms = dalloc<unsigned long long>(3 * (size_t)nrhs);
   dfree(F->d_xrm);
   F->d_xrm = dalloc<double>((size_t)P->n * 16 * ((nrhs + 15) / 16));
   dfree(F->d_maxpart);
   F->d_maxpart = dalloc<double>((size_t)2 * ((nrhs + 15) / 16) * kMaxParts * 16);
   F->ws_nrhs = nrhs;
}

// Norm equilibration of the CSR values into F->d_scaling (pivot order),
// following SPRAL's equilib_scale_sym as the reference calls it
// (src/spldlt_factorize_mod.F90:808-835): s = 1; per sweep the largest
// scaled entry of every row, s_i /= sqrt(max_i), stop after the sweep whose
// maxima are all within tol of 1, at most max_it sweeps.
void equilibrate(const DevicePlan* P, DeviceFactors* F, hipStream_t s, int max_it = 10, double tol = 1e-8) {
   const int n = P->n;
   k_fill<<<grid1(n), kThreads, 0, s>>>(n, 1.0, F->d_scaling);
   const int parts = (int)grid1(n);
   double* buf = dalloc<double>((size_t)n + parts + 1);
   double* rmax = buf;
   double* part = buf + n;
   for (int it = 0; it < max_it; ++it) {
      k_equil_rowmax<<<grid1(n), kThreads, 0, s>>>(n, P->d_csr_ptr, P->d_csr_col, F->d_csr_val, P->d_invp,
                                                   F->d_scaling, rmax);
      k_equil_update<<<parts, kThreads, 0, s>>>(n, rmax, F->d_scaling, part);
      k_max_reduce<<<1, kThreads, 0, s>>>(part, parts, part + parts);
      double dev = 0;
      HIPCHECK(hipMemcpyAsync(&dev, part + parts, sizeof(double), hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
      if (dev < tol) break;
   }
   dfree(buf);
}

}  // namespace

int plan_check(const Analysis& a, int small_max, int wide_ncol, int num_cus, int64_t* out, std::string* why) {
   std::unique_ptr<DevicePlan> P(build_plan_host(a, small_max, wide_ncol, num_cus, plan_knobs(ext_options().precision)));
   const int64_t bad = check_plan_indices(P.get(), a, why);
   if (out) {
      out[0] = bad;
      out[1] = (int64_t)P->fprog.size();
      out[2] = (int64_t)P->fwdprog.size();
      out[3] = (int64_t)P->bwdprog.size();
      out[4] = P->arena_elems;
      out[5] = P->chain_flags;
      out[6] = P->nlevels;
      out[7] = P->cb_slots;
      out[8] = (int64_t)P->strip.size();
      out[9] = (int64_t)P->diag.size();
   }
   return 0;
}

// ---------------------------------------------------------------------------
// Factorization driver
// ---------------------------------------------------------------------------
FactorResult device_factor(const Analysis& a, PlanPtr* plan, DeviceFactors** fact, bool posdef,
                           const double* aval, const double* scaling, const sylver_options_t& opts,
                           bool equilib, double* scaling_out) {
   FactorResult res;
   res.matrix_rank = a.n;
   if (device_count() <= 0) {
      res.flag = SYLVER_ERROR_CUDA_UNKNOWN;
      res.hip_error = (int)hipErrorNoDevice;
      return res;
   }
   const sylver_amd_options_t& xo = ext_options();
   try {
      int small_max = xo.small_front_max > 0 ? xo.small_front_max : 256;
      int wide_ncol = xo.bf16_min_ncol > 0 ? xo.bf16_min_ncol : 1024;
      // a new plan for new knobs or variants; factorizations built on the old
      // one keep it alive through their own reference
      int prec = xo.precision;
      if (prec != SYLVER_AMD_PREC_FP64 && prec != SYLVER_AMD_PREC_FP32 && prec != SYLVER_AMD_PREC_MIXED)
         prec = SYLVER_AMD_PREC_FP64;
      const PlanKnobs knobs = plan_knobs(prec);
      if (*plan && ((*plan)->small_max != small_max || (*plan)->wide_ncol != wide_ncol ||
                    (*plan)->fuse_small != knobs.fuse_small || (*plan)->strip_trsm != knobs.strip_trsm ||
                    (*plan)->level_asm != knobs.level_asm || (*plan)->asm_elems != knobs.asm_elems ||
                    (*plan)->variant_gen != a.variant_gen))
         plan->reset();
      if (!*plan) plan->reset(build_plan(a, small_max, wide_ncol, knobs));
      DevicePlan* P = plan->get();
      // Mixed precision without the BF16 update kernels runs the FP32 path.
      const bool low = prec != SYLVER_AMD_PREC_FP64;
      // the indefinite path's low tier is FP32 (Schur updates of the
      // contribution blocks on FP32 MFMA; no BF16 tier)
      if (!posdef && low) prec = SYLVER_AMD_PREC_FP32;
      DeviceFactors* F = *fact;
      if (F && (F->plan.get() != P || F->precision != prec || F->posdef != posdef)) {
         destroy_factors(F);
         F = nullptr;
         *fact = nullptr;
      }
      if (!F) {
         F = new DeviceFactors();
         *fact = F;
         F->plan = *plan;
         F->precision = prec;
         size_t es = low ? sizeof(float) : sizeof(double);
         HIPCHECK(hipStreamCreateWithFlags(&F->stream, hipStreamNonBlocking));
         HIPCHECK(hipStreamCreateWithFlags(&F->stream2, hipStreamNonBlocking));
         if (posdef) {  // the indefinite path lays out its fronts level by level
            const size_t ab = std::max<int64_t>(P->arena_elems, 1) * es;
            HIPCHECK(hipMalloc(&F->arena, ab));
            // once, outside any capture: padding rows (ld > m) and entries
            // above the diagonal that no kernel writes stay zero, so the
            // solves' full-height block loads never meet stale memory
            HIPCHECK(hipMemsetAsync(F->arena, 0, ab, F->stream));
            HIPCHECK(hipMalloc(&F->dinv, std::max<int64_t>(P->dinv_elems, 1) * es));
            HIPCHECK(hipMalloc(&F->dinvT, std::max<int64_t>(P->dinv_elems, 1) * es));
         }
         F->d_aval = dalloc<double>(std::max<int64_t>(P->clean_nnz, 1));
         F->d_scaling = dalloc<double>(std::max(a.n, 1));
         F->d_csr_val = dalloc<double>(std::max<int64_t>(P->csr_nnz, 1));
         F->d_info = dalloc<int>(std::max(P->nnodes, 1));
         F->xev.resize(P->nevents);
         for (auto& e : F->xev) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      }
      F->posdef = posdef;
      hipStream_t s = F->stream;
      hipEvent_t e0, e1, e2;
      HIPCHECK(hipEventCreate(&e0));
      HIPCHECK(hipEventCreate(&e1));
      HIPCHECK(hipEventCreate(&e2));
      HIPCHECK(hipEventRecord(e0, s));
      if (P->clean_nnz > 0)
         HIPCHECK(hipMemcpyAsync(F->d_aval, aval, P->clean_nnz * sizeof(double), hipMemcpyHostToDevice, s));
      F->has_scaling = scaling != nullptr;
      if (scaling) HIPCHECK(hipMemcpyAsync(F->d_scaling, scaling, a.n * sizeof(double), hipMemcpyHostToDevice, s));
      HIPCHECK(hipMemsetAsync(F->d_info, 0, std::max(P->nnodes, 1) * sizeof(int), s));
      // FP64 values of the whole matrix (CSR, pivot order) for the
      // equilibration, the norm and the refinement residual
      if (P->csr_nnz > 0)
         k_gather_vals<<<grid1(P->csr_nnz), kThreads, 0, s>>>(P->csr_nnz, P->d_csr_src, F->d_aval, nullptr,
                                                               nullptr, F->d_csr_val);
      if (equilib) {
         if (P->csr_nnz <= 0 || !P->d_invp) throw HipError{hipErrorInvalidValue, __LINE__};
         equilibrate(P, F, s);
         F->has_scaling = true;
         if (scaling_out)
            HIPCHECK(hipMemcpyAsync(scaling_out, F->d_scaling, a.n * sizeof(double), hipMemcpyDeviceToHost, s));
      }
      HIPCHECK(hipEventRecord(e1, s));
      const bool record = xo.record_timing != 0 && posdef;
      const bool use_graph = xo.use_graph != 0 && !record && posdef;
      std::vector<hipEvent_t> events;
      if (record) {
         events.resize(2 * P->fprog.size());
         for (auto& e : events) HIPCHECK(hipEventCreate(&e));
      }
      // the captured program bakes in whether the A-scatter scales
      if (use_graph && F->graph && F->graph_scaled != F->has_scaling) {
         (void)hipGraphExecDestroy(F->graph);
         F->graph = nullptr;
      }
      if (use_graph && !F->graph) {
         hipGraph_t g;
         HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
         if (low)
            run_factor_program<float>(P, F, s);
         else
            run_factor_program<double>(P, F, s);
         HIPCHECK(hipStreamEndCapture(s, &g));
         HIPCHECK(hipGraphInstantiate(&F->graph, g, nullptr, nullptr, 0));
         (void)hipGraphDestroy(g);
         F->graph_scaled = F->has_scaling;
      }
      if (!posdef) {
         factor_indef(a, P, F, opts, res);
      } else if (use_graph && F->graph) {
         HIPCHECK(hipGraphLaunch(F->graph, s));
      } else if (low) {
         run_factor_program<float>(P, F, s, record ? &events : nullptr);
      } else {
         run_factor_program<double>(P, F, s, record ? &events : nullptr);
      }
      HIPCHECK(hipEventRecord(e2, s));
      std::vector<int> info(std::max(P->nnodes, 1));
      HIPCHECK(hipMemcpyAsync(info.data(), F->d_info, info.size() * sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
      float t1 = 0, t2 = 0;
      HIPCHECK(hipEventElapsedTime(&t1, e0, e1));
      HIPCHECK(hipEventElapsedTime(&t2, e1, e2));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      (void)hipEventDestroy(e2);
      F->upload_ms = t1;
      F->factor_ms = t2;
      if (record) {
         const double es = low ? 4.0 : 8.0;
         for (int k = 0; k < 16; ++k) {
            F->kind_ms[k] = F->kind_flops[k] = F->kind_bytes[k] = 0;
            F->kind_launches[k] = 0;
         }
         for (size_t i = 0; i < P->fprog.size(); ++i) {
            float t = 0;
            HIPCHECK(hipEventElapsedTime(&t, events[2 * i], events[2 * i + 1]));
            const Launch& L = P->fprog[i];
            // the assembled one-workgroup fronts count with the one-workgroup kind
            // (kinds past the first 16 are counted with their nearest: the
            // level assembly with the extend-add)
            const int kind = L.kind == L_SMALL_ASM ? (int)L_SMALL
                             : L.kind == L_DFACT ? (int)L_DIAG
                             : L.kind == L_STRIP ? (int)L_TRSM
                             : L.kind == L_ASM   ? (int)L_EA
                                                 : L.kind;
            if (kind >= 16) continue;  // L_SYNC: no kernel
            F->kind_ms[kind] += t;
            F->kind_flops[kind] += L.flops;
            F->kind_bytes[kind] += L.words * es + L.bytes;
            F->kind_launches[kind] += 1;
         }
         for (auto& e : events) (void)hipEventDestroy(e);
      }
      F->nlaunches = posdef ? (int)P->fprog.size() : 0;
      res.upload_ms = t1;
      res.factor_ms = t2;
      if (posdef)
         for (int i = 0; i < P->nnodes; ++i)
            if (info[i] != 0) {
               res.flag = SYLVER_ERROR_NOT_POS_DEF;
               break;
            }
      // ||A||_inf of the unscaled matrix (full symmetric row sums) for the
      // refinement's backward error, from the device CSR
      F->anorm = 0;
      if (P->csr_nnz > 0) {
         const int parts = (int)std::min<int64_t>(kMaxParts, std::max<int64_t>(1, (a.n + kThreads - 1) / kThreads));
         double* part = dalloc<double>(parts + 1);
         k_csr_rowabs<<<parts, kThreads, 0, s>>>(a.n, P->d_csr_ptr, F->d_csr_val, part);
         k_max_reduce<<<1, kThreads, 0, s>>>(part, parts, part + parts);
         HIPCHECK(hipMemcpyAsync(&F->anorm, part + parts, sizeof(double), hipMemcpyDeviceToHost, s));
         HIPCHECK(hipStreamSynchronize(s));
         dfree(part);
      }
   } catch (const HipError& e) {
      res.flag = SYLVER_ERROR_CUDA_UNKNOWN;
      res.hip_error = (int)e.err;
      if (opts.print_level > 0)
         std::fprintf(stderr, "sylver_amd: HIP error %d (%s) at numeric.hip:%d\n", (int)e.err,
                      hipGetErrorString(e.err), e.line);
   }
   return res;
}

// ---------------------------------------------------------------------------
// Solve driver
// ---------------------------------------------------------------------------
namespace {

template <typename T>
void solve_once_eager(const DevicePlan* P, DeviceFactors* F, int job, const double* b, double* xsol,
                      int nrhs, hipStream_t s) {
   if (job == 0 || job == 1) run_solve_program<T>(P, F, P->fwdprog, b, xsol, nrhs, s);
   if (job == 3) {
      // backward only: the input is the current solution vector
      if (b != xsol)
         k_copy<<<grid1((int64_t)P->n * nrhs), kThreads, 0, s>>>((int64_t)P->n * nrhs, b, xsol);
   }
   if (job == 0 || job == 3) run_solve_program<T>(P, F, P->bwdprog, xsol, xsol, nrhs, s);
   if (job == 2 || job == 4) {
      // posdef: D = I
      if (job == 2 && b != xsol)
         k_copy<<<grid1((int64_t)P->n * nrhs), kThreads, 0, s>>>((int64_t)P->n * nrhs, b, xsol);
      if (job == 4) {
         if (b != xsol)
            k_copy<<<grid1((int64_t)P->n * nrhs), kThreads, 0, s>>>((int64_t)P->n * nrhs, b, xsol);
         run_solve_program<T>(P, F, P->bwdprog, xsol, xsol, nrhs, s);
      }
   }
}

// The solve program replayed from a hipGraph (one capture per job, buffers
// and right-hand-side count; refinement replays the same graph every step).
template <typename T>
void solve_once(const DevicePlan* P, DeviceFactors* F, int job, const double* b, double* xsol,
                int nrhs, hipStream_t s) {
   if (!ext_options().use_graph) {
      solve_once_eager<T>(P, F, job, b, xsol, nrhs, s);
      return;
   }
   const int es = (int)sizeof(T);
   for (const auto& g : F->sgraphs)
      if (g.job == job && g.nrhs == nrhs && g.es == es && g.b == b && g.x == xsol) {
         HIPCHECK(hipGraphLaunch(g.exec, s));
         return;
      }
   hipGraph_t graph;
   HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
   solve_once_eager<T>(P, F, job, b, xsol, nrhs, s);
   HIPCHECK(hipStreamEndCapture(s, &graph));
   hipGraphExec_t exec;
   HIPCHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
   (void)hipGraphDestroy(graph);
   F->sgraphs.push_back(DeviceFactors::SolveGraph{job, nrhs, es, b, xsol, exec});
   HIPCHECK(hipGraphLaunch(exec, s));
}

// The refinement's preconditioner: d = S (L L^T)^{-1} S r with the
// low-precision factors of S A S (S = I without a scaling). `r` is not
// modified; d may not alias r.
template <typename T>
void precond(const DevicePlan* P, DeviceFactors* F, const double* r, double* d, int nrhs, hipStream_t s) {
   auto solve = [&](const double* rhs) {
      if (F->posdef)
         solve_once<T>(P, F, 0, rhs, d, nrhs, s);
      else
         solve_indef(P, F, 0, rhs, d, nrhs, s);
   };
   if (!F->has_scaling) {
      solve(r);
      return;
   }
   const int64_t nn = (int64_t)P->n * nrhs;
   k_scale_rows<<<grid1(nn), kThreads, 0, s>>>(P->n, nrhs, F->d_scaling, r, F->d_sc);
   solve(F->d_sc);
   k_scale_rows<<<grid1(nn), kThreads, 0, s>>>(P->n, nrhs, F->d_scaling, d, d);
}

// One refinement update, fused: x += d, r = b - A x, and the scaled
// backward error max_c |r_c|_inf / (|A|_inf |x_c|_inf + |b_c|_inf), worst
// over the columns (max |b| per column already in norms[2 nrhs ..], from
// the start of the solve).
double refine_update_err(const DevicePlan* P, DeviceFactors* F, const double* d, int nrhs, hipStream_t s) {
   const int n = P->n, nch = (nrhs + 15) / 16;
   // grid-stride launches of at most kMaxParts workgroups per chunk, each
   // leaving its per-column maxima for one reduction launch
   const unsigned ga = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 63) / 64, kMaxParts));
   const unsigned gr =
       (unsigned)std::max<int64_t>(1, std::min<int64_t>((4 * (int64_t)n + kThreads - 1) / kThreads, kMaxParts));
   double* px = F->d_maxpart;
   double* pr = F->d_maxpart + (int64_t)nch * kMaxParts * 16;
   k_axpy_rm<<<dim3(ga, nch), kThreads, 0, s>>>(n, nrhs, d, F->d_x, F->d_xrm, px);
   k_residual_rm<<<dim3(gr, nch), kThreads, 0, s>>>(n, P->d_csr_ptr, P->d_csr_col, F->d_csr_val, F->d_xrm, F->d_b,
                                                    F->d_r, nrhs, pr);
   k_max_parts<<<dim3(16 * nch), kThreads, 0, s>>>(pr, (int)gr, nrhs, F->d_norms);
   k_max_parts<<<dim3(16 * nch), kThreads, 0, s>>>(px, (int)ga, nrhs, F->d_norms + nrhs);
   std::vector<unsigned long long> h(3 * nrhs);
   HIPCHECK(hipMemcpyAsync(h.data(), F->d_norms, h.size() * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, s));
   HIPCHECK(hipStreamSynchronize(s));
   double worst = 0;
   for (int c = 0; c < nrhs; ++c) {
      double rn, xn, bn;
      std::memcpy(&rn, &h[c], 8);
      std::memcpy(&xn, &h[nrhs + c], 8);
      std::memcpy(&bn, &h[2 * nrhs + c], 8);
      double den = F->anorm * xn + bn;
      double e = den > 0 ? rn / den : rn;
      if (std::isnan(e)) return e;
      worst = std::max(worst, e);
   }
   return worst;
}

// Per-column dot products <x, V_z> for z < nvec (reproducible two-pass
// reduction), returned on the host as h[z * nrhs + c].
void device_dots(DeviceFactors* F, int n, int nrhs, const double* x, const double* V, int64_t vstride,
                 int nvec, std::vector<double>& h, hipStream_t s) {
   dim3 g(kDotBlocks, nrhs, nvec);
   k_dot_partial<<<g, kThreads, 0, s>>>(n, x, V, vstride, F->d_part);
   const int cnt = nvec * nrhs;
   k_dot_final<<<(cnt + 63) / 64, 64, 0, s>>>(F->d_part, cnt, F->d_coef);
   h.resize(cnt);
   HIPCHECK(hipMemcpyAsync(h.data(), F->d_coef, cnt * sizeof(double), hipMemcpyDeviceToHost, s));
   HIPCHECK(hipStreamSynchronize(s));
}

void ensure_gmres_ws(const DevicePlan* P, DeviceFactors* F, int nrhs, int m) {
   if (F->gm_nrhs >= nrhs && F->gm_m >= m) return;
   for (void* p : {(void*)F->d_V, (void*)F->d_W, (void*)F->d_tmp, (void*)F->d_part, (void*)F->d_coef})
      dfree(p);
   const size_t nn = (size_t)P->n * nrhs;
   F->d_V = dalloc<double>(nn * (m + 1));
   F->d_W = dalloc<double>(nn);
   F->d_tmp = dalloc<double>(nn);
   F->d_part = dalloc<double>((size_t)kDotBlocks * nrhs * (m + 1));
   F->d_coef = dalloc<double>((size_t)nrhs * (m + 1));
   F->gm_nrhs = nrhs;
   F->gm_m = m;
}

// One GMRES-IR correction: approximately solves A d = r by GMRES in FP64 on
// the left-preconditioned system inv(L L^T) A d = inv(L L^T) r (the
// preconditioner is the low-precision factorization), Gram-Schmidt applied
// twice per step; stops when every column's preconditioned residual fell by
// `rtol` or after m steps. d is written to `dout`.
template <typename T>
int gmres_correction(const DevicePlan* P, DeviceFactors* F, const double* r, double* dout, int nrhs,
                     int m, double rtol, hipStream_t s) {
   const int n = P->n;
   const int64_t nn = (int64_t)n * nrhs;
   ensure_gmres_ws(P, F, nrhs, m);
   double* V = F->d_V;
   precond<T>(P, F, r, V, nrhs, s);
   std::vector<double> h, beta(nrhs), coef(nrhs);
   device_dots(F, n, nrhs, V, V, 0, 1, h, s);
   for (int c = 0; c < nrhs; ++c) {
      beta[c] = std::sqrt(h[c]);
      coef[c] = beta[c] > 0 ? 1.0 / beta[c] : 0.0;
   }
   HIPCHECK(hipMemcpyAsync(F->d_coef, coef.data(), nrhs * sizeof(double), hipMemcpyHostToDevice, s));
   k_scale_cols<<<grid1(nn), kThreads, 0, s>>>(n, nrhs, F->d_coef, V, V);
   // Hessenberg columns and Givens rotations per right-hand side
   std::vector<std::vector<double>> H(nrhs, std::vector<double>((size_t)(m + 1) * m, 0.0));
   std::vector<std::vector<double>> cs(nrhs, std::vector<double>(m)), sn(nrhs, std::vector<double>(m));
   std::vector<std::vector<double>> g(nrhs, std::vector<double>(m + 1, 0.0));
   for (int c = 0; c < nrhs; ++c) g[c][0] = beta[c];
   int steps = 0;
   for (int j = 0; j < m; ++j) {
      double* Vj = V + (int64_t)j * nn;
      double* Vn = V + (int64_t)(j + 1) * nn;
      k_spmv<<<grid1(n), kThreads, 0, s>>>(n, P->d_csr_ptr, P->d_csr_col, F->d_csr_val, Vj, F->d_tmp, nrhs);
      precond<T>(P, F, F->d_tmp, F->d_W, nrhs, s);
      std::vector<double> hsum((size_t)(j + 1) * nrhs, 0.0);
      for (int pass = 0; pass < 2; ++pass) {
         device_dots(F, n, nrhs, F->d_W, V, nn, j + 1, h, s);
         for (size_t k = 0; k < h.size(); ++k) hsum[k] += h[k];
         HIPCHECK(hipMemcpyAsync(F->d_coef, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
         k_multi_axpy<<<grid1(nn), kThreads, 0, s>>>(n, nrhs, j + 1, F->d_coef, V, nn, F->d_W);
      }
      device_dots(F, n, nrhs, F->d_W, F->d_W, 0, 1, h, s);
      bool done = true;
      for (int c = 0; c < nrhs; ++c) {
         const double hn = std::sqrt(h[c]);
         coef[c] = hn > 0 ? 1.0 / hn : 0.0;
         std::vector<double>& Hc = H[c];
         for (int i = 0; i <= j; ++i) Hc[i + (size_t)j * (m + 1)] = hsum[(size_t)i * nrhs + c];
         Hc[j + 1 + (size_t)j * (m + 1)] = hn;
         for (int i = 0; i < j; ++i) {
            const double a = Hc[i + (size_t)j * (m + 1)], b = Hc[i + 1 + (size_t)j * (m + 1)];
            Hc[i + (size_t)j * (m + 1)] = cs[c][i] * a + sn[c][i] * b;
            Hc[i + 1 + (size_t)j * (m + 1)] = -sn[c][i] * a + cs[c][i] * b;
         }
         const double a = Hc[j + (size_t)j * (m + 1)], b = Hc[j + 1 + (size_t)j * (m + 1)];
         const double rr = std::hypot(a, b);
         cs[c][j] = rr > 0 ? a / rr : 1.0;
         sn[c][j] = rr > 0 ? b / rr : 0.0;
         Hc[j + (size_t)j * (m + 1)] = rr;
         Hc[j + 1 + (size_t)j * (m + 1)] = 0.0;
         g[c][j + 1] = -sn[c][j] * g[c][j];
         g[c][j] = cs[c][j] * g[c][j];
         if (beta[c] > 0 && std::fabs(g[c][j + 1]) > rtol * beta[c]) done = false;
      }
      HIPCHECK(hipMemcpyAsync(F->d_coef, coef.data(), nrhs * sizeof(double), hipMemcpyHostToDevice, s));
      k_scale_cols<<<grid1(nn), kThreads, 0, s>>>(n, nrhs, F->d_coef, F->d_W, Vn);
      steps = j + 1;
      if (done) break;
   }
   // y = H^{-1} g per column, d = sum_i y_i V_i
   std::vector<double> y((size_t)steps * nrhs, 0.0);
   for (int c = 0; c < nrhs; ++c) {
      for (int i = steps - 1; i >= 0; --i) {
         double v = g[c][i];
         for (int k = i + 1; k < steps; ++k) v -= H[c][i + (size_t)k * (m + 1)] * y[(size_t)k * nrhs + c];
         const double d = H[c][i + (size_t)i * (m + 1)];
         y[(size_t)i * nrhs + c] = d != 0 ? v / d : 0.0;
      }
   }
   for (auto& v : y) v = -v;  // k_multi_axpy subtracts
   k_fill<<<grid1(nn), kThreads, 0, s>>>(nn, 0.0, dout);
   HIPCHECK(hipMemcpyAsync(F->d_coef, y.data(), y.size() * sizeof(double), hipMemcpyHostToDevice, s));
   k_multi_axpy<<<grid1(nn), kThreads, 0, s>>>(n, nrhs, steps, F->d_coef, V, nn, dout);
   return steps;
}

}  // namespace

SolveResult device_solve(DeviceFactors* F, int job, int nrhs, double* x, int ldx, bool user_order) {
   SolveResult res;
   if (!F || !F->plan) {
      res.flag = SYLVER_ERROR_CALL_SEQUENCE;
      return res;
   }
   const DevicePlan* P = F->plan.get();
   const sylver_amd_options_t& xo = ext_options();
   try {
      const int n = P->n;
      ensure_solve_ws(P, F, nrhs);
      hipStream_t s = F->stream;
      hipEvent_t e0, e1;
      HIPCHECK(hipEventCreate(&e0));
      HIPCHECK(hipEventCreate(&e1));
      const int64_t nn = (int64_t)n * nrhs;
      // the caller's x, staged (ld n); permuted (and for jobs 0/1 scaled)
      // into pivot order on the device
      const bool permute = user_order && P->d_invp;
      const bool scaled = user_order && F->has_scaling;
      const bool low = F->precision != SYLVER_AMD_PREC_FP64;
      const bool refine = low && job == 0 && xo.refine != SYLVER_AMD_REFINE_NONE;
      HIPCHECK(hipMemcpy2DAsync(permute ? F->d_xin : F->d_b, (size_t)n * sizeof(double), x,
                                (size_t)ldx * sizeof(double), (size_t)n * sizeof(double), nrhs,
                                hipMemcpyHostToDevice, s));
      HIPCHECK(hipEventRecord(e0, s));
      // refinement works on the unscaled system and applies S inside its
      // preconditioner; the plain solves take S on the way in and out
      const bool scale_in = scaled && !refine && (job == 0 || job == 1);
      const bool scale_out = scaled && !refine && (job == 0 || job == 3 || job == 4);
      if (permute)
         k_perm_in<<<grid1(nn), kThreads, 0, s>>>(n, nrhs, P->d_invp, scale_in ? F->d_scaling : nullptr, F->d_xin,
                                                  n, F->d_b);
      else if (scale_in)
         k_scale_rows<<<grid1(nn), kThreads, 0, s>>>(n, nrhs, F->d_scaling, F->d_b, F->d_b);
      if (F->ind && !F->posdef && !refine) {
         solve_indef(P, F, job, F->d_b, F->d_sol, nrhs, s);
         res.refine_mode = SYLVER_AMD_REFINE_NONE;
      } else if (!refine) {
         if (low)
            solve_once<float>(P, F, job, F->d_b, F->d_sol, nrhs, s);
         else
            solve_once<double>(P, F, job, F->d_b, F->d_sol, nrhs, s);
         res.refine_mode = SYLVER_AMD_REFINE_NONE;
      } else {
         // FP64 refinement: x0 = 0, r0 = b; classical steps
         // x_{k+1} = x_k + M^{-1} r_k, r_{k+1} = b - A x_{k+1}; in AUTO mode a
         // step that does not halve the backward error switches the remaining
         // steps to GMRES-IR corrections.
         k_fill<<<grid1(nn), kThreads, 0, s>>>(nn, 0.0, F->d_x);
         // max |b| per column, once
         k_fill_u64<<<1, 64, 0, s>>>(F->d_norms + 2 * nrhs, nrhs, 0ull);
         k_colmax<<<dim3(std::min<unsigned>(grid1(n), 256u), nrhs), kThreads, 0, s>>>(n, F->d_b, nrhs,
                                                                                     F->d_norms + 2 * nrhs);
         k_copy<<<grid1(nn), kThreads, 0, s>>>(nn, F->d_b, F->d_r);
         double err = 1e300;
         int it = 0;
         bool gm = xo.refine == SYLVER_AMD_REFINE_GMRES, converged = false, broken = false;
         for (; it < xo.refine_max_iter; ++it) {
            if (gm)
               gmres_correction<float>(P, F, F->d_r, F->d_sol, nrhs, 20, 1e-10, s);
            else
               precond<float>(P, F, F->d_r, F->d_sol, nrhs, s);
            double e = refine_update_err(P, F, F->d_sol, nrhs, s);
            if (!std::isfinite(e)) {
               err = e;
               broken = true;
               ++it;
               break;
            }
            if (!gm && xo.refine == SYLVER_AMD_REFINE_AUTO && it > 0 && !(e < 0.5 * err)) gm = true;
            err = e;
            if (e <= xo.refine_tol) {
               converged = true;
               ++it;
               break;
            }
         }
         k_copy<<<grid1(nn), kThreads, 0, s>>>(nn, F->d_x, F->d_sol);
         res.refine_mode = gm ? SYLVER_AMD_REFINE_GMRES : SYLVER_AMD_REFINE_IR;
         res.refine_iters = it;
         res.bwd_err = err;
         if (broken)
            res.flag = SYLVER_AMD_ERROR_REFINE;
         else if (!converged)
            res.flag = SYLVER_AMD_WARNING_REFINE;
      }
      double* out = F->d_sol;
      if (permute) {
         k_perm_out<<<grid1(nn), kThreads, 0, s>>>(n, nrhs, P->d_invp, scale_out ? F->d_scaling : nullptr, F->d_sol,
                                                   F->d_xin, n);
         out = F->d_xin;
      } else if (scale_out) {
         k_scale_rows<<<grid1(nn), kThreads, 0, s>>>(n, nrhs, F->d_scaling, F->d_sol, F->d_sol);
      }
      HIPCHECK(hipEventRecord(e1, s));
      HIPCHECK(hipMemcpy2DAsync(x, (size_t)ldx * sizeof(double), out, (size_t)n * sizeof(double),
                                (size_t)n * sizeof(double), nrhs, hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
      float t = 0;
      HIPCHECK(hipEventElapsedTime(&t, e0, e1));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      if (F->d_chain) {
         // waits of the chained solves that gave up (never expected: the
         // launches cannot deadlock); the result is then not to be used
         int werr = 0;
         HIPCHECK(hipMemcpy(&werr, F->d_chain, sizeof(int), hipMemcpyDeviceToHost));
         if (werr != 0) {
            HIPCHECK(hipMemset(F->d_chain, 0, sizeof(int)));
            res.flag = SYLVER_ERROR_CUDA_UNKNOWN;
            res.hip_error = (int)hipErrorLaunchTimeOut;
         }
      }
      res.solve_ms = t;
      F->solve_ms = t;
      F->refine_iters = res.refine_iters;
      F->bwd_err = res.bwd_err;
   } catch (const HipError& e) {
      res.flag = SYLVER_ERROR_CUDA_UNKNOWN;
      res.hip_error = (int)e.err;
   }
   return res;
}

// ---------------------------------------------------------------------------
// Statistics and read-back
// ---------------------------------------------------------------------------
int factors_stats(const DeviceFactors* F, sylver_amd_stats_t* st) {
   if (!F || !st) return SYLVER_ERROR_CALL_SEQUENCE;
   std::memset(st, 0, sizeof(*st));
   st->factor_ms = F->factor_ms;
   st->upload_ms = F->upload_ms;
   st->solve_ms = F->solve_ms;
   st->flops = F->plan ? (double)F->plan->num_flops : 0;
   st->nlevels = F->plan ? F->plan->nlevels : 0;
   st->nlaunches = F->nlaunches;
   st->precision = F->precision;
   st->refine_iters = F->refine_iters;
   st->bwd_err = F->bwd_err;
   for (int k = 0; k < 16; ++k) {
      st->kind_ms[k] = F->kind_ms[k];
      st->kind_flops[k] = F->kind_flops[k];
      st->kind_bytes[k] = F->kind_bytes[k];
      st->kind_launches[k] = F->kind_launches[k];
   }
   return 0;
}

int plan_device_symbolic(const DevicePlan* P, int* sptr, int* sparent, int64_t* rptr, int* rlist,
                         int64_t* nptr, int64_t* nlist, int* cmap, int64_t* cmap_off) {
   if (!P) return SYLVER_ERROR_CALL_SEQUENCE;
   try {
      int nn = P->nnodes;
      auto cp = [](void* dst, const void* src, size_t bytes) {
         if (dst && bytes) HIPCHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
      };
      std::vector<int64_t> hr(nn + 1), hn(nn + 1);
      HIPCHECK(hipMemcpy(hr.data(), P->d_rptr, (nn + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(hn.data(), P->d_nptr, (nn + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
      cp(sptr, P->d_sptr, (nn + 1) * sizeof(int));
      cp(sparent, P->d_sparent, nn * sizeof(int));
      if (rptr) std::memcpy(rptr, hr.data(), (nn + 1) * sizeof(int64_t));
      cp(rlist, P->d_rlist, (hr[nn] - 1) * sizeof(int));
      if (nptr) std::memcpy(nptr, hn.data(), (nn + 1) * sizeof(int64_t));
      cp(nlist, P->d_nlist, 2 * (hn[nn] - 1) * sizeof(int64_t));
      cp(cmap, P->d_cmap, P->hcmap.size() * sizeof(int));
      cp(cmap_off, P->d_cmap_off, (nn + 1) * sizeof(int64_t));
      return 0;
   } catch (const HipError&) {
      return SYLVER_ERROR_CUDA_UNKNOWN;
   }
}

int factors_node_factor(const DeviceFactors* F, int node, double* l) {
   if (!F || !F->plan || node < 0 || node >= F->plan->nnodes || !l) return SYLVER_ERROR_CALL_SEQUENCE;
   try {
      const NodeDesc& d = F->plan->hnodes[node];
      const bool dbl = F->precision == SYLVER_AMD_PREC_FP64;
      const size_t es = dbl ? sizeof(double) : sizeof(float);
      std::vector<char> tmp((size_t)d.m * d.n * es + 1);
      const char* src = static_cast<const char*>(F->arena) + (size_t)d.off * es;
      HIPCHECK(hipMemcpy2D(tmp.data(), (size_t)d.m * es, src, (size_t)d.ld * es, (size_t)d.m * es, d.n,
                           hipMemcpyDeviceToHost));
      for (size_t i = 0; i < (size_t)d.m * d.n; ++i)
         l[i] = dbl ? reinterpret_cast<const double*>(tmp.data())[i]
                    : (double)reinterpret_cast<const float*>(tmp.data())[i];
      return 0;
   } catch (const HipError&) {
      return SYLVER_ERROR_CUDA_UNKNOWN;
   }
}

int factors_node_assembled(const DeviceFactors* F, int node, double* front) {
   (void)F;
   (void)node;
   (void)front;
   return SYLVER_ERROR_UNIMPLEMENTED;
}

int dev_alloc(void** ptr, size_t bytes) {
   if (!ptr) return SYLVER_ERROR_CALL_SEQUENCE;
   return hipMalloc(ptr, bytes ? bytes : 1) == hipSuccess ? 0 : SYLVER_ERROR_CUDA_UNKNOWN;
}
int dev_free(void* ptr) { return hipFree(ptr) == hipSuccess ? 0 : SYLVER_ERROR_CUDA_UNKNOWN; }
int dev_copy(void* dst, const void* src, size_t bytes, bool h2d) {
   hipError_t e = hipMemcpy(dst, src, bytes, h2d ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost);
   return e == hipSuccess ? 0 : SYLVER_ERROR_CUDA_UNKNOWN;
}

static double g_last_batched_ms = 0;
double last_batched_ms() { return g_last_batched_ms; }
void set_last_batched_ms(double ms) { g_last_batched_ms = ms; }

// ---------------------------------------------------------------------------
// Batched dense fronts (no tree): every front through the one-workgroup
// kernel in one variable-size launch.
// ---------------------------------------------------------------------------
int batched_front_fill(int nfronts, const int* nrow, const int64_t* off, void* data, int dtype,
                       unsigned long long seed) {
   if (nfronts < 0 || (nfronts > 0 && (!nrow || !off || !data))) return SYLVER_ERROR_CALL_SEQUENCE;
   if (nfronts == 0) return 0;
   if (device_count() <= 0) return SYLVER_ERROR_CUDA_UNKNOWN;
   try {
      int* dn = dalloc<int>(nfronts);
      int64_t* doff = dalloc<int64_t>(nfronts);
      HIPCHECK(hipMemcpy(dn, nrow, nfronts * sizeof(int), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(doff, off, nfronts * sizeof(int64_t), hipMemcpyHostToDevice));
      dim3 g(8, nfronts);
      if (dtype == 0)
         k_front_fill<double><<<g, kThreads>>>(dn, nullptr, doff, static_cast<double*>(data), seed);
      else
         k_front_fill<float><<<g, kThreads>>>(dn, nullptr, doff, static_cast<float*>(data), seed);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipDeviceSynchronize());
      dfree(dn);
      dfree(doff);
      return 0;
   } catch (const HipError&) {
      return SYLVER_ERROR_CUDA_UNKNOWN;
   }
}

// Device time of the posdef factorization of K independent synthetic SPD
// fronts (m_i rows, n_i pivots, values from k_front_fill) with the given
// per-front variants (1 one-workgroup kernel, 0 tiled path, -1 default):
// the plan's compute launches only, median of `reps` runs, values refilled
// before each run. The autotuner's measurement.
template <typename T>
double time_fronts_t(DevicePlan* P, DeviceFactors* F, const std::vector<int>& nrow, const std::vector<int>& ld,
                     const std::vector<int64_t>& off, int reps) {
   const int K = (int)nrow.size();
   int* dn = dalloc<int>(K);
   int* dl = dalloc<int>(K);
   int64_t* doff = dalloc<int64_t>(K);
   HIPCHECK(hipMemcpy(dn, nrow.data(), K * sizeof(int), hipMemcpyHostToDevice));
   HIPCHECK(hipMemcpy(dl, ld.data(), K * sizeof(int), hipMemcpyHostToDevice));
   HIPCHECK(hipMemcpy(doff, off.data(), K * sizeof(int64_t), hipMemcpyHostToDevice));
   hipEvent_t e0, e1;
   HIPCHECK(hipEventCreate(&e0));
   HIPCHECK(hipEventCreate(&e1));
   std::vector<double> t;
   for (int r = 0; r < reps + 1; ++r) {
      k_front_fill<T><<<dim3(8, K), kThreads, 0, F->stream>>>(dn, dl, doff, static_cast<T*>(F->arena), 12345);
      HIPCHECK(hipEventRecord(e0, F->stream));
      run_factor_program<T>(P, F, F->stream, nullptr, true);
      HIPCHECK(hipEventRecord(e1, F->stream));
      HIPCHECK(hipStreamSynchronize(F->stream));
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
      if (r > 0) t.push_back(ms);  // the first run warms up
   }
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   dfree(dn);
   dfree(dl);
   dfree(doff);
   std::sort(t.begin(), t.end());
   return t[t.size() / 2];
}

int time_fronts(int K, const int* m, const int* n, const signed char* variant, int precision, int reps,
                double* ms) {
   if (K <= 0 || !m || !n || !ms || reps < 1) return SYLVER_ERROR_CALL_SEQUENCE;
   for (int i = 0; i < K; ++i)
      if (n[i] < 1 || m[i] < n[i]) return SYLVER_ERROR_CALL_SEQUENCE;
   if (device_count() <= 0) return SYLVER_ERROR_CUDA_UNKNOWN;
   PlanPtr P;
   DeviceFactors* F = nullptr;
   try {
      // a forest of K roots: rows 1..m_i of a matrix of order max m_i
      Analysis a;
      a.nnodes = K;
      a.fronts.assign(K + 1, SymbolicFront{});
      int64_t roff = 0;
      for (int i = 0; i < K; ++i) {
         a.n = std::max(a.n, m[i]);
         SymbolicFront& f = a.fronts[i];
         f.nrow = m[i];
         f.ncol = n[i];
         f.parent = K;
         f.rlist_off = roff;
         roff += m[i];
         for (int r = 1; r <= m[i]; ++r) a.rlist.push_back(r);
      }
      if (variant) a.variant.assign(variant, variant + K);
      P.reset(build_plan(a, 256, 1 << 30, PlanKnobs{}));
      const bool low = precision != SYLVER_AMD_PREC_FP64;
      const size_t es = low ? sizeof(float) : sizeof(double);
      F = new DeviceFactors();
      F->plan = P;
      F->precision = low ? SYLVER_AMD_PREC_FP32 : SYLVER_AMD_PREC_FP64;
      F->posdef = true;
      HIPCHECK(hipMalloc(&F->arena, std::max<int64_t>(P->arena_elems, 1) * es));
      HIPCHECK(hipMalloc(&F->dinv, std::max<int64_t>(P->dinv_elems, 1) * es));
      HIPCHECK(hipMalloc(&F->dinvT, std::max<int64_t>(P->dinv_elems, 1) * es));
      F->d_info = dalloc<int>(K);
      HIPCHECK(hipMemset(F->d_info, 0, K * sizeof(int)));
      HIPCHECK(hipStreamCreateWithFlags(&F->stream, hipStreamNonBlocking));
      HIPCHECK(hipStreamCreateWithFlags(&F->stream2, hipStreamNonBlocking));
      F->xev.resize(P->nevents);
      for (auto& e : F->xev) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      std::vector<int> nrow(m, m + K), ld(K);
      std::vector<int64_t> off(K);
      for (int i = 0; i < K; ++i) {
         ld[i] = P->hnodes[i].ld;
         off[i] = P->hnodes[i].off;
      }
      *ms = low ? time_fronts_t<float>(P.get(), F, nrow, ld, off, reps)
                : time_fronts_t<double>(P.get(), F, nrow, ld, off, reps);
      std::vector<int> info(K);
      HIPCHECK(hipMemcpy(info.data(), F->d_info, K * sizeof(int), hipMemcpyDeviceToHost));
      destroy_factors(F);
      for (int v : info)
         if (v != 0) return SYLVER_ERROR_NOT_POS_DEF;
      return 0;
   } catch (const HipError&) {
      destroy_factors(F);
      return SYLVER_ERROR_CUDA_UNKNOWN;
   }
}

// Batched partial factorization of independent fronts in the caller's device
// memory (front f at data + off[f], ld = nrow[f]) through the launch plan:
// per front the one-workgroup kernel or the tiled path (variant[f] = 1 / 0,
// -1 or NULL the default rule), all fronts of a variant batched in the same
// launches. info[f] (device) as batched_front_potrf.
int batched_front_potrf_v(int nfronts, const int* nrow, const int* ncol, const int64_t* off, void* data,
                          int dtype, const signed char* variant, int* info) {
   if (nfronts < 0 || (nfronts > 0 && (!nrow || !ncol || !off || !data || !info)))
      return SYLVER_ERROR_CALL_SEQUENCE;
   if (nfronts == 0) return 0;
   for (int f = 0; f < nfronts; ++f)
      if (nrow[f] < ncol[f] || ncol[f] < 1) return SYLVER_ERROR_CALL_SEQUENCE;
   if (device_count() <= 0) return SYLVER_ERROR_CUDA_UNKNOWN;
   PlanPtr P;
   DeviceFactors* F = nullptr;
   try {
      Analysis a;
      a.nnodes = nfronts;
      a.fronts.assign(nfronts + 1, SymbolicFront{});
      a.fixed_off.assign(off, off + nfronts);
      a.fixed_ld.assign(nrow, nrow + nfronts);
      int64_t roff = 0;
      for (int i = 0; i < nfronts; ++i) {
         a.n = std::max(a.n, nrow[i]);
         SymbolicFront& f = a.fronts[i];
         f.nrow = nrow[i];
         f.ncol = ncol[i];
         f.parent = nfronts;
         f.rlist_off = roff;
         roff += nrow[i];
         for (int r = 1; r <= nrow[i]; ++r) a.rlist.push_back(r);
      }
      if (variant) a.variant.assign(variant, variant + nfronts);
      P.reset(build_plan(a, 256, 1 << 30, PlanKnobs{}));
      const bool low = dtype != 0;
      const size_t es = low ? sizeof(float) : sizeof(double);
      F = new DeviceFactors();
      F->plan = P;
      F->precision = low ? SYLVER_AMD_PREC_FP32 : SYLVER_AMD_PREC_FP64;
      F->posdef = true;
      F->arena = data;  // the caller's memory: detached before the factors are freed
      HIPCHECK(hipMalloc(&F->dinv, std::max<int64_t>(P->dinv_elems, 1) * es));
      HIPCHECK(hipMalloc(&F->dinvT, std::max<int64_t>(P->dinv_elems, 1) * es));
      F->d_info = dalloc<int>(nfronts);
      HIPCHECK(hipMemset(F->d_info, 0, nfronts * sizeof(int)));
      HIPCHECK(hipStreamCreateWithFlags(&F->stream, hipStreamNonBlocking));
      HIPCHECK(hipStreamCreateWithFlags(&F->stream2, hipStreamNonBlocking));
      F->xev.resize(P->nevents);
      for (auto& e : F->xev) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      hipEvent_t e0, e1;
      HIPCHECK(hipEventCreate(&e0));
      HIPCHECK(hipEventCreate(&e1));
      HIPCHECK(hipEventRecord(e0, F->stream));
      if (low)
         run_factor_program<float>(P.get(), F, F->stream, nullptr, true);
      else
         run_factor_program<double>(P.get(), F, F->stream, nullptr, true);
      HIPCHECK(hipEventRecord(e1, F->stream));
      HIPCHECK(hipStreamSynchronize(F->stream));
      float t = 0;
      HIPCHECK(hipEventElapsedTime(&t, e0, e1));
      g_last_batched_ms = t;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      HIPCHECK(hipMemcpy(info, F->d_info, nfronts * sizeof(int), hipMemcpyDeviceToDevice));
      F->arena = nullptr;
      destroy_factors(F);
      return 0;
   } catch (const HipError&) {
      if (F) F->arena = nullptr;
      destroy_factors(F);
      return SYLVER_ERROR_CUDA_UNKNOWN;
   }
}

int batched_front_potrf(int nfronts, const int* nrow, const int* ncol, const int64_t* off, void* data,
                        int dtype, int* info, void* stream) {
   if (nfronts < 0 || (nfronts > 0 && (!nrow || !ncol || !off || !data || !info)))
      return SYLVER_ERROR_CALL_SEQUENCE;
   if (nfronts == 0) return 0;
   if (device_count() <= 0) return SYLVER_ERROR_CUDA_UNKNOWN;
   try {
      std::vector<NodeDesc> nd(nfronts);
      std::vector<int> list(nfronts);
      for (int f = 0; f < nfronts; ++f) {
         if (nrow[f] < ncol[f] || ncol[f] < 0) return SYLVER_ERROR_CALL_SEQUENCE;
         std::memset(&nd[f], 0, sizeof(NodeDesc));
         nd[f].m = nrow[f];
         nd[f].n = ncol[f];
         nd[f].ld = nrow[f];
         nd[f].off = off[f];
         nd[f].first_child = nd[f].next_child = -1;
         list[f] = f;
      }
      hipStream_t s = static_cast<hipStream_t>(stream);
      NodeDesc* dnd = dalloc<NodeDesc>(nfronts);
      int* dl = dalloc<int>(nfronts);
      HIPCHECK(hipMemcpyAsync(dnd, nd.data(), nfronts * sizeof(NodeDesc), hipMemcpyHostToDevice, s));
      HIPCHECK(hipMemcpyAsync(dl, list.data(), nfronts * sizeof(int), hipMemcpyHostToDevice, s));
      HIPCHECK(hipMemsetAsync(info, 0, nfronts * sizeof(int), s));
      hipEvent_t e0, e1;
      HIPCHECK(hipEventCreate(&e0));
      HIPCHECK(hipEventCreate(&e1));
      HIPCHECK(hipEventRecord(e0, s));
      if (dtype == 0)
         k_small_front<double><<<nfronts, kThreads, 0, s>>>(dl, dnd, static_cast<double*>(data), info);
      else
         k_small_front<float><<<nfronts, kThreads, 0, s>>>(dl, dnd, static_cast<float*>(data), info);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipEventRecord(e1, s));
      HIPCHECK(hipStreamSynchronize(s));
      float t = 0;
      HIPCHECK(hipEventElapsedTime(&t, e0, e1));
      g_last_batched_ms = t;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      dfree(dnd);
      dfree(dl);
      return 0;
   } catch (const HipError&) {
      return SYLVER_ERROR_CUDA_UNKNOWN;
   }
}

// ---------------------------------------------------------------------------
// Allocation-free batched factorization. The per-front sizes and offsets are
// staged in a fixed device array instead of allocated descriptors: int64
// offsets in the first kNoAllocMaxFronts slots, then nrow and ncol as ints.
// Each device holds its own copy for the life of the process; every call on
// a device reuses it, so calls that overlap on one device race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) int64_t scratch[2 * kNoAllocMaxFronts];  // 1 MiB

// Device address of `scratch` per device, taken once per device with that
// device current (the array's name in host code is not a device address).
static std::mutex g_scratch_mu;
static int64_t* g_scratch_ptr[kMaxScratchDevices] = {};

static int64_t* scratch_ptr(int dev) {
   std::lock_guard<std::mutex> lock(g_scratch_mu);
   if (!g_scratch_ptr[dev]) {
      void* p = nullptr;
      HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      if (!p) throw HipError{hipErrorInvalidSymbol, __LINE__};
      g_scratch_ptr[dev] = static_cast<int64_t*>(p);
   }
   return g_scratch_ptr[dev];
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_small_front_sized(const int64_t* __restrict__ off,
                                                                const int* __restrict__ nrow,
                                                                const int* __restrict__ ncol,
                                                                T* __restrict__ data, int* __restrict__ info) {
   const int f = blockIdx.x;
   NodeDesc nd{};
   nd.m = nrow[f];
   nd.n = ncol[f];
   nd.ld = nd.m;
   nd.off = off[f];
   nd.first_child = nd.next_child = -1;
   if (threadIdx.x == 0) info[f] = 0;  // front_potrf_block opens with a barrier
   front_potrf_block<T>(nd, data + nd.off, info + f);
}

int batched_front_potrf_noalloc(int nfronts, const int* nrow, const int* ncol, const int64_t* off, void* data,
                                int dtype, int* info, void* stream) {
   if (nfronts < 0 || nfronts > kNoAllocMaxFronts || (nfronts > 0 && (!nrow || !ncol || !off || !data || !info)))
      return SYLVER_ERROR_CALL_SEQUENCE;
   if (dtype != 0 && dtype != 1) return SYLVER_ERROR_CALL_SEQUENCE;
   for (int f = 0; f < nfronts; ++f)
      if (nrow[f] < ncol[f] || ncol[f] < 0 || off[f] < 0) return SYLVER_ERROR_CALL_SEQUENCE;
   if (nfronts == 0) return 0;
   if (device_count() <= 0) return SYLVER_ERROR_CUDA_UNKNOWN;
   try {
      int dev = 0;
      HIPCHECK(hipGetDevice(&dev));
      if (dev < 0 || dev >= kMaxScratchDevices) return SYLVER_ERROR_CUDA_UNKNOWN;
      int64_t* s_off = scratch_ptr(dev);
      int* s_nrow = reinterpret_cast<int*>(s_off + kNoAllocMaxFronts);
      int* s_ncol = s_nrow + kNoAllocMaxFronts;
      hipStream_t s = static_cast<hipStream_t>(stream);
      HIPCHECK(hipMemcpyAsync(s_off, off, nfronts * sizeof(int64_t), hipMemcpyHostToDevice, s));
      HIPCHECK(hipMemcpyAsync(s_nrow, nrow, nfronts * sizeof(int), hipMemcpyHostToDevice, s));
      HIPCHECK(hipMemcpyAsync(s_ncol, ncol, nfronts * sizeof(int), hipMemcpyHostToDevice, s));
      if (dtype == 0)
         k_small_front_sized<double><<<nfronts, kThreads, 0, s>>>(s_off, s_nrow, s_ncol,
                                                                   static_cast<double*>(data), info);
      else
         k_small_front_sized<float><<<nfronts, kThreads, 0, s>>>(s_off, s_nrow, s_ncol,
                                                                  static_cast<float*>(data), info);
      HIPCHECK(hipGetLastError());
      // the host arrays are the caller's and the scratch is shared: finish
      // before returning
      HIPCHECK(hipStreamSynchronize(s));
      return 0;
   } catch (const HipError&) {
      return SYLVER_ERROR_CUDA_UNKNOWN;
   }
}

}  // namespace sylver_amd
