// Internal to the device numeric phase: error handling, allocation helpers,
// launch bookkeeping and the plan / factor state shared by the posdef
// (numeric.hip) and the indefinite (indef.hip) paths.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "kernels.hpp"
#include "numeric.hpp"

namespace sylver_amd {

namespace {

struct HipError {
   hipError_t err;
   int line;
};

#define HIPCHECK(x)                                   \
   do {                                              \
      hipError_t e_ = (x);                           \
      if (e_ != hipSuccess) throw HipError{e_, __LINE__}; \
   } while (0)

template <typename T>
T* dalloc(size_t n) {
   void* p = nullptr;
   if (n == 0) n = 1;
   HIPCHECK(hipMalloc(&p, n * sizeof(T)));
   return static_cast<T*>(p);
}

template <typename T>
T* dupload(const std::vector<T>& v) {
   T* p = dalloc<T>(v.size());
   if (!v.empty()) HIPCHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
   return p;
}

void dfree(void* p) {
   if (p) (void)hipFree(p);
}

enum LaunchKind {
   L_ZERO,
   L_INIT,
   L_EA,
   L_SMALL,
   L_DIAG,
   L_TRSM,
   L_UPDATE,
   L_DINV,
   L_UPDATE_WIDE,
   L_FGATHER,
   L_FDIAG,
   L_FUPD,
   L_BGATHER,
   L_BDIAG,
   L_BUPD,
   L_BRED,
   L_FSMALL,
   L_BSMALL,
   L_FCHAIN,  // level-wide forward solve of the level's larger fronts (k_chain)
   L_BCHAIN,
   L_CLEAR,   // exchange-buffer fill at the start of a chained sweep
   L_SMALL_ASM,  // one-workgroup fronts assembled (zero, A, children) and factored in one launch
   L_DFACT,  // diagonal blocks factored without their inverse (k_diag_factor), before L_STRIP
   L_ASM,    // a level's assembly (zero, A entries, children) by receiving columns (k_assemble)
   L_STRIP,  // panel TRSM by 64-row strips, by substitution (k_strip)
   L_SYNC  // no kernel: only the stream wait / event record of the entry
};

struct Launch {
   int kind;
   int count;
   int64_t off;
   int stream = 0;    // 0: the factor stream, 1: the look-ahead side stream
   int wait_ev = -1;  // event (index) the launch's stream waits for first
   int rec_ev = -1;   // event (index) recorded on the launch's stream after it
   double flops = 0;  // algorithmic flops of the launch
   double words = 0;  // algorithmic memory traffic in front elements
   double bytes = 0;  // algorithmic memory traffic independent of the element size
   int node1 = -1;    // solve steps: the one front all tasks belong to (or -1)
   int tile = 128;    // Schur updates: tile size of the launch's tasks (128 or 64)
   int smax_m = 0;    // small-front solves: largest m and n of the launch's fronts
   int smax_n = 0;
};

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

}  // namespace

// Switches that shape the factor program (extension options).
struct PlanKnobs {
   bool fuse_small = false;  // one-workgroup fronts assembled inside their factor launch
   bool strip_trsm = false;  // panel TRSM by substitution after a factor-only diagonal launch
   bool level_asm = false;   // a level's assembly in one launch by receiving columns
   int asm_elems = kAsmLdsBytes / 4;  // entries of an assembly run staged in LDS
};
PlanKnobs plan_knobs(int prec);

struct IndefState;  // APP LDL^T factor state (indef.hip)
void destroy_indef(IndefState* s);
struct DevicePlan;
struct DeviceFactors;
// Indefinite path (indef.hip): factorization driven level by level, solves.
void factor_indef(const Analysis& a, const DevicePlan* P, DeviceFactors* F, const sylver_options_t& opts,
                  FactorResult& res);
void solve_indef(const DevicePlan* P, DeviceFactors* F, int job, const double* b, double* sol, int nrhs,
                 hipStream_t s);

struct DevicePlan {
   int variant_gen = -1;  // Analysis::variant_gen the plan was built with
   int n = 0, nnodes = 0, nlevels = 0;
   int small_max = 256;
   bool fuse_small = false;  // one-workgroup fronts assembled inside their factor launch (L_SMALL_ASM)
   bool strip_trsm = false;  // large-front blocks of few fronts: L_DFACT + L_STRIP (no inverse on the chain)
   bool level_asm = false;   // assembly by receiving columns, one launch per level (L_ASM)
   int asm_elems = 0;        // entries of an L_ASM run staged in LDS
   int wide_ncol = 1024;  // fronts with more pivots take the wide-front update launches
   std::vector<NodeDesc> hnodes;
   std::vector<std::vector<int>> levels;
   std::vector<int> hcmap;
   std::vector<int64_t> hcmap_off;
   int64_t arena_elems = 0, dinv_elems = 0, solve_rows = 0;
   // bit-exact device copies of the analyse outputs
   int* d_sptr = nullptr;
   int* d_sparent = nullptr;
   int64_t* d_rptr = nullptr;
   int* d_rlist = nullptr;
   int64_t* d_nptr = nullptr;
   int64_t* d_nlist = nullptr;
   int* d_cmap = nullptr;
   int64_t* d_cmap_off = nullptr;
   NodeDesc* d_nodes = nullptr;
   // task lists
   std::vector<ColTask> zero;
   std::vector<InitTask> init;
   std::vector<EaTask> ea;
   std::vector<int> nlistv;  // node lists (small fronts, gathers)
   std::vector<DiagTask> diag;
   std::vector<StripTask> strip;
   std::vector<AsmTask> asmt;
   std::vector<int4> asmch;      // child column ranges of the assembly runs
   std::vector<int64_t> hasort;  // each front's A-map pairs sorted by position (L_ASM)
   std::vector<GemmTask> gemm;
   std::vector<SolveTask> stask;
   std::vector<int4> gch;  // forward gather: {child, i0, i1} ranges per task
   std::vector<ChainTask> ctask;
   std::vector<int4> cgch;   // forward gather ranges of the chained tasks
   int64_t chain_flags = 0;  // diagonal blocks of the fronts solved by k_chain (exchange slots)
   ColTask* d_zero = nullptr;
   InitTask* d_init = nullptr;
   EaTask* d_ea = nullptr;
   int* d_nlistv = nullptr;
   DiagTask* d_diag = nullptr;
   StripTask* d_strip = nullptr;
   AsmTask* d_asmt = nullptr;
   int4* d_asmch = nullptr;
   int64_t* d_asort = nullptr;
   GemmTask* d_gemm = nullptr;
   SolveTask* d_stask = nullptr;
   int4* d_gch = nullptr;
   ChainTask* d_ctask = nullptr;
   int4* d_cgch = nullptr;
   std::vector<Launch> fprog, fwdprog, bwdprog;
   int64_t num_flops = 0;
   int64_t cb_slots = 0;  // partial-sum slots of the largest backward CB launch
   int nevents = 0;       // cross-stream events of the factor program
   // CSR of the whole matrix in pivot order, values gathered per factorize
   int64_t csr_nnz = 0;
   int64_t* d_csr_ptr = nullptr;
   int* d_csr_col = nullptr;
   int64_t* d_csr_src = nullptr;
   int64_t clean_nnz = 0;
   // pivot order -> caller's order: invp[i] = 1-based variable at pivot i
   int* d_invp = nullptr;
   // host CSR until the upload
   std::vector<int64_t> hcsr_ptr, hcsr_src;
   std::vector<int> hcsr_col;

   ~DevicePlan() {
      for (void* p : {(void*)d_sptr, (void*)d_sparent, (void*)d_rptr, (void*)d_rlist,
                      (void*)d_nptr, (void*)d_nlist, (void*)d_cmap, (void*)d_cmap_off,
                      (void*)d_nodes, (void*)d_zero, (void*)d_init, (void*)d_ea,
                      (void*)d_nlistv, (void*)d_diag, (void*)d_strip, (void*)d_asmt, (void*)d_asmch, (void*)d_asort, (void*)d_gemm, (void*)d_stask, (void*)d_gch, (void*)d_ctask, (void*)d_cgch,
                      (void*)d_csr_ptr, (void*)d_csr_col, (void*)d_csr_src, (void*)d_invp})
         dfree(p);
   }
};

struct DeviceFactors {
   PlanPtr plan;  // the plan the factors (and their captured graphs) were built with
   int precision = SYLVER_AMD_PREC_FP64;
   bool posdef = true;
   void* arena = nullptr;
   void* dinv = nullptr;
   void* dinvT = nullptr;
   double* d_aval = nullptr;
   double* d_scaling = nullptr;
   double* d_csr_val = nullptr;
   int* d_info = nullptr;
   double anorm = 0;
   hipStream_t stream = nullptr;
   hipStream_t stream2 = nullptr;    // look-ahead: far trailing updates
   std::vector<hipEvent_t> xev;      // cross-stream events of the factor program
   hipGraphExec_t graph = nullptr;
   bool graph_scaled = false;  // the captured A-scatter applies the scaling
   bool has_scaling = false;
   // solve workspaces
   int ws_nrhs = 0;
   double* d_xloc = nullptr;
   double* d_b = nullptr;
   double* d_x = nullptr;
   double* d_r = nullptr;
   double* d_sol = nullptr;
   double* d_xin = nullptr;      // the caller's x (caller's order), staged for the permutation
   double* d_sc = nullptr;       // scaled right-hand sides of the preconditioner (S r)
   double* d_part_cb = nullptr;  // backward CB partial sums: cb_slots x 64 x nrhs
   int* d_chain = nullptr;       // k_chain: count of waits that gave up
   double* d_ybuf = nullptr;     // k_chain: exchange buffer of the diagonal blocks of y
   unsigned long long* d_norms = nullptr;
   double* d_xrm = nullptr;      // refinement: x in row-major 16-column chunks
   double* d_maxpart = nullptr;  // refinement: per-workgroup column maxima of x and r
   // GMRES-IR workspace (allocated on first use)
   int gm_nrhs = 0, gm_m = 0;
   double* d_V = nullptr;      // (gm_m + 1) Krylov vectors, n x nrhs each
   double* d_W = nullptr;
   double* d_tmp = nullptr;
   double* d_part = nullptr;
   double* d_coef = nullptr;
   // statistics
   double factor_ms = 0, upload_ms = 0, solve_ms = 0, bwd_err = -1;
   int refine_iters = 0;
   int nlaunches = 0;
   double kind_ms[16] = {0};
   double kind_flops[16] = {0};
   double kind_bytes[16] = {0};
   int kind_launches[16] = {0};
   // indefinite (APP LDL^T) factors; null on the posdef path
   IndefState* ind = nullptr;
   // captured solve programs, keyed by job, buffers and right-hand sides
   struct SolveGraph {
      int job, nrhs, es;
      const double* b;
      double* x;
      hipGraphExec_t exec;
   };
   std::vector<SolveGraph> sgraphs;
   void clear_solve_graphs() {
      for (auto& g : sgraphs) (void)hipGraphExecDestroy(g.exec);
      sgraphs.clear();
   }

   ~DeviceFactors() {
      clear_solve_graphs();
      if (ind) destroy_indef(ind);
      if (graph) (void)hipGraphExecDestroy(graph);
      dfree(d_part_cb);
      for (void* p : {arena, dinv, dinvT, (void*)d_aval, (void*)d_scaling, (void*)d_csr_val,
                      (void*)d_info, (void*)d_xloc, (void*)d_b, (void*)d_x, (void*)d_r,
                      (void*)d_sol, (void*)d_norms, (void*)d_V, (void*)d_W, (void*)d_tmp,
                      (void*)d_part, (void*)d_coef, (void*)d_chain, (void*)d_ybuf,
                      (void*)d_xrm, (void*)d_maxpart, (void*)d_xin, (void*)d_sc})
         dfree(p);
      for (auto e : xev) (void)hipEventDestroy(e);
      if (stream2) (void)hipStreamDestroy(stream2);
      if (stream) (void)hipStreamDestroy(stream);
   }
};


}  // namespace sylver_amd
