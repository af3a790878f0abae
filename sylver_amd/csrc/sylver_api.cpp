// C ABI of the library: the reference's public entry points
// (include/sylver/sylver.h:73-123, implemented there in Fortran in
// src/interfaces/C/sylver_ciface.F90:304-826 and the modules it drives),
// re-expressed in C++ over this library's CPU symbolic analysis and device
// numeric phase. Nothing here throws across the boundary.
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/sylver/sylver.h"
#include "../../include/sylver_amd.h"
#include "analyse.hpp"
#include "numeric.hpp"
#include "scaling.hpp"

using namespace sylver_amd;

namespace {

struct Akeep {
   uint32_t magic = 0x414b4550;  // "AKEP"
   Analysis a;
   PlanPtr plan;  // shared with the factorizations built on it
   ~Akeep() { magic = 0; }
};

struct Fkeep {
   uint32_t magic = 0x464b4550;  // "FKEP"
   Akeep* ak = nullptr;
   DeviceFactors* f = nullptr;
   bool posdef = true;
   bool factored = false;
   ~Fkeep() {
      destroy_factors(f);
      f = nullptr;
      magic = 0;
   }
};

int g_ncpu = 1, g_ngpu = 1;
bool g_init = false;

Akeep* as_akeep(void* p) {
   Akeep* a = static_cast<Akeep*>(p);
   return (a && a->magic == 0x414b4550) ? a : nullptr;
}
Fkeep* as_fkeep(void* p) {
   Fkeep* f = static_cast<Fkeep*>(p);
   return (f && f->magic == 0x464b4550) ? f : nullptr;
}

// Numeric-tree seam objects (src/SymbolicTree.cxx:110-137,
// src/NumericTreePosdef.cxx:19-83, src/NumericTree.cxx:27-136).
struct SeamSymbolic {
   Akeep ak;  // holds an Analysis rebuilt from the caller's arrays
};
struct SeamNumeric {
   uint32_t magic = 0x534e554d;  // "SNUM"
   SeamSymbolic* sym = nullptr;
   DeviceFactors* f = nullptr;
   bool posdef = true;
   ~SeamNumeric() {
      destroy_factors(f);
      magic = 0;
   }
};

// The indefinite factors behind a handle: an fkeep of spldlt_factorize or a
// numeric tree of spldlt_create_numeric_tree_dbl (posdef = false).
const DeviceFactors* indef_factors_of(const void* p) {
   if (!p) return nullptr;
   const uint32_t magic = *static_cast<const uint32_t*>(p);
   if (magic == 0x464b4550) {
      const Fkeep* fk = static_cast<const Fkeep*>(p);
      return (fk->f && !fk->posdef) ? fk->f : nullptr;
   }
   if (magic == 0x534e554d) {
      const SeamNumeric* t = static_cast<const SeamNumeric*>(p);
      return (t->f && !t->posdef) ? t->f : nullptr;
   }
   return nullptr;
}

void clear_inform(sylver_inform_t* inform) {
   if (!inform) return;
   std::memset(inform, 0, sizeof(*inform));
}

void write_tail(sylver_inform_t* inform, const sylver_amd_inform_tail_t& t) {
   static_assert(sizeof(sylver_amd_inform_tail_t) <= sizeof(((sylver_inform_t*)0)->unused),
                 "inform tail must fit the reserved bytes");
   std::memcpy(inform->unused, &t, sizeof(t));
}

void read_tail(const sylver_inform_t* inform, sylver_amd_inform_tail_t& t) {
   std::memcpy(&t, inform->unused, sizeof(t));
}

}  // namespace

extern "C" {

void sylver_init(int ncpu, int ngpu) {
   g_ncpu = ncpu;
   g_ngpu = ngpu;
   g_init = true;
}

void sylver_finalize(void) {
   device_sync();
   g_init = false;
}

// Defaults: src/sylver_datatypes_mod.F90:97-198.
void sylver_default_options(sylver_options_t* o) {
   if (!o) return;
   std::memset(o, 0, sizeof(*o));
   o->array_base = 0;
   o->print_level = 0;
   o->unit_diagnostics = 6;
   o->unit_error = 6;
   o->unit_warning = 6;
   o->ordering = 1;
   o->nemin = 32;
   o->prune_tree = true;
   o->min_gpu_work = 5000000000L;
   o->scaling = 0;
   o->pivot_method = 2;
   o->small = 1e-20;
   o->u = 0.01;
   o->small_subtree_threshold = 4000000L;
   o->nb = 256;
   o->cpu_topology = 1;
   o->action = true;
   o->use_gpu = true;
   o->gpu_perf_coeff = 1.0;
   o->failed_pivot_method = 1;
   o->scheduler = 1;
}

// src/spldlt_analyse_mod.F90:587-890 (spldlt_analyse): check, order,
// analyse; akeep created if null, replaced otherwise.
void spldlt_analyse(int n, int* order, long const* ptr, int const* row, double const* val,
                    void** akeep, bool check, sylver_options_t const* options,
                    sylver_inform_t* inform) {
   sylver_inform_t dummy;
   if (!inform) inform = &dummy;
   clear_inform(inform);
   if (!akeep || !options) {
      inform->flag = SYLVER_ERROR_CALL_SEQUENCE;
      return;
   }
   if (*akeep) {
      Akeep* old = as_akeep(*akeep);
      delete old;
      *akeep = nullptr;
   }
   Akeep* ak = nullptr;
   try {
      ak = new Akeep();
      int nemin = options->nemin < 1 ? 32 : options->nemin;
      // ordering 0: the caller's order; 1: nested dissection; 2: matching-based
      // (src/spldlt_analyse_mod.F90:740-820)
      int flag = analyse(n, options->ordering == 0 ? order : nullptr, reinterpret_cast<const int64_t*>(ptr),
                         row, val, check, options->ordering, nemin, ak->a);
      inform->flag = flag;
      if (flag < 0) {
         delete ak;
         return;
      }
      const Analysis& a = ak->a;
      inform->matrix_dup = a.matrix_dup;
      inform->matrix_outrange = a.matrix_outrange;
      inform->matrix_missing_diag = a.matrix_missing_diag;
      inform->maxdepth = a.maxdepth;
      inform->maxfront = a.maxfront;
      inform->num_factor = a.num_factor;
      inform->num_flops = a.num_flops;
      inform->num_sup = a.nnodes;
      inform->matrix_rank = n;
      *akeep = ak;
   } catch (const std::bad_alloc&) {
      delete ak;
      inform->flag = SYLVER_ERROR_ALLOCATION;
   } catch (...) {
      delete ak;
      inform->flag = SYLVER_ERROR_UNKNOWN;
   }
}

// src/spldlt_factorize_mod.F90:569-898 (spldlt_factorize) and factor_core
// (:473-567): values through the conversion map, optional user scaling, then
// the numeric tree on the device.
void spldlt_factorize(bool posdef, long const* ptr, int const* row, double const* val,
                      double* scale, void* akeep, void** fkeep,
                      sylver_options_t const* options, sylver_inform_t* inform) {
   (void)ptr;
   (void)row;
   sylver_inform_t dummy;
   if (!inform) inform = &dummy;
   clear_inform(inform);
   Akeep* ak = as_akeep(akeep);
   if (!ak || !fkeep || !options) {
      inform->flag = SYLVER_ERROR_CALL_SEQUENCE;
      return;
   }
   if (!val) {
      inform->flag = SYLVER_ERROR_VAL;
      return;
   }
   const Analysis& a = ak->a;
   try {
      Fkeep* fk = as_fkeep(*fkeep);
      if (!fk) {
         fk = new Fkeep();
         *fkeep = fk;
      }
      fk->ak = ak;
      fk->posdef = posdef;
      fk->factored = false;
      std::vector<double> clean;
      const double* aval = val;
      if (a.has_conv_map) {
         clean = clean_values(a, val);
         aval = clean.data();
      }
      // Scaling (src/spldlt_factorize_mod.F90:707-835), kept in pivot order
      // (fkeep%scaling(i) = scaling(invp(i))); computed scalings are copied
      // to the caller's scale array (caller's order) when one is given.
      //   <= 0  the caller's scale array, when one is given (case(:0))
      //      1  matching-based, shortest augmenting paths (MC64-like)
      //      2  matching-based, auction
      //      3  the matching-based ordering's scaling saved by the analysis
      //   >= 4  norm equilibration (MC77-like), computed on the device
      int warn = 0;
      if (!a.scaling.empty() && options->scaling != 3) warn = SYLVER_WARNING_MATCH_ORD_NO_SCALE;
      std::vector<double> sc, user(a.n);
      const double* scaling = nullptr;
      bool equilib = false;
      auto to_pivot = [&](const double* s_user) {
         sc.resize(a.n);
         for (int i = 0; i < a.n; ++i) sc[i] = s_user[a.invp[i] - 1];
         scaling = sc.data();
      };
      const int mode = options->scaling;
      if (mode <= 0) {
         if (scale) to_pivot(scale);
      } else if (mode == 1 || mode == 2) {
         int st = mode == 1 ? hungarian_scale_sym(a.n, a.ptr.data(), a.row.data(), aval, options->action,
                                                  user.data(), nullptr)
                            : auction_scale_sym(a.n, a.ptr.data(), a.row.data(), aval, user.data(), nullptr);
         if (st == -2) {
            inform->flag = SYLVER_ERROR_SINGULAR;
            return;
         }
         to_pivot(user.data());
         if (scale) std::memcpy(scale, user.data(), a.n * sizeof(double));
      } else if (mode == 3) {
         if (a.scaling.empty()) {
            inform->flag = SYLVER_ERROR_NO_SAVED_SCALING;
            return;
         }
         to_pivot(a.scaling.data());
      } else {
         equilib = true;
         sc.resize(a.n);
      }
      FactorResult r = device_factor(a, &ak->plan, &fk->f, posdef, aval, scaling, *options, equilib,
                                     equilib ? sc.data() : nullptr);
      if (equilib && scale && r.flag >= 0)
         for (int i = 0; i < a.n; ++i) scale[a.invp[i] - 1] = sc[i];
      inform->flag = (r.flag == 0 && warn) ? warn : r.flag;
      inform->cuda_error = r.hip_error;
      inform->maxdepth = a.maxdepth;
      inform->maxfront = a.maxfront;
      inform->num_factor = a.num_factor;
      inform->num_flops = a.num_flops;
      inform->num_sup = a.nnodes;
      inform->num_neg = r.num_neg;
      inform->num_two = r.num_two;
      inform->num_delay = r.num_delay;
      inform->matrix_rank = r.matrix_rank;
      if (!posdef && r.maxfront > 0) {  // actual sizes with delayed columns
         inform->maxfront = r.maxfront;
         inform->num_factor = r.num_factor;
         inform->num_flops = (long)r.num_flops;
      }
      sylver_amd_inform_tail_t t;
      std::memset(&t, 0, sizeof(t));
      t.factor_ms = r.factor_ms;
      t.upload_ms = r.upload_ms;
      t.gflops = r.factor_ms > 0 ? (double)a.num_flops / (r.factor_ms * 1e6) : 0.0;
      t.bwd_err = -1;
      t.precision = ext_options().precision;
      sylver_amd_stats_t st;
      if (fk->f && factors_stats(fk->f, &st) == 0) t.nlevels = st.nlevels;
      write_tail(inform, t);
      fk->factored = r.flag >= 0;
   } catch (const std::bad_alloc&) {
      inform->flag = SYLVER_ERROR_ALLOCATION;
   } catch (...) {
      inform->flag = SYLVER_ERROR_UNKNOWN;
   }
}

// src/spldlt_factorize_mod.F90:1169-1286 (checks) and :901-1060 (permute,
// scale, tree solves, unscale, unpermute).
void spldlt_solve(int job, int nrhs, double* x, int ldx, void* akeep, void* fkeep,
                  sylver_options_t const* options, sylver_inform_t* inform) {
   (void)options;
   sylver_inform_t dummy;
   if (!inform) inform = &dummy;
   sylver_amd_inform_tail_t tail;
   read_tail(inform, tail);
   clear_inform(inform);
   Akeep* ak = as_akeep(akeep);
   Fkeep* fk = as_fkeep(fkeep);
   if (!ak || !fk || !fk->factored || fk->ak != ak) {
      inform->flag = SYLVER_ERROR_CALL_SEQUENCE;
      return;
   }
   if (job < 0 || job > 4) {
      inform->flag = SYLVER_ERROR_JOB_OOR;
      return;
   }
   if (fk->posdef && (job == 2 || job == 4)) {
      inform->flag = SYLVER_ERROR_JOB_OOR;
      return;
   }
   const Analysis& a = ak->a;
   if (nrhs < 1 || ldx < a.n || (!x && a.n > 0)) {
      inform->flag = SYLVER_ERROR_X_SIZE;
      return;
   }
   try {
      // permutation, scaling and their inverses run on the device
      SolveResult s = device_solve(fk->f, job, nrhs, x, ldx, true);
      inform->flag = s.flag;
      inform->cuda_error = s.hip_error;
      tail.solve_ms = s.solve_ms;
      tail.refine_iters = s.refine_iters;
      tail.refine_mode = s.refine_mode;
      tail.bwd_err = s.bwd_err;
      write_tail(inform, tail);
   } catch (const std::bad_alloc&) {
      inform->flag = SYLVER_ERROR_ALLOCATION;
   } catch (...) {
      inform->flag = SYLVER_ERROR_UNKNOWN;
   }
}

void spldlt_free_akeep(void** akeep) {
   if (!akeep) return;
   delete as_akeep(*akeep);
   *akeep = nullptr;
}

void spldlt_free_fkeep(void** fkeep) {
   if (!fkeep) return;
   delete as_fkeep(*fkeep);
   *fkeep = nullptr;
}

// ---------------------------------------------------------------------------
// Extension entry points
// ---------------------------------------------------------------------------

void sylver_amd_default_options(sylver_amd_options_t* o) {
   if (!o) return;
   std::memset(o, 0, sizeof(*o));
   o->precision = SYLVER_AMD_PREC_FP64;
   o->bf16_min_ncol = 1024;
   o->refine = SYLVER_AMD_REFINE_AUTO;
   o->refine_tol = 1e-15;
   o->refine_max_iter = 30;
   o->use_graph = 1;
   o->tile_nb = 0;
   o->small_front_max = 0;
   o->indef_panel_min = 0;
   o->small_assembly = 0;
   o->strip_trsm = 1;
   o->level_assembly = 1;
   o->record_timing = 0;
}

void sylver_amd_set_options(sylver_amd_options_t const* o) {
   if (o) ext_options() = *o;
}

void sylver_amd_get_options(sylver_amd_options_t* o) {
   if (o) *o = ext_options();
}

int sylver_amd_fkeep_stats(void const* fkeep, sylver_amd_stats_t* st) {
   const Fkeep* fk = as_fkeep(const_cast<void*>(fkeep));
   if (!fk || !fk->f || !st) return SYLVER_ERROR_CALL_SEQUENCE;
   return factors_stats(fk->f, st);
}

int sylver_amd_akeep_sizes(void const* akeep, int* n, int* nnodes, long* nrlist, long* nnlist) {
   const Akeep* ak = as_akeep(const_cast<void*>(akeep));
   if (!ak) return SYLVER_ERROR_CALL_SEQUENCE;
   if (n) *n = ak->a.n;
   if (nnodes) *nnodes = ak->a.nnodes;
   if (nrlist) *nrlist = (long)ak->a.rlist.size();
   if (nnlist) *nnlist = (long)(ak->a.nlist.size() / 2);
   return 0;
}

int sylver_amd_akeep_symbolic(void const* akeep, int* order, int* invp, int* sptr, int* sparent,
                              long* rptr, int* rlist, long* nptr, long* nlist) {
   const Akeep* ak = as_akeep(const_cast<void*>(akeep));
   if (!ak) return SYLVER_ERROR_CALL_SEQUENCE;
   const Analysis& a = ak->a;
   auto cp = [](auto* dst, const auto& v) {
      if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
   };
   cp(order, a.order);
   cp(invp, a.invp);
   cp(sptr, a.sptr);
   cp(sparent, a.sparent);
   cp(reinterpret_cast<int64_t*>(rptr), a.rptr);
   cp(rlist, a.rlist);
   cp(reinterpret_cast<int64_t*>(nptr), a.nptr);
   cp(reinterpret_cast<int64_t*>(nlist), a.nlist);
   return 0;
}

int sylver_amd_fkeep_device_symbolic(void const* fkeep, int* sptr, int* sparent, long* rptr,
                                     int* rlist, long* nptr, long* nlist, int* cmap,
                                     long* cmap_off) {
   const Fkeep* fk = as_fkeep(const_cast<void*>(fkeep));
   if (!fk || !fk->ak || !fk->ak->plan) return SYLVER_ERROR_CALL_SEQUENCE;
   return plan_device_symbolic(fk->ak->plan.get(), sptr, sparent, reinterpret_cast<int64_t*>(rptr),
                               rlist, reinterpret_cast<int64_t*>(nptr),
                               reinterpret_cast<int64_t*>(nlist), cmap,
                               reinterpret_cast<int64_t*>(cmap_off));
}

int sylver_amd_fkeep_node_factor(void const* fkeep, int node, double* l) {
   const Fkeep* fk = as_fkeep(const_cast<void*>(fkeep));
   if (!fk || !fk->f) return SYLVER_ERROR_CALL_SEQUENCE;
   return factors_node_factor(fk->f, node, l);
}

int sylver_amd_fkeep_indef_node(void const* fkeep, int node, int* dims, int* rows, double* l,
                                double* dinv) {
   const DeviceFactors* f = indef_factors_of(fkeep);
   if (!f) return SYLVER_ERROR_CALL_SEQUENCE;
   return indef_node(f, node, dims, rows, l, dinv);
}

int sylver_amd_fkeep_indef_stats(void const* fkeep, int* out) {
   const DeviceFactors* f = indef_factors_of(fkeep);
   if (!f || !out) return SYLVER_ERROR_CALL_SEQUENCE;
   return indef_stats(f, out);
}

int sylver_amd_fkeep_node_assembled(void const* fkeep, int node, double* front) {
   const Fkeep* fk = as_fkeep(const_cast<void*>(fkeep));
   if (!fk || !fk->f) return SYLVER_ERROR_CALL_SEQUENCE;
   return factors_node_assembled(fk->f, node, front);
}

int sylver_amd_batched_front_potrf_v(int nfronts, int const* nrow, int const* ncol, long const* offset,
                                     void* data, int dtype, signed char const* variant, int* info) {
   try {
      return batched_front_potrf_v(nfronts, nrow, ncol, reinterpret_cast<const int64_t*>(offset), data, dtype,
                                   variant, info);
   } catch (...) {
      return SYLVER_ERROR_UNKNOWN;
   }
}

int sylver_amd_set_front_variants(void* akeep, signed char const* variant, int nnodes) {
   Akeep* ak = as_akeep(akeep);
   if (!ak) return SYLVER_ERROR_CALL_SEQUENCE;
   if (variant && nnodes != ak->a.nnodes) return SYLVER_ERROR_CALL_SEQUENCE;
   if (variant)
      ak->a.variant.assign(variant, variant + nnodes);
   else
      ak->a.variant.clear();
   ++ak->a.variant_gen;
   return 0;
}

int sylver_amd_time_fronts(int k, int const* m, int const* n, signed char const* variant, int precision,
                           int reps, double* ms) {
   try {
      return time_fronts(k, m, n, variant, precision, reps, ms);
   } catch (...) {
      return SYLVER_ERROR_UNKNOWN;
   }
}

int sylver_amd_mlp_init(float* params, unsigned long long seed) { return mlp_init(params, seed); }

int sylver_amd_mlp_grad(float const* params, float const* x, float const* y, int b, float* grads, float* loss) {
   try {
      return mlp_grad(params, x, y, b, grads, loss);
   } catch (...) {
      return SYLVER_ERROR_UNKNOWN;
   }
}

int sylver_amd_mlp_train(float* params, float const* x, float const* y, int n, int epochs, int batch, float lr,
                         unsigned long long seed, float* loss_hist) {
   try {
      return mlp_train(params, x, y, n, epochs, batch, lr, seed, loss_hist);
   } catch (...) {
      return SYLVER_ERROR_UNKNOWN;
   }
}

int sylver_amd_mlp_predict(float const* params, float const* x, int n, float* out) {
   try {
      return mlp_predict(params, x, n, out);
   } catch (...) {
      return SYLVER_ERROR_UNKNOWN;
   }
}

int sylver_amd_batched_front_fill(int nfronts, int const* nrow, long const* offset, void* data,
                                  int dtype, unsigned long long seed) {
   try {
      return batched_front_fill(nfronts, nrow, reinterpret_cast<const int64_t*>(offset), data, dtype,
                                seed);
   } catch (...) {
      return SYLVER_ERROR_UNKNOWN;
   }
}

int sylver_amd_batched_front_potrf(int nfronts, int const* nrow, int const* ncol,
                                   long const* offset, void* data, int dtype, int* info,
                                   void* stream) {
   try {
      return batched_front_potrf(nfronts, nrow, ncol, reinterpret_cast<const int64_t*>(offset),
                                 data, dtype, info, stream);
   } catch (...) {
      return SYLVER_ERROR_UNKNOWN;
   }
}

int sylver_amd_batched_front_potrf_noalloc(int nfronts, int const* nrow, int const* ncol,
                                           long const* offset, void* data, int dtype, int* info,
                                           void* stream) {
   try {
      return batched_front_potrf_noalloc(nfronts, nrow, ncol, reinterpret_cast<const int64_t*>(offset),
                                         data, dtype, info, stream);
   } catch (...) {
      return SYLVER_ERROR_UNKNOWN;
   }
}

int sylver_amd_batched_front_lu(int nfronts, int const* m, int const* k, long const* offset, void* data, int dtype,
                                int mode, double u, int* rperm, int* cperm, int* nelim, void* stream) {
   try {
      return batched_front_lu(nfronts, m, k, reinterpret_cast<const int64_t*>(offset), data, dtype, mode, u, rperm,
                              cperm, nelim, stream);
   } catch (...) {
      return SYLVER_ERROR_UNKNOWN;
   }
}

int sylver_amd_hungarian_scale_sym(int n, long const* ptr, int const* row, double const* val,
                                   int scale_if_singular, double* scaling, int* match) {
   if (n < 0 || (n > 0 && (!ptr || !row || !val || !scaling))) return SYLVER_ERROR_CALL_SEQUENCE;
   try {
      return hungarian_scale_sym(n, reinterpret_cast<const int64_t*>(ptr), row, val, scale_if_singular != 0,
                                 scaling, match);
   } catch (...) {
      return SYLVER_ERROR_ALLOCATION;
   }
}

int sylver_amd_auction_scale_sym(int n, long const* ptr, int const* row, double const* val, double* scaling,
                                 int* match) {
   if (n < 0 || (n > 0 && (!ptr || !row || !val || !scaling))) return SYLVER_ERROR_CALL_SEQUENCE;
   try {
      return auction_scale_sym(n, reinterpret_cast<const int64_t*>(ptr), row, val, scaling, match);
   } catch (...) {
      return SYLVER_ERROR_ALLOCATION;
   }
}

int sylver_amd_match_order(int n, long const* ptr, int const* row, double const* val, int* order,
                           double* scaling) {
   if (n < 0 || (n > 0 && (!ptr || !row || !val || !order || !scaling))) return SYLVER_ERROR_CALL_SEQUENCE;
   try {
      return match_order(n, reinterpret_cast<const int64_t*>(ptr), row, val, order, scaling);
   } catch (...) {
      return SYLVER_ERROR_ALLOCATION;
   }
}

int sylver_amd_plan_check(void const* akeep, int small_front_max, int bf16_min_ncol, int num_cus, long* out,
                          char* why, int why_len) {
   const Akeep* ak = as_akeep(const_cast<void*>(akeep));
   if (!ak || !out || num_cus < 1) return SYLVER_ERROR_CALL_SEQUENCE;
   try {
      std::string msg;
      int64_t o[10];
      plan_check(ak->a, small_front_max > 0 ? small_front_max : 256, bf16_min_ncol > 0 ? bf16_min_ncol : 1024,
                 num_cus, o, &msg);
      for (int k = 0; k < 10; ++k) out[k] = (long)o[k];
      if (why && why_len > 0) {
         std::strncpy(why, msg.c_str(), (size_t)why_len - 1);
         why[why_len - 1] = 0;
      }
      return 0;
   } catch (const std::exception& e) {
      if (why && why_len > 0) {
         std::strncpy(why, e.what(), (size_t)why_len - 1);
         why[why_len - 1] = 0;
      }
      return SYLVER_ERROR_UNKNOWN;
   }
}

int sylver_amd_device_count(void) { return device_count(); }

int sylver_amd_dev_alloc(void** ptr, size_t bytes) { return dev_alloc(ptr, bytes); }
int sylver_amd_dev_free(void* ptr) { return dev_free(ptr); }
int sylver_amd_memcpy_h2d(void* dst, const void* src, size_t bytes) {
   return dev_copy(dst, src, bytes, true);
}
int sylver_amd_memcpy_d2h(void* dst, const void* src, size_t bytes) {
   return dev_copy(dst, src, bytes, false);
}
double sylver_amd_last_batched_ms(void) { return last_batched_ms(); }

// ---------------------------------------------------------------------------
// Internal numeric-tree seam (src/SymbolicTree.cxx:110-137,
// src/NumericTreePosdef.cxx:19-83, src/NumericTree.cxx:27-136).
// ---------------------------------------------------------------------------
void* spldlt_create_symbolic_tree(void* akeep, int n, int nnodes, int const* sptr,
                                  int const* sparent, long const* rptr, int const* rlist,
                                  long const* nptr, long const* nlist, int nsubtrees,
                                  int const* subtrees, int const* small, int const* contrib_dest,
                                  int const* exec_loc) {
   (void)akeep;
   (void)nsubtrees;
   (void)subtrees;
   (void)small;
   (void)contrib_dest;
   (void)exec_loc;
   try {
      auto* s = new SeamSymbolic();
      Analysis& a = s->ak.a;
      a.n = n;
      a.nnodes = nnodes;
      a.sptr.assign(sptr, sptr + nnodes + 1);
      a.sparent.assign(sparent, sparent + nnodes);
      a.rptr.assign(reinterpret_cast<const int64_t*>(rptr),
                    reinterpret_cast<const int64_t*>(rptr) + nnodes + 1);
      a.rlist.assign(rlist, rlist + (a.rptr[nnodes] - 1));
      a.nptr.assign(reinterpret_cast<const int64_t*>(nptr),
                    reinterpret_cast<const int64_t*>(nptr) + nnodes + 1);
      a.nlist.assign(reinterpret_cast<const int64_t*>(nlist),
                     reinterpret_cast<const int64_t*>(nlist) + 2 * (a.nptr[nnodes] - 1));
      a.invp.resize(n);
      a.order.resize(n);
      for (int i = 0; i < n; ++i) a.invp[i] = a.order[i] = i + 1;
      build_symbolic_tree(a);
      for (int k = 0; k < nnodes; ++k) {
         int64_t m = a.rptr[k + 1] - a.rptr[k];
         int64_t c = a.sptr[k + 1] - a.sptr[k];
         a.num_flops += node_flops(m, c);
         a.num_factor += c * (c + 1) / 2 + c * (m - c);
      }
      return s;
   } catch (...) {
      return nullptr;
   }
}

void spldlt_destroy_symbolic_tree(void* t) { delete static_cast<SeamSymbolic*>(t); }

static void* seam_create(bool posdef, void* symbolic_tree, double* aval, double const* scaling,
                         sylver_options_c* options, sylver_inform_c* inform) {
   sylver_inform_c dummy;
   if (!inform) inform = &dummy;
   std::memset(inform, 0, sizeof(*inform));
   auto* s = static_cast<SeamSymbolic*>(symbolic_tree);
   if (!s || !aval) {
      inform->flag = SYLVER_ERROR_CALL_SEQUENCE;
      return nullptr;
   }
   try {
      sylver_options_t o;
      sylver_default_options(&o);
      if (options) {
         o.print_level = options->print_level;
         o.action = options->action;
         o.small = options->small;
         o.u = options->u;
         o.nb = options->nb;
         o.pivot_method = options->pivot_method;
         o.failed_pivot_method = options->failed_pivot_method;
      }
      auto* t = new SeamNumeric();
      t->sym = s;
      t->posdef = posdef;
      FactorResult r = device_factor(s->ak.a, &s->ak.plan, &t->f, posdef, aval, scaling, o);
      inform->flag = r.flag;
      inform->num_delay = r.num_delay;
      inform->num_neg = r.num_neg;
      inform->num_two = r.num_two;
      inform->num_zero = r.num_zero;
      inform->maxfront = (!posdef && r.maxfront > 0) ? r.maxfront : s->ak.a.maxfront;
      inform->not_first_pass = r.not_first_pass;
      inform->not_second_pass = r.not_second_pass;
      if (r.flag < 0) {
         delete t;
         return nullptr;
      }
      return t;
   } catch (...) {
      inform->flag = SYLVER_ERROR_UNKNOWN;
      return nullptr;
   }
}

void* spldlt_create_numeric_tree_posdef_dbl(void* fkeep, void* symbolic_tree, double* aval,
                                            double const* scaling, void** child_contrib,
                                            sylver_options_c* options, sylver_inform_c* inform) {
   (void)fkeep;
   (void)child_contrib;
   return seam_create(true, symbolic_tree, aval, scaling, options, inform);
}

void* spldlt_create_numeric_tree_dbl(bool posdef, void* fkeep, void* symbolic_tree, double* aval,
                                     double const* scaling, void** child_contrib,
                                     sylver_options_c* options, sylver_inform_c* inform) {
   (void)fkeep;
   (void)child_contrib;
   return seam_create(posdef, symbolic_tree, aval, scaling, options, inform);
}

void spldlt_destroy_numeric_tree_posdef_dbl(void* t) { delete static_cast<SeamNumeric*>(t); }
void spldlt_destroy_numeric_tree_dbl(bool posdef, void* t) {
   (void)posdef;
   delete static_cast<SeamNumeric*>(t);
}

static int seam_solve(const void* tree, int job, int nrhs, double* x, int ldx) {
   auto* t = static_cast<const SeamNumeric*>(tree);
   if (!t || !t->sym || !t->f) return SYLVER_ERROR_CALL_SEQUENCE;
   if (nrhs < 1 || ldx < t->sym->ak.a.n || (!x && t->sym->ak.a.n > 0)) return SYLVER_ERROR_X_SIZE;
   // the numeric tree's solves: x already in pivot order, no scaling (the
   // reference's driver applies both around these calls)
   SolveResult s = device_solve(t->f, job, nrhs, x, ldx, false);
   return s.flag;
}

int spldlt_tree_solve_fwd_posdef_dbl(void const* t, int nrhs, double* x, int ldx) {
   return seam_solve(t, 1, nrhs, x, ldx);
}
int spldlt_tree_solve_bwd_posdef_dbl(void const* t, int nrhs, double* x, int ldx) {
   return seam_solve(t, 3, nrhs, x, ldx);
}
int spldlt_tree_solve_fwd_dbl(bool posdef, void const* t, int nrhs, double* x, int ldx) {
   (void)posdef;
   return seam_solve(t, 1, nrhs, x, ldx);
}
int spldlt_tree_solve_bwd_dbl(bool posdef, void const* t, int nrhs, double* x, int ldx) {
   (void)posdef;
   return seam_solve(t, 3, nrhs, x, ldx);
}
int spldlt_tree_solve_diag_dbl(bool posdef, void const* t, int nrhs, double* x, int ldx) {
   (void)posdef;
   return seam_solve(t, 2, nrhs, x, ldx);
}
int spldlt_tree_solve_diag_bwd_dbl(bool posdef, void const* t, int nrhs, double* x, int ldx) {
   (void)posdef;
   return seam_solve(t, 4, nrhs, x, ldx);
}

}  // extern "C"
