// Device numeric phase: plan (device-resident symbolic structures and the
// level schedule), multifrontal factorization, solves and refinement.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "analyse.hpp"
#include "../../include/sylver_amd.h"

namespace sylver_amd {

struct DevicePlan;     // pattern-dependent device state (one per akeep)
struct DeviceFactors;  // value-dependent device state (one per fkeep)

struct FactorResult {
   int flag = 0;
   int hip_error = 0;
   int num_neg = 0;
   int num_two = 0;
   int num_delay = 0;
   int num_zero = 0;
   int matrix_rank = 0;
   int not_first_pass = 0;
   int not_second_pass = 0;
   int maxfront = 0;          // indefinite path: largest front including delays
   int64_t num_factor = 0;    // indefinite path: entries of L including delays
   double num_flops = 0;      // indefinite path: flops including delays
   double factor_ms = 0;
   double upload_ms = 0;
};

struct SolveResult {
   int flag = 0;          // 0, or SYLVER_AMD_WARNING_REFINE / an error flag
   int hip_error = 0;
   double solve_ms = 0;
   int refine_iters = 0;
   int refine_mode = 0;
   double bwd_err = -1;
};

// Global extension options (process-wide).
sylver_amd_options_t& ext_options();

// Number of HIP devices (0 on a host without a GPU).
int device_count();

// Plan lifetime: shared by the akeep (the plan of its current pattern and
// variants) and every factorization built on it, so that a plan rebuilt for
// new variants never frees the one older factors still use.
using PlanPtr = std::shared_ptr<DevicePlan>;
// Factor lifetime (owned by the fkeep).
void destroy_factors(DeviceFactors* f);

// Numeric factorization on the device. `aval` are the cleaned matrix values
// (host, FP64, addressed by nlist); `scaling` (pivot order, may be null) is
// applied as S A S in the A-scatter. equilib: the scaling is instead
// computed on the device by norm equilibration of the values (MC77-like:
// at most 10 sweeps, stop when every scaled row's largest entry is within
// 1e-8 of 1) and, when scaling_out is given, returned there (pivot order).
// Creates or reuses *plan and *fact.
FactorResult device_factor(const Analysis& a, PlanPtr* plan, DeviceFactors** fact,
                           bool posdef, const double* aval, const double* scaling,
                           const sylver_options_t& opts, bool equilib = false,
                           double* scaling_out = nullptr);

// Full solve (job 0) or partial solves (1 fwd, 2 diag, 3 bwd, 4 diag+bwd) of
// nrhs right-hand sides on the host (column-major, ld = ldx), with the
// factors' own plan. user_order: x is in the caller's variable order and the
// permutation to pivot order, the scaling (x2 = S x before the forward solve,
// x = S x2 after the backward one, as the reference's solve driver does) and
// the inverse permutation run on the device; otherwise x is in pivot order
// and neither is applied (the internal seam). With low-precision factors
// and job 0, FP64 refinement against the original (unscaled) matrix is
// applied, preconditioned by S (L L^T)^{-1} S.
SolveResult device_solve(DeviceFactors* fact, int job, int nrhs, double* x, int ldx, bool user_order);

// Builds the host side of the plan of `a` (no device needed) and checks
// every index its launches will use against the buffers they address.
// out[8] = {violations, factor launches, forward-solve launches,
// backward-solve launches, arena elements, exchange slots, levels,
// partial-sum slots}; *why describes the first violation.
int plan_check(const Analysis& a, int small_max, int wide_ncol, int num_cus, int64_t* out, std::string* why);

// Waits for all device work of the process's current device (no-op without a
// device).
void device_sync();

// Statistics and read-back.
int factors_stats(const DeviceFactors* f, sylver_amd_stats_t* st);
int plan_device_symbolic(const DevicePlan* plan, int* sptr, int* sparent, int64_t* rptr,
                         int* rlist, int64_t* nptr, int64_t* nlist, int* cmap,
                         int64_t* cmap_off);
int factors_node_factor(const DeviceFactors* f, int node, double* l);
int factors_node_assembled(const DeviceFactors* f, int node, double* front);

// Indefinite factors: statistics {num_delay, num_neg, num_two, num_zero,
// maxfront, not_first_pass, not_second_pass, 0} and per-node read-back.
int indef_stats(const DeviceFactors* F, int* out);
int indef_node(const DeviceFactors* F, int node, int* dims, int* rows, double* L, double* dinv);

int dev_alloc(void** ptr, size_t bytes);
int dev_free(void* ptr);
int dev_copy(void* dst, const void* src, size_t bytes, bool h2d);
double last_batched_ms();
void set_last_batched_ms(double ms);

int batched_front_potrf(int nfronts, const int* nrow, const int* ncol, const int64_t* off,
                        void* data, int dtype, int* info, void* stream);
int batched_front_potrf_v(int nfronts, const int* nrow, const int* ncol, const int64_t* off, void* data,
                          int dtype, const signed char* variant, int* info);
// batched_front_potrf without allocations: sizes and offsets staged in a
// fixed per-device array of kNoAllocMaxFronts fronts.
constexpr int kNoAllocMaxFronts = 1 << 16;
constexpr int kMaxScratchDevices = 64;
// Dense front LU (unsymmetric path), batched: lu.hip.
int batched_front_lu(int nfronts, const int* m, const int* k, const int64_t* off, void* data, int dtype, int mode,
                     double u, int* rperm, int* cperm, int* nelim, void* stream);
int batched_front_potrf_noalloc(int nfronts, const int* nrow, const int* ncol, const int64_t* off, void* data,
                                int dtype, int* info, void* stream);
int time_fronts(int K, const int* m, const int* n, const signed char* variant, int precision, int reps,
                double* ms);
// Cost model (costmodel.hip): an 8 -> 64 -> 64 -> 2 MLP in FP32.
constexpr int kMlpIn = 8, kMlpHid = 64, kMlpOut = 2;
constexpr int kMlpParams = kMlpIn * kMlpHid + kMlpHid + kMlpHid * kMlpHid + kMlpHid + kMlpHid * kMlpOut + kMlpOut;
int mlp_init(float* params, unsigned long long seed);
int mlp_grad(const float* params, const float* X, const float* Y, int B, float* grads, float* loss);
int mlp_train(float* params, const float* X, const float* Y, int N, int epochs, int batch, float lr,
              unsigned long long seed, float* loss_hist);
int mlp_predict(const float* params, const float* X, int N, float* out);
int batched_front_fill(int nfronts, const int* nrow, const int64_t* off, void* data, int dtype,
                       unsigned long long seed);

}  // namespace sylver_amd
