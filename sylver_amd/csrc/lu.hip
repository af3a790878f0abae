// Dense front LU of the unsymmetric path, batched on the device: the
// reference's kernels lu_nopiv_factor (src/kernels/lu_nopiv.hxx:13-41),
// lu_pp_factor (src/kernels/lu_pp.hxx:38-73) and lu_tpp_factor
// (src/kernels/lu_tpp.hxx:33-118) over many fronts in one launch.
//
// One workgroup per front, the m x m front (m <= kLuMaxRows) held in LDS,
// one thread per row. Elimination step at pivot e: every thread reads
// a_ee and forms d = 1 / a_ee, the thread of row i > e forms l = a_ie d,
// stores it and updates its row, a_ij = fma(-l, a_ej, a_ij) for j > e --
// row e is not written in the step, so one barrier per step suffices. Pivot
// searches are workgroup reductions of (|a_ic|, i) keeping the largest
// value and, on ties, the smaller row (the reference's find_col_abs_max,
// include/sylver/kernels/common.hxx:63-104, keeps the first row reaching
// the maximum). Row and column interchanges are parallel over the other
// index. The arithmetic and its order are those of the CPU restatement in
// the oracle, so the factors agree bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "numeric.hpp"

namespace sylver_amd {

namespace {

constexpr int kLuMaxRows = 128;
constexpr int kLuLd = kLuMaxRows + 1;

struct LuTask {
   int64_t off;   // the front's first element in the caller's buffer (ld = m)
   int m, k;      // order and columns to eliminate
   int64_t perm;  // offset of the front's rperm / cperm entries
};

template <typename T>
__device__ inline T lu_abs(T v) {
   return v < T(0) ? -v : v;
}

template <typename T>
__global__ __launch_bounds__(kLuMaxRows) void k_lu_front(const LuTask* __restrict__ tasks, T* __restrict__ data,
                                                         int mode, double u, int* __restrict__ rperm,
                                                         int* __restrict__ cperm, int* __restrict__ nelim_out) {
   __shared__ T A[kLuMaxRows * kLuLd];
   __shared__ int rp[kLuMaxRows], cp[kLuMaxRows];
   __shared__ T red_v[2];
   __shared__ int red_i[2];
   const LuTask t = tasks[blockIdx.x];
   const int m = t.m, k = t.k;
   T* F = data + t.off;
   const int i = threadIdx.x, lane = i & 63, wave = i >> 6;
   for (int e = i; e < m * m; e += kLuMaxRows) A[(e % m) + (e / m) * kLuLd] = F[e];
   if (i < m) rp[i] = cp[i] = i;
   __syncthreads();
   auto eliminate = [&](int e) {
      if (i > e && i < m) {
         const T d = T(1) / A[e + e * kLuLd];
         const T l = A[i + e * kLuLd] * d;
         A[i + e * kLuLd] = l;
         for (int j = e + 1; j < m; ++j) A[i + j * kLuLd] = fma(-l, A[e + j * kLuLd], A[i + j * kLuLd]);
      }
      __syncthreads();
   };
   // largest |a_ic| over rows [from, to], the smaller row on ties
   auto col_abs_max = [&](int c, int from, int to, int& idx, T& val) {
      T v = (i >= from && i <= to) ? lu_abs(A[i + c * kLuLd]) : T(-1);
      int ix = i;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
         const T v2 = __shfl_xor(v, o);
         const int i2 = __shfl_xor(ix, o);
         if (v2 > v || (v2 == v && i2 < ix)) {
            v = v2;
            ix = i2;
         }
      }
      if (lane == 0) {
         red_v[wave] = v;
         red_i[wave] = ix;
      }
      __syncthreads();
      T v0 = red_v[0];
      int i0 = red_i[0];
      if (red_v[1] > v0 || (red_v[1] == v0 && red_i[1] < i0)) {
         v0 = red_v[1];
         i0 = red_i[1];
      }
      idx = i0;
      val = v0;
      __syncthreads();  // red_v / red_i are reused
   };
   // rows r1 <-> r2 and columns c1 <-> c2 over the whole front, with the orders
   auto swap_rc = [&](int r1, int r2, int c1, int c2) {
      if (r1 != r2 && i < m) {
         const T x = A[r1 + i * kLuLd];
         A[r1 + i * kLuLd] = A[r2 + i * kLuLd];
         A[r2 + i * kLuLd] = x;
      }
      __syncthreads();
      if (c1 != c2 && i < m) {
         const T x = A[i + c1 * kLuLd];
         A[i + c1 * kLuLd] = A[i + c2 * kLuLd];
         A[i + c2 * kLuLd] = x;
      }
      if (i == 0) {
         const int a = rp[r1];
         rp[r1] = rp[r2];
         rp[r2] = a;
         const int b = cp[c1];
         cp[c1] = cp[c2];
         cp[c2] = b;
      }
      __syncthreads();
   };
   int nelim = 0;
   if (mode == 0) {
      for (int e = 0; e < k; ++e) eliminate(e);
      nelim = k;
   } else if (mode == 1) {
      for (int e = 0; e < k; ++e) {
         int p;
         T v;
         col_abs_max(e, e, m - 1, p, v);
         swap_rc(e, p, e, e);
         eliminate(e);
      }
      nelim = k;
   } else {
      const T tu = T(u);
      for (int c = 0; c < k;) {
         int idx;
         T maxc;
         col_abs_max(c, nelim, m - 1, idx, maxc);
         int r = -1;
         if (lu_abs(A[c + c * kLuLd]) >= tu * maxc) {
            r = c;
         } else {
            int p;
            T maxp;
            col_abs_max(c, nelim, k - 1, p, maxp);
            if (lu_abs(A[p + c * kLuLd]) >= tu * maxc) r = p;
         }
         if (r < 0) {
            ++c;
            continue;
         }
         swap_rc(nelim, r, nelim, c);
         eliminate(nelim);
         ++nelim;
         c = nelim;
      }
   }
   for (int e = i; e < m * m; e += kLuMaxRows) F[e] = A[(e % m) + (e / m) * kLuLd];
   if (i < m) {
      rperm[t.perm + i] = rp[i];
      cperm[t.perm + i] = cp[i];
   }
   if (i == 0) nelim_out[blockIdx.x] = nelim;
}

}  // namespace

int batched_front_lu(int nfronts, const int* m, const int* k, const int64_t* off, void* data, int dtype, int mode,
                     double u, int* rperm, int* cperm, int* nelim, void* stream) {
   if (nfronts < 0 || (nfronts > 0 && (!m || !k || !off || !data || !rperm || !cperm || !nelim)) || mode < 0 ||
       mode > 2 || (dtype != 0 && dtype != 1))
      return SYLVER_ERROR_CALL_SEQUENCE;
   if (nfronts == 0) return 0;
   std::vector<LuTask> tasks(nfronts);
   int64_t perm = 0;
   for (int f = 0; f < nfronts; ++f) {
      if (m[f] < 1 || m[f] > kLuMaxRows || k[f] < 0 || k[f] > m[f] || off[f] < 0)
         return SYLVER_ERROR_CALL_SEQUENCE;
      tasks[f] = LuTask{off[f], m[f], k[f], perm};
      perm += m[f];
   }
   int dev_count = 0;
   if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1) return SYLVER_ERROR_CUDA_UNKNOWN;
   hipStream_t s = static_cast<hipStream_t>(stream);
   LuTask* dt = nullptr;
   int *drp = nullptr, *dcp = nullptr, *dne = nullptr;
   int rc = 0;
   auto ok = [&](hipError_t e) {
      if (e != hipSuccess && rc == 0) rc = SYLVER_ERROR_CUDA_UNKNOWN;
      return rc == 0;
   };
   hipEvent_t e0 = nullptr, e1 = nullptr;
   if (ok(hipMalloc(&dt, nfronts * sizeof(LuTask))) && ok(hipMalloc(&drp, perm * sizeof(int))) &&
       ok(hipMalloc(&dcp, perm * sizeof(int))) && ok(hipMalloc(&dne, nfronts * sizeof(int))) &&
       ok(hipEventCreate(&e0)) && ok(hipEventCreate(&e1)) &&
       ok(hipMemcpyAsync(dt, tasks.data(), nfronts * sizeof(LuTask), hipMemcpyHostToDevice, s)) &&
       ok(hipEventRecord(e0, s))) {
      if (dtype == 0)
         k_lu_front<double><<<nfronts, kLuMaxRows, 0, s>>>(dt, static_cast<double*>(data), mode, u, drp, dcp, dne);
      else
         k_lu_front<float><<<nfronts, kLuMaxRows, 0, s>>>(dt, static_cast<float*>(data), mode, u, drp, dcp, dne);
      if (ok(hipGetLastError()) && ok(hipEventRecord(e1, s)) &&
          ok(hipMemcpyAsync(rperm, drp, perm * sizeof(int), hipMemcpyDeviceToHost, s)) &&
          ok(hipMemcpyAsync(cperm, dcp, perm * sizeof(int), hipMemcpyDeviceToHost, s)) &&
          ok(hipMemcpyAsync(nelim, dne, nfronts * sizeof(int), hipMemcpyDeviceToHost, s)) &&
          ok(hipStreamSynchronize(s))) {
         float ms = 0;
         if (ok(hipEventElapsedTime(&ms, e0, e1))) set_last_batched_ms(ms);  // the kernel alone
      }
   }
   for (hipEvent_t e : {e0, e1})
      if (e) (void)hipEventDestroy(e);
   for (void* p : {(void*)dt, (void*)drp, (void*)dcp, (void*)dne})
      if (p) (void)hipFree(p);
   return rc;
}

}  // namespace sylver_amd
