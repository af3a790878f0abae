// Hand-written CDNA4 (gfx950) kernels of the multifrontal numeric phase.
//
// Front storage: each front is a column-major nrow x nrow block (ld = nrow)
// in one arena; columns [0, ncol) end up holding L, the trailing
// (nrow-ncol)^2 block is the contribution block (CB). Only the lower
// triangle is meaningful.
//
// All kernels use 256 threads (4 wave64) and take a device task list; one
// workgroup per task. Reductions that must be reproducible never use
// atomics: the extend-add is split into launches in which every parent gets
// at most one child (see NumericTreePosdef.hxx:228-268 in the reference for
// the serialized order this reproduces).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <cstdint>

namespace sylver_amd {

constexpr int kThreads = 256;
constexpr int kNB = 128;      // block size of the large-front path and of the solve
constexpr int kSB = 16;       // inner block of the one-workgroup front kernel
constexpr int kMaxRhsChunk = 16;

struct NodeDesc {
   int m;            // rows of the front
   int n;            // pivots of the front
   int parent;       // 0-based, nnodes for the virtual root
   int ld;           // leading dimension of the front (== m)
   int64_t off;      // offset of the front in the arena (elements)
   int64_t rlist_off;
   int64_t amap_off; // pair index into nlist
   int64_t num_a;
   int64_t cmap_off; // offset of this node's map into its parent (q entries)
   int64_t solve_off;// offset (elements) of the node's solve vector (m x rhs chunk)
   int64_t dinv_off; // offset of the inverted diagonal blocks
   int first_child;
   int next_child;
};

struct ColTask { int node, c0, c1, pad; };
struct InitTask { int node, pad; int64_t p0, p1; };
struct EaTask { int child, j0, j1, pad; };
struct DiagTask { int node, k0, kb, pad; };
struct GemmTask {
   int64_t a_off, b_off, c_off;
   int lda, ldb, ldc;
   int M, N, K;
   int pad0, pad1;
};
struct SolveTask { int node, k0, kb, r0, r1, write; };

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
// Operand layout of the 16x16x4 variants: lane l holds A[l%16][l/16] and
// B[l/16][l%16]. Output item v of lane l is D[row(l, v)][l%16]; the FP64
// variant interleaves rows (l/16 + 4v), the FP32 one keeps them contiguous
// (4(l/16) + v) -- both measured on gfx950.
template <> struct Mfma<double> {
   typedef v4d acc_t;
   __device__ static inline acc_t op(double a, double b, acc_t c) {
      return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
   }
   __device__ static inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct Mfma<float> {
   typedef v4f acc_t;
   __device__ static inline acc_t op(float a, float b, acc_t c) {
      return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
   }
   __device__ static inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

__device__ inline double dsqrt(double x) { return __builtin_sqrt(x); }
__device__ inline float dsqrt(float x) { return __builtin_sqrtf(x); }

// ---------------------------------------------------------------------------
// Zero the lower triangle of the columns [c0, c1) of a front.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_zero(const ColTask* __restrict__ tasks,
                                                   const NodeDesc* __restrict__ nodes,
                                                   T* __restrict__ arena) {
   const ColTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   T* F = arena + nd.off;
   for (int c = t.c0 + wave; c < t.c1; c += 4) {
      T* col = F + (int64_t)c * nd.ld;
      for (int r = c + lane; r < nd.m; r += 64) col[r] = T(0);
   }
}

// ---------------------------------------------------------------------------
// Scatter the original entries into the fronts: restates init_a_block
// (reference src/kernels/assemble.hxx:161-214): amap holds 1-based pairs
// (source index in aval, (col)*nrow + row).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_init_a(const InitTask* __restrict__ tasks,
                                                     const NodeDesc* __restrict__ nodes,
                                                     const int64_t* __restrict__ nlist,
                                                     const int* __restrict__ rlist,
                                                     const double* __restrict__ aval,
                                                     const double* __restrict__ scaling,
                                                     T* __restrict__ arena) {
   const InitTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   T* F = arena + nd.off;
   for (int64_t p = t.p0 + threadIdx.x; p < t.p1; p += kThreads) {
      const int64_t src = nlist[2 * (nd.amap_off + p)] - 1;
      const int64_t dest = nlist[2 * (nd.amap_off + p) + 1] - 1;
      const int c = (int)(dest / nd.m);
      const int r = (int)(dest - (int64_t)c * nd.m);
      double v = aval[src];
      if (scaling)
         v = scaling[rlist[nd.rlist_off + r] - 1] * v * scaling[rlist[nd.rlist_off + c] - 1];
      F[(int64_t)c * nd.ld + r] = (T)v;
   }
}

// ---------------------------------------------------------------------------
// Extend-add of one child's contribution block into its parent front:
// parent[cmap[i], cmap[j]] += child_cb[i, j] for i >= j (reference
// assemble_block / assemble_contrib_block, src/kernels/assemble.hxx:230-517).
// Columns [j0, j1) of the child's CB per task. No atomics: the launch holds at
// most one child per parent and the map is injective within a child.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_extend_add(const EaTask* __restrict__ tasks,
                                                         const NodeDesc* __restrict__ nodes,
                                                         const int* __restrict__ cmap,
                                                         T* __restrict__ arena) {
   const EaTask t = tasks[blockIdx.x];
   const NodeDesc c = nodes[t.child];
   const NodeDesc p = nodes[c.parent];
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const int q = c.m - c.n;
   const int* __restrict__ cm = cmap + c.cmap_off;
   for (int j = t.j0 + wave; j < t.j1; j += 4) {
      const int pj = cm[j];
      T* dst = arena + p.off + (int64_t)pj * p.ld;
      const T* src = arena + c.off + (int64_t)(c.n + j) * c.ld + c.n;
      for (int i = j + lane; i < q; i += 64) dst[cm[i]] += src[i];
   }
}

// ---------------------------------------------------------------------------
// In-LDS dense kernels on a kb x kb block (kb <= kNB), stored column-major
// with leading dimension LD.
// ---------------------------------------------------------------------------
// Unblocked right-looking Cholesky. Returns 0 or the 1-based column of the
// first non-positive (or NaN) pivot. All threads must call it.
template <typename T, int LD>
__device__ int lds_potrf(T* L, int kb) {
   const int tid = threadIdx.x;
   for (int j = 0; j < kb; ++j) {
      __syncthreads();
      T d = L[j + j * LD];
      if (!(d > T(0))) return j + 1;
      d = dsqrt(d);
      __syncthreads();
      if (tid == 0) L[j + j * LD] = d;
      for (int i = j + 1 + tid; i < kb; i += kThreads) L[i + j * LD] /= d;
      __syncthreads();
      // trailing update: rows i (< kNB) on tid % kNB, columns strided
      const int i = j + 1 + (tid % kNB);
      if (i < kb) {
         const T lij = L[i + j * LD];
         for (int c = j + 1 + tid / kNB; c <= i; c += kThreads / kNB)
            L[i + c * LD] -= lij * L[c + j * LD];
      }
   }
   __syncthreads();
   return 0;
}

// In-place inverse of a lower triangular kb x kb block (upper part set to 0).
// tmp holds kNB elements.
template <typename T, int LD>
__device__ void lds_trtri(T* L, T* tmp, int kb) {
   const int tid = threadIdx.x;
   __syncthreads();
   for (int j = kb - 1; j >= 0; --j) {
      if (tid < kb - j - 1) tmp[tid] = L[j + 1 + tid + j * LD];
      __syncthreads();
      const T ajj = T(1) / L[j + j * LD];
      for (int i = j + 1 + tid; i < kb; i += kThreads) {
         T s = T(0);
         for (int c = j + 1; c <= i; ++c) s += L[i + c * LD] * tmp[c - j - 1];
         L[i + j * LD] = -ajj * s;
      }
      __syncthreads();
      if (tid == 0) L[j + j * LD] = ajj;
      __syncthreads();
   }
   for (int e = tid; e < kb * kb; e += kThreads) {
      int i = e % kb, j = e / kb;
      if (i < j) L[i + j * LD] = T(0);
   }
   __syncthreads();
}

// ---------------------------------------------------------------------------
// Diagonal block of the large-front path: POTRF of the kb x kb block at
// (k0, k0), L written back, its inverse written to dinv (ld = kb) for the
// TRSM-by-GEMM and the solves.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_diag(const DiagTask* __restrict__ tasks,
                                                   const NodeDesc* __restrict__ nodes,
                                                   T* __restrict__ arena, T* __restrict__ dinv,
                                                   int* __restrict__ info, int do_potrf) {
   constexpr int LD = kNB + 1;
   __shared__ T L[kNB * LD];
   __shared__ T tmp[kNB];
   const DiagTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   const int kb = t.kb;
   T* blk = arena + nd.off + t.k0 + (int64_t)t.k0 * nd.ld;
   for (int e = threadIdx.x; e < kb * kb; e += kThreads) {
      int i = e % kb, j = e / kb;
      L[i + j * LD] = (i >= j) ? blk[i + (int64_t)j * nd.ld] : T(0);
   }
   __syncthreads();
   if (do_potrf) {
      int bad = lds_potrf<T, LD>(L, kb);
      if (bad) {
         if (threadIdx.x == 0) info[t.node] = t.k0 + bad;
         return;
      }
      for (int e = threadIdx.x; e < kb * kb; e += kThreads) {
         int i = e % kb, j = e / kb;
         if (i >= j) blk[i + (int64_t)j * nd.ld] = L[i + j * LD];
      }
   }
   lds_trtri<T, LD>(L, tmp, kb);
   T* out = dinv + nd.dinv_off + (int64_t)(t.k0 / kNB) * kNB * kNB;
   for (int e = threadIdx.x; e < kb * kb; e += kThreads) {
      int i = e % kb, j = e / kb;
      out[i + j * kb] = L[i + j * LD];
   }
}

// ---------------------------------------------------------------------------
// C = beta*C + alpha*A*B^T on one BM x BN tile per workgroup (MFMA 16x16x4,
// 2x2 waves). Used for the TRSM of a panel by the inverted diagonal block
// (in place: BN covers all columns, so each workgroup owns its rows) and for
// the Schur-complement update of the trailing block (lower tiles only).
// ---------------------------------------------------------------------------
template <typename T, int BM, int BN>
__global__ __launch_bounds__(kThreads) void k_gemm_nt(const GemmTask* __restrict__ tasks,
                                                      const T* A, const T* B, T* C, T alpha,
                                                      T beta) {
   constexpr int BK = 16;
   constexpr int PAD = 4;
   constexpr int WM = BM / 2, WN = BN / 2;
   constexpr int FM = WM / 16, FN = WN / 16;
   typedef typename Mfma<T>::acc_t acc_t;
   __shared__ T As[BK][BM + PAD];
   __shared__ T Bs[BK][BN + PAD];
   const GemmTask t = tasks[blockIdx.x];
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int wm = wave >> 1, wn = wave & 1;
   acc_t acc[FM][FN];
   for (int a = 0; a < FM; ++a)
      for (int b = 0; b < FN; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
   const T* a = A + t.a_off;
   const T* b = B + t.b_off;
   for (int k0 = 0; k0 < t.K; k0 += BK) {
      for (int e = tid; e < BK * BM; e += kThreads) {
         int i = e % BM, kk = e / BM;
         T v = T(0);
         if (i < t.M && k0 + kk < t.K) v = a[i + (int64_t)(k0 + kk) * t.lda];
         As[kk][i] = v;
      }
      for (int e = tid; e < BK * BN; e += kThreads) {
         int j = e % BN, kk = e / BN;
         T v = T(0);
         if (j < t.N && k0 + kk < t.K) v = b[j + (int64_t)(k0 + kk) * t.ldb];
         Bs[kk][j] = v;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
         T af[FM], bf[FN];
#pragma unroll
         for (int f = 0; f < FM; ++f) af[f] = As[kk + (lane >> 4)][wm * WM + f * 16 + (lane & 15)];
#pragma unroll
         for (int f = 0; f < FN; ++f) bf[f] = Bs[kk + (lane >> 4)][wn * WN + f * 16 + (lane & 15)];
#pragma unroll
         for (int x = 0; x < FM; ++x)
#pragma unroll
            for (int y = 0; y < FN; ++y) acc[x][y] = Mfma<T>::op(af[x], bf[y], acc[x][y]);
      }
      __syncthreads();
   }
   T* c = C + t.c_off;
#pragma unroll
   for (int x = 0; x < FM; ++x)
#pragma unroll
      for (int y = 0; y < FN; ++y) {
         const int j = wn * WN + y * 16 + (lane & 15);
         if (j >= t.N) continue;
#pragma unroll
         for (int v = 0; v < 4; ++v) {
            const int i = wm * WM + x * 16 + Mfma<T>::row(lane, v);
            if (i >= t.M) continue;
            T* p = c + i + (int64_t)j * t.ldc;
            T old = (beta == T(0)) ? T(0) : *p;
            *p = beta * old + alpha * acc[x][y][v];
         }
      }
}

// ---------------------------------------------------------------------------
// One workgroup factorizes a whole front (partial Cholesky over its first n
// columns, Schur complement into the CB): blocked right-looking with inner
// block kSB; the diagonal block in LDS, the panel TRSM one row per thread,
// the trailing update on MFMA 16x16 tiles read through L2.
// Restates factor_front_posdef's arithmetic (reference src/factor.hxx:19-148,
// src/kernels/factor.hxx:34-263) with a different blocking.
// ---------------------------------------------------------------------------
template <typename T>
__device__ void front_potrf_block(const NodeDesc& nd, T* __restrict__ F, int* info_slot) {
   constexpr int B = kSB;
   __shared__ T D[B][B + 1];
   typedef typename Mfma<T>::acc_t acc_t;
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int m = nd.m, n = nd.n, ld = nd.ld;
   for (int k = 0; k < n; k += B) {
      const int kb = min(B, n - k);
      __syncthreads();
      for (int e = tid; e < kb * kb; e += kThreads) {
         int i = e % kb, j = e / kb;
         D[i][j] = F[(k + i) + (int64_t)(k + j) * ld];
      }
      // diagonal block Cholesky
      for (int j = 0; j < kb; ++j) {
         __syncthreads();
         T d = D[j][j];
         if (!(d > T(0))) {
            if (tid == 0) *info_slot = k + j + 1;
            return;
         }
         d = dsqrt(d);
         __syncthreads();
         if (tid == 0) D[j][j] = d;
         if (tid > j && tid < kb) D[tid][j] /= d;
         __syncthreads();
         {
            const int i = tid & 15, c = tid >> 4;
            if (c > j && i >= c && i < kb) D[i][c] -= D[i][j] * D[c][j];
         }
      }
      __syncthreads();
      for (int e = tid; e < kb * kb; e += kThreads) {
         int i = e % kb, j = e / kb;
         if (i >= j) F[(k + i) + (int64_t)(k + j) * ld] = D[i][j];
      }
      // panel TRSM: rows below the diagonal block, one row per thread
      for (int r = k + kb + tid; r < m; r += kThreads) {
         T x[B];
#pragma unroll
         for (int c = 0; c < B; ++c) x[c] = (c < kb) ? F[r + (int64_t)(k + c) * ld] : T(0);
#pragma unroll
         for (int c = 0; c < B; ++c) {
            if (c < kb) {
               T s = x[c];
#pragma unroll
               for (int cc = 0; cc < c; ++cc) s -= x[cc] * D[c][cc];
               x[c] = s / D[c][c];
            }
         }
#pragma unroll
         for (int c = 0; c < B; ++c)
            if (c < kb) F[r + (int64_t)(k + c) * ld] = x[c];
      }
      __syncthreads();
      // trailing update of the lower triangle of rows/cols [k+kb, m)
      const int r0 = k + kb;
      const int R = m - r0;
      if (R <= 0) continue;
      const int nt = (R + 15) / 16;
      const int ntiles = nt * (nt + 1) / 2;
      const T* P = F + r0 + (int64_t)k * ld;  // panel rows r0.., columns k..k+kb
      for (int tile = wave; tile < ntiles; tile += 4) {
         int ti = (int)((__builtin_sqrtf(8.0f * tile + 1.0f) - 1.0f) * 0.5f);
         while (ti * (ti + 1) / 2 > tile) --ti;
         while ((ti + 1) * (ti + 2) / 2 <= tile) ++ti;
         const int tj = tile - ti * (ti + 1) / 2;
         acc_t acc = acc_t{0, 0, 0, 0};
         const int ia = ti * 16 + (lane & 15);
         const int ib = tj * 16 + (lane & 15);
         for (int kk = 0; kk < kb; kk += 4) {
            const int kc = kk + (lane >> 4);
            T av = (ia < R && kc < kb) ? P[ia + (int64_t)kc * ld] : T(0);
            T bv = (ib < R && kc < kb) ? P[ib + (int64_t)kc * ld] : T(0);
            acc = Mfma<T>::op(av, bv, acc);
         }
         const int j = tj * 16 + (lane & 15);
         if (j < R) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
               const int i = ti * 16 + Mfma<T>::row(lane, v);
               if (i < R && i >= j) {
                  T* p = F + (r0 + i) + (int64_t)(r0 + j) * ld;
                  *p -= acc[v];
               }
            }
         }
      }
   }
   __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_small_front(const int* __restrict__ node_list,
                                                          const NodeDesc* __restrict__ nodes,
                                                          T* __restrict__ arena,
                                                          int* __restrict__ info) {
   const int node = node_list[blockIdx.x];
   const NodeDesc nd = nodes[node];
   front_potrf_block<T>(nd, arena + nd.off, info + node);
}

// ---------------------------------------------------------------------------
// Solves. Vectors are FP64 whatever the factor precision. Each node owns a
// column-major m x nrhs work vector at solve_off (ld = m).
// ---------------------------------------------------------------------------
// Forward gather: own rows from b, CB rows zero, then children's CB parts
// added in child-list order (deterministic).
__global__ __launch_bounds__(kThreads) void k_fwd_gather(const int* __restrict__ node_list,
                                                         const NodeDesc* __restrict__ nodes,
                                                         const int* __restrict__ rlist,
                                                         const int* __restrict__ cmap,
                                                         const double* __restrict__ b, int ldb,
                                                         double* __restrict__ xloc, int nrhs) {
   const int node = node_list[blockIdx.x];
   const NodeDesc nd = nodes[node];
   double* xl = xloc + nd.solve_off * nrhs;
   const int m = nd.m;
   for (int e = threadIdx.x; e < m * nrhs; e += kThreads) {
      int r = e % m, c = e / m;
      xl[r + (int64_t)c * m] = (r < nd.n) ? b[rlist[nd.rlist_off + r] - 1 + (int64_t)c * ldb] : 0.0;
   }
   for (int ch = nd.first_child; ch != -1;) {
      __syncthreads();
      const NodeDesc cn = nodes[ch];
      const int q = cn.m - cn.n;
      const double* cx = xloc + cn.solve_off * nrhs;
      for (int e = threadIdx.x; e < q * nrhs; e += kThreads) {
         int i = e % q, c = e / q;
         xl[cmap[cn.cmap_off + i] + (int64_t)c * m] += cx[cn.n + i + (int64_t)c * cn.m];
      }
      ch = cn.next_child;
   }
}

// Forward step on block [k0, k0+kb): y = inv(L_kk) x_k (written to xsol by the
// task with write=1), then rows [r0, r1) of the work vector -= L[r, k] y.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fwd_step(const SolveTask* __restrict__ tasks,
                                                       const NodeDesc* __restrict__ nodes,
                                                       const T* __restrict__ arena,
                                                       const T* __restrict__ dinv,
                                                       const int* __restrict__ rlist,
                                                       double* __restrict__ xloc,
                                                       double* __restrict__ xsol, int ldsol,
                                                       int nrhs) {
   __shared__ double xs[kNB * kMaxRhsChunk];
   __shared__ double ys[kNB * kMaxRhsChunk];
   const SolveTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   const int m = nd.m, kb = t.kb;
   double* xl = xloc + nd.solve_off * nrhs;
   const T* Li = dinv + nd.dinv_off + (int64_t)(t.k0 / kNB) * kNB * kNB;
   const T* F = arena + nd.off;
   for (int c0 = 0; c0 < nrhs; c0 += kMaxRhsChunk) {
      const int nc = min(kMaxRhsChunk, nrhs - c0);
      __syncthreads();
      for (int e = threadIdx.x; e < kb * nc; e += kThreads) {
         int i = e % kb, c = e / kb;
         xs[i + c * kNB] = xl[t.k0 + i + (int64_t)(c0 + c) * m];
      }
      __syncthreads();
      for (int e = threadIdx.x; e < kb * nc; e += kThreads) {
         int i = e % kb, c = e / kb;
         double s = 0.0;
         for (int l = 0; l <= i; ++l) s += (double)Li[i + l * kb] * xs[l + c * kNB];
         ys[i + c * kNB] = s;
      }
      __syncthreads();
      if (t.write)
         for (int e = threadIdx.x; e < kb * nc; e += kThreads) {
            int i = e % kb, c = e / kb;
            xsol[rlist[nd.rlist_off + t.k0 + i] - 1 + (int64_t)(c0 + c) * ldsol] = ys[i + c * kNB];
         }
      for (int r = t.r0 + threadIdx.x; r < t.r1; r += kThreads) {
         double acc[kMaxRhsChunk];
#pragma unroll
         for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] = 0.0;
         for (int l = 0; l < kb; ++l) {
            const double lv = (double)F[r + (int64_t)(t.k0 + l) * nd.ld];
#pragma unroll
            for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] += lv * ys[l + c * kNB];
         }
#pragma unroll
         for (int c = 0; c < kMaxRhsChunk; ++c)
            if (c < nc) xl[r + (int64_t)(c0 + c) * m] -= acc[c];
      }
   }
}

// Backward gather: every row of the node's work vector from the current
// solution (own rows hold the forward result, CB rows the final values).
__global__ __launch_bounds__(kThreads) void k_bwd_gather(const int* __restrict__ node_list,
                                                         const NodeDesc* __restrict__ nodes,
                                                         const int* __restrict__ rlist,
                                                         const double* __restrict__ xsol,
                                                         int ldsol, double* __restrict__ xloc,
                                                         int nrhs) {
   const int node = node_list[blockIdx.x];
   const NodeDesc nd = nodes[node];
   double* xl = xloc + nd.solve_off * nrhs;
   const int m = nd.m;
   for (int e = threadIdx.x; e < m * nrhs; e += kThreads) {
      int r = e % m, c = e / m;
      xl[r + (int64_t)c * m] = xsol[rlist[nd.rlist_off + r] - 1 + (int64_t)c * ldsol];
   }
}

// Wave-wide sum.
__device__ inline double wave_sum(double v) {
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
   return v;
}

// Backward CB step: x[j] -= sum_{r >= n} L[r, j] x[r] for the columns
// [r0, r1) of the task (one wave per column).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_bwd_cb(const SolveTask* __restrict__ tasks,
                                                     const NodeDesc* __restrict__ nodes,
                                                     const T* __restrict__ arena,
                                                     double* __restrict__ xloc, int nrhs) {
   const SolveTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const int m = nd.m;
   double* xl = xloc + nd.solve_off * nrhs;
   const T* F = arena + nd.off;
   for (int j = t.r0 + wave; j < t.r1; j += 4) {
      for (int c0 = 0; c0 < nrhs; c0 += kMaxRhsChunk) {
         const int nc = min(kMaxRhsChunk, nrhs - c0);
         double acc[kMaxRhsChunk];
#pragma unroll
         for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] = 0.0;
         for (int r = nd.n + lane; r < m; r += 64) {
            const double lv = (double)F[r + (int64_t)j * nd.ld];
#pragma unroll
            for (int c = 0; c < kMaxRhsChunk; ++c)
               if (c < nc) acc[c] += lv * xl[r + (int64_t)(c0 + c) * m];
         }
#pragma unroll
         for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] = wave_sum(acc[c]);
         if (lane < nc) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < kMaxRhsChunk; ++c)
               if (c == lane) s = acc[c];
            xl[j + (int64_t)(c0 + lane) * m] -= s;
         }
      }
   }
}

// Backward step on block [k0, k0+kb): x_k = inv(L_kk)^T x_k (written to xsol
// by the task with write=1), then columns [r0, r1) (< k0) of the work vector
// -= L[k0:k0+kb, j]^T x_k (one wave per column).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_bwd_step(const SolveTask* __restrict__ tasks,
                                                       const NodeDesc* __restrict__ nodes,
                                                       const T* __restrict__ arena,
                                                       const T* __restrict__ dinv,
                                                       const int* __restrict__ rlist,
                                                       double* __restrict__ xloc,
                                                       double* __restrict__ xsol, int ldsol,
                                                       int nrhs) {
   __shared__ double xs[kNB * kMaxRhsChunk];
   __shared__ double ys[kNB * kMaxRhsChunk];
   const SolveTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   const int m = nd.m, kb = t.kb;
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   double* xl = xloc + nd.solve_off * nrhs;
   const T* Li = dinv + nd.dinv_off + (int64_t)(t.k0 / kNB) * kNB * kNB;
   const T* F = arena + nd.off;
   for (int c0 = 0; c0 < nrhs; c0 += kMaxRhsChunk) {
      const int nc = min(kMaxRhsChunk, nrhs - c0);
      __syncthreads();
      for (int e = threadIdx.x; e < kb * nc; e += kThreads) {
         int i = e % kb, c = e / kb;
         xs[i + c * kNB] = xl[t.k0 + i + (int64_t)(c0 + c) * m];
      }
      __syncthreads();
      for (int e = threadIdx.x; e < kb * nc; e += kThreads) {
         int i = e % kb, c = e / kb;
         double s = 0.0;
         for (int l = i; l < kb; ++l) s += (double)Li[l + i * kb] * xs[l + c * kNB];
         ys[i + c * kNB] = s;
      }
      __syncthreads();
      if (t.write)
         for (int e = threadIdx.x; e < kb * nc; e += kThreads) {
            int i = e % kb, c = e / kb;
            xsol[rlist[nd.rlist_off + t.k0 + i] - 1 + (int64_t)(c0 + c) * ldsol] = ys[i + c * kNB];
         }
      for (int j = t.r0 + wave; j < t.r1; j += 4) {
         double acc[kMaxRhsChunk];
#pragma unroll
         for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] = 0.0;
         for (int l = lane; l < kb; l += 64) {
            const double lv = (double)F[t.k0 + l + (int64_t)j * nd.ld];
#pragma unroll
            for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] += lv * ys[l + c * kNB];
         }
#pragma unroll
         for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] = wave_sum(acc[c]);
         if (lane < nc) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < kMaxRhsChunk; ++c)
               if (c == lane) s = acc[c];
            xl[j + (int64_t)(c0 + lane) * m] -= s;
         }
      }
   }
}

// ---------------------------------------------------------------------------
// FP64 refinement helpers.
// ---------------------------------------------------------------------------
// r = b - A x with A in CSR (both triangles, pivot order); one row per thread.
__global__ __launch_bounds__(kThreads) void k_residual(int n, const int64_t* __restrict__ rowptr,
                                                       const int* __restrict__ col,
                                                       const double* __restrict__ val,
                                                       const double* __restrict__ x,
                                                       const double* __restrict__ b,
                                                       double* __restrict__ r, int nrhs) {
   const int i = blockIdx.x * kThreads + threadIdx.x;
   if (i >= n) return;
   for (int c = 0; c < nrhs; ++c) {
      double s = b[i + (int64_t)c * n];
      for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p)
         s -= val[p] * x[col[p] + (int64_t)c * n];
      r[i + (int64_t)c * n] = s;
   }
}

// Per-column max |v| into out[c] (as ordered bits of a non-negative double).
__global__ __launch_bounds__(kThreads) void k_colmax(int n, const double* __restrict__ v, int nrhs,
                                                     unsigned long long* __restrict__ out) {
   const int c = blockIdx.y;
   double m = 0.0;
   for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
      m = fmax(m, fabs(v[i + (int64_t)c * n]));
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
   __shared__ double wm[4];
   if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
   __syncthreads();
   if (threadIdx.x == 0) {
      m = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
      atomicMax(out + c, (unsigned long long)__double_as_longlong(m));
   }
}

__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
   int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kThreads) void k_axpy(int64_t n, double a, const double* __restrict__ x,
                                                   double* __restrict__ y) {
   int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (i < n) y[i] += a * x[i];
}

__global__ __launch_bounds__(kThreads) void k_copy(int64_t n, const double* __restrict__ x,
                                                   double* __restrict__ y) {
   int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (i < n) y[i] = x[i];
}

__global__ __launch_bounds__(kThreads) void k_fill(int64_t n, double v, double* __restrict__ y) {
   int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (i < n) y[i] = v;
}

// CSR values gathered from the cleaned input values.
__global__ __launch_bounds__(kThreads) void k_gather_vals(int64_t nnz, const int64_t* __restrict__ src,
                                                          const double* __restrict__ aval,
                                                          const double* __restrict__ rowscale,
                                                          const double* __restrict__ colscale_dummy,
                                                          double* __restrict__ out) {
   (void)colscale_dummy;
   (void)rowscale;
   int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (i < nnz) out[i] = aval[src[i]];
}

}  // namespace sylver_amd
