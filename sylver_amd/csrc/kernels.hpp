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
#include <type_traits>

namespace sylver_amd {

constexpr int kThreads = 256;
constexpr int kNB = 128;      // block size of the large-front path and of the solve
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
// A run of one front's columns [c0, c1) for the level assembly (k_assemble):
// staged != 0 when it is assembled in LDS; its original entries are the
// front's sorted A-map pairs [p0, p1); the children's contribution columns
// that land in it are the ranges asmch[g0 .. g0 + gn) = {child, j0, j1, 0}
// (columns j0 .. j1 - 1 of the child's contribution block), in child order.
struct AsmTask {
   int node, c0, c1, staged;
   int64_t p0, p1;
   int g0, gn, pad0, pad1;
};
struct DiagTask { int node, k0, kb, pad; };
struct GemmTask {
   int64_t a_off, b_off, c_off;
   int lda, ldb, ldc;
   int M, N, K;
   int pad0, pad1;
};
struct SolveTask { int node, k0, kb, r0, r1, write; };

// Zero pad for masked loads. A masked-off element is read from here instead
// of being skipped: `ok ? *p : 0` compiles to a branch around the load with
// a wait for it at the join, which serializes a batch of loads; selecting
// the address keeps every load unconditional and in flight together. The
// pad is never written (device globals start zeroed) and is not const, so
// the compiler cannot fold the load back into a select.
__device__ __align__(16) double g_zero_pad[4];

template <typename P>
__device__ __forceinline__ const P* live_or_zero(bool ok, const P* p) {
   return ok ? p : reinterpret_cast<const P*>(g_zero_pad);
}

// Raw buffer loads and stores: byte offset from `base`; a load at or beyond
// `bytes` reads zero and a store there is dropped, which serves as the mask
// (one 32-bit offset per lane instead of a 64-bit address per element).
template <typename T>
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const T* base, uint32_t bytes) {
   return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ inline float buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, float*) {
   return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ inline double buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, double*) {
   return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
__device__ inline void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, float v) {
   __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)off, 0, 0);
}
__device__ inline void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, double v) {
   typedef unsigned int u2 __attribute__((ext_vector_type(2)));
   __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)off, 0, 0);
}

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

// 1/sqrt(x) for x > 0 from the hardware estimate (v_rsq) refined by Newton
// steps y <- y + y (1 - x y^2) / 2 (two in FP64, one in FP32): a few
// dependent operations instead of a correctly rounded square root followed
// by a division, on the critical path of the diagonal-block Cholesky;
// within a few ulps. Values in the normal range only (pivots of an SPD front).
__device__ inline double rsqrt_nr(double x) {
   double y = __builtin_amdgcn_rsq(x);
#pragma unroll
   for (int k = 0; k < 2; ++k) {
      const double r = __builtin_fma(-x * y, y, 1.0);
      y = __builtin_fma(0.5 * y, r, y);
   }
   return y;
}
__device__ inline float rsqrt_nr(float x) {
   const float y = __builtin_amdgcn_rsqf(x);
   const float r = __builtin_fmaf(-x * y, y, 1.0f);
   return __builtin_fmaf(0.5f * y, r, y);
}

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
      // 8 rows per lane per batch: all loads issued before the stores
      for (int i0 = j; i0 < q; i0 += 8 * 64) {
         int ci[8];
         T sv[8], dv[8];
#pragma unroll
         for (int u = 0; u < 8; ++u) {
            const int i = min(i0 + u * 64 + lane, q - 1);
            ci[u] = cm[i];
            sv[u] = src[i];
         }
#pragma unroll
         for (int u = 0; u < 8; ++u) dv[u] = dst[ci[u]];
#pragma unroll
         for (int u = 0; u < 8; ++u)
            if (i0 + u * 64 + lane < q) dst[ci[u]] = dv[u] + sv[u];
      }
   }
}

// ---------------------------------------------------------------------------
// Assembly of a level in one launch, by the receiving columns: each task
// owns columns [c0, c1) of one front and writes nothing else. It zeroes
// their lower triangle, writes the original entries that fall in them
// (asort: the front's A-map sorted by position, (col)*nrow + row), then adds
// the contribution-block columns of every child that land in them, children
// in list order. Per entry these are the operations, in the order, of
// k_zero, k_init_a and one k_extend_add launch per child rank, so the front
// is bitwise the same; the barriers between children order their sums
// (a child's map is injective, two children may meet on an entry). Child
// maps are increasing (posdef: no delayed rows), so the child columns of a
// task are a range; the plan holds every task's ranges (and its range of
// original entries), so the kernel starts on its loads at once.
// ---------------------------------------------------------------------------

// Runs of at most kAsmLdsBytes of entries (AsmTask::staged = 1) are
// assembled in LDS and written to the front once: the zero fill and the
// parent's read-modify-write per child entry stay on chip, the front is
// written once. The run is its columns' lower trapezoid, column c (rows c ..
// m-1) at offset (c - c0) m - (c - c0)(c + c0 - 1) / 2. Longer runs (a
// single column of more rows) work on the front directly. 16 KiB per
// workgroup keeps ten of them resident per CU (measured at 80^3 against 32
// and 64 KiB runs: FP32 0.4 ms, FP64 1.0 ms faster per factorization than
// 8192-entry runs).
constexpr int kAsmLdsBytes = 16384;

template <typename T>
__global__ __launch_bounds__(kThreads) void k_assemble(const AsmTask* __restrict__ tasks,
                                                       const NodeDesc* __restrict__ nodes,
                                                       const int64_t* __restrict__ asort,
                                                       const int* __restrict__ rlist,
                                                       const double* __restrict__ aval,
                                                       const double* __restrict__ scaling,
                                                       const int* __restrict__ cmap, const int4* __restrict__ asmch,
                                                       T* __restrict__ arena) {
   extern __shared__ __align__(16) unsigned char asm_lds[];
   T* run = reinterpret_cast<T*>(asm_lds);
   const AsmTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const bool staged = t.staged != 0;
   T* F = arena + nd.off;
   const int m = nd.m, c0 = t.c0;
   // run offset of column c (staged)
   auto col_off = [&](int c) { return (c - c0) * m - (int)(((int64_t)(c - c0) * (c + c0 - 1)) / 2); };
   if (staged) {
      const int area = col_off(t.c1);
      for (int e = threadIdx.x; e < area; e += kThreads) run[e] = T(0);
   } else {
      for (int c = c0 + wave; c < t.c1; c += 4) {
         T* col = F + (int64_t)c * nd.ld;
         for (int r = c + lane; r < m; r += 64) col[r] = T(0);
      }
   }
   __syncthreads();
   {
      // original entries: four per thread in flight (pairs, then values)
      constexpr int AU = 4;
      const int64_t* am = asort + 2 * nd.amap_off;
      for (int64_t pb = t.p0; pb < t.p1; pb += AU * kThreads) {
         int64_t src[AU], dest[AU];
#pragma unroll
         for (int u = 0; u < AU; ++u) {
            const int64_t p = min(pb + u * kThreads + (int64_t)threadIdx.x, t.p1 - 1);
            src[u] = am[2 * p] - 1;
            dest[u] = am[2 * p + 1] - 1;
         }
         double v[AU];
#pragma unroll
         for (int u = 0; u < AU; ++u) v[u] = aval[src[u]];
#pragma unroll
         for (int u = 0; u < AU; ++u) {
            if (pb + u * kThreads + (int64_t)threadIdx.x >= t.p1) continue;
            const int c = (int)(dest[u] / m);
            const int r = (int)(dest[u] - (int64_t)c * m);
            double w = v[u];
            if (scaling) w = scaling[rlist[nd.rlist_off + r] - 1] * w * scaling[rlist[nd.rlist_off + c] - 1];
            if (staged)
               run[col_off(c) + (r - c)] = (T)w;
            else
               F[(int64_t)c * nd.ld + r] = (T)w;
         }
      }
   }
   for (int g = t.g0; g < t.g0 + t.gn; ++g) {
      const int4 cr = asmch[g];
      const NodeDesc cn = nodes[cr.x];
      __syncthreads();
      const int q = cn.m - cn.n;
      const int* __restrict__ cm = cmap + cn.cmap_off;
      for (int j = cr.y + wave; j < cr.z; j += 4) {
         const int pj = cm[j];
         const T* src = arena + cn.off + (int64_t)(cn.n + j) * cn.ld + cn.n;
         if (staged) {
            T* dst = run + col_off(pj) - pj;  // row r of column pj at dst[r]
            for (int i0 = j; i0 < q; i0 += 8 * 64) {
               int ci[8];
               T sv[8];
#pragma unroll
               for (int u = 0; u < 8; ++u) {
                  const int i = min(i0 + u * 64 + lane, q - 1);
                  ci[u] = cm[i];
                  sv[u] = src[i];
               }
#pragma unroll
               for (int u = 0; u < 8; ++u)
                  if (i0 + u * 64 + lane < q) dst[ci[u]] = dst[ci[u]] + sv[u];
            }
         } else {
            T* dst = F + (int64_t)pj * nd.ld;
            // 8 rows per lane per batch: all loads issued before the stores
            for (int i0 = j; i0 < q; i0 += 8 * 64) {
               int ci[8];
               T sv[8], dv[8];
#pragma unroll
               for (int u = 0; u < 8; ++u) {
                  const int i = min(i0 + u * 64 + lane, q - 1);
                  ci[u] = cm[i];
                  sv[u] = src[i];
               }
#pragma unroll
               for (int u = 0; u < 8; ++u) dv[u] = dst[ci[u]];
#pragma unroll
               for (int u = 0; u < 8; ++u)
                  if (i0 + u * 64 + lane < q) dst[ci[u]] = dv[u] + sv[u];
            }
         }
      }
   }
   if (!staged) return;
   __syncthreads();
   // the run to the front, column by column (rows c .. m-1)
   for (int c = c0 + wave; c < t.c1; c += 4) {
      const T* s = run + col_off(c) - c;
      T* col = F + (int64_t)c * nd.ld;
      for (int r = c + lane; r < m; r += 64) col[r] = s[r];
   }
}

// ---------------------------------------------------------------------------
// In-LDS dense kernels on the diagonal block of the large-front path. The
// block (kb <= kNB) is held padded to a multiple of 16 with identity, so
// every loop below works on whole 16x16 sub-blocks: the Cholesky factor and
// the inverse of [A 0; 0 I] are [L 0; 0 I] and [inv(L) 0; 0 I].
// ---------------------------------------------------------------------------
// Value of v in lane `src` (a compile-time constant in the callers) for the
// whole wave, through v_readlane (a scalar operand afterwards).
__device__ __forceinline__ float lane_bcast(float v, int src) {
   return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double lane_bcast(double v, int src) {
   const long long u = __double_as_longlong(v);
   const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffff), src);
   const int hi = __builtin_amdgcn_readlane((int)(u >> 32), src);
   return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// In-place inverse of a 16x16 lower triangular block by one wave: lane j < 16
// computes column j of the inverse by forward substitution.
template <typename T, int LD>
__device__ void wave_trtri16(T* D) {
   const int lane = threadIdx.x & 63;
   T x[16], rd[16];
#pragma unroll
   for (int i = 0; i < 16; ++i) rd[i] = T(1) / D[i + i * LD];
#pragma unroll
   for (int i = 0; i < 16; ++i) {
      T s = (i == lane) ? T(1) : T(0);
#pragma unroll
      for (int k = 0; k < i; ++k) s -= D[i + k * LD] * x[k];
      x[i] = (i >= lane) ? s * rd[i] : T(0);
   }
   __builtin_amdgcn_wave_barrier();
   if (lane < 16)
#pragma unroll
      for (int i = 0; i < 16; ++i)
         if (i >= lane) D[i + lane * LD] = x[i];
}

// C(16x16) += A(16x16) * B(16x16)^T or A*B, operands in LDS.
// A element (r, k) at A[r + k*lda]; B element (k, c) at B[c + k*ldb] when
// transB (B given as rows c), else at B[k + c*ldb].
template <typename T, bool TRANSB>
__device__ inline typename Mfma<T>::acc_t lds_mma16(const T* A, int lda, const T* B, int ldb,
                                                    typename Mfma<T>::acc_t acc) {
   const int lane = threadIdx.x & 63;
#pragma unroll
   for (int kc = 0; kc < 16; kc += 4) {
      const int k = kc + (lane >> 4), rc = lane & 15;
      const T a = A[rc + k * lda];
      const T b = TRANSB ? B[rc + k * ldb] : B[k + rc * ldb];
      acc = Mfma<T>::op(a, b, acc);
   }
   return acc;
}

// Cholesky of a 16-column panel in LDS by one wave: lanes 0..15 hold the
// rows of the diagonal block, lanes 16..63 48 rows below it (rows o + 16 +
// 48 w + ...; w = the wave's index among the panel's waves), each row in
// registers. Column step j: l_ij = a_ij / sqrt(d), a_ic -= (l_ij / sqrt(d))
// a_cj for c > j -- one FMA per element with a_cj broadcast from lane c by
// v_readlane, so a step waits on no memory round trip. The rows below the
// diagonal block are simply more rows of the same elimination, so a panel
// of up to 64 rows takes the time of the 16x16 block alone. Every wave
// of a panel factors the diagonal block (the same values, the same
// operations, so the same results); wave 0 stores it.
template <typename T, int LD>
__device__ __forceinline__ int wave_panel16(T* L, int o, int kb16, int w) {
   const int lane = threadIdx.x & 63;
   const int row = lane < 16 ? o + lane : o + 16 + w * 48 + (lane - 16);
   const bool live = row < kb16;
   T r[16];
#pragma unroll
   for (int c = 0; c < 16; ++c) r[c] = (live && (lane >= 16 || c <= lane)) ? L[row + (o + c) * LD] : T(0);
   // Branch-free, so the whole panel is one block the compiler can schedule:
   // the next pivot (row j + 1's update, then its broadcast) is formed first
   // and its reciprocal square root overlaps the rest of step j. After a
   // non-positive pivot the steps go on with meaningless values; the first
   // such column is what is returned.
   int bad = 0;
   T d = lane_bcast(r[0], 0);
#pragma unroll
   for (int j = 0; j < 16; ++j) {
      T col[16];
#pragma unroll
      for (int c = j + 1; c < 16; ++c) col[c] = lane_bcast(r[j], c);
      bad = (bad == 0 && !(d > T(0))) ? j + 1 : bad;
      const T inv = rsqrt_nr(d);
      const T lj = r[j] * inv;
      const T t = lj * inv;
      r[j] = (lane == j) ? d * inv : lj;
      if (j + 1 < 16) {
         r[j + 1] = fma(-t, col[j + 1], r[j + 1]);
         d = lane_bcast(r[j + 1], j + 1);
      }
#pragma unroll
      for (int c = j + 2; c < 16; ++c) r[c] = fma(-t, col[c], r[c]);
   }
   if (live) {
      if (lane < 16) {
         if (w == 0)
#pragma unroll
            for (int c = 0; c < 16; ++c)
               if (c <= lane) L[row + (o + c) * LD] = r[c];
      } else {
#pragma unroll
         for (int c = 0; c < 16; ++c) L[row + (o + c) * LD] = r[c];
      }
   }
   return bad;
}

// Waves that factor the panel at column o of a kb16-row block (48 rows below
// the diagonal block each).
__device__ inline int panel_waves(int kb16, int o) {
   const int below = kb16 - o - 16;
   return below > 0 ? (below + 47) / 48 : 1;
}

// Row and column of lower 16x16 tile `tile` in packed row order.
__device__ inline void packed_tile(int tile, int& ti, int& tj) {
   ti = 0;
   while ((ti + 1) * (ti + 2) / 2 <= tile) ++ti;
   tj = tile - ti * (ti + 1) / 2;
}

// Lower 16x16 tiles (ti, tj) and (ti2, tj2) of the trailing block at t0 minus
// the products of panel o: both products are formed before either is
// applied, so their LDS and MFMA latencies overlap (the tiles are disjoint).
template <typename T, int LD>
__device__ void panel_update2(T* L, int o, int t0, int ti, int tj, int ti2, int tj2, bool two) {
   typedef typename Mfma<T>::acc_t acc_t;
   const int lane = threadIdx.x & 63;
   acc_t acc = acc_t{0, 0, 0, 0}, acc2 = acc_t{0, 0, 0, 0};
   acc = lds_mma16<T, true>(L + t0 + ti * 16 + o * LD, LD, L + t0 + tj * 16 + o * LD, LD, acc);
   if (two) acc2 = lds_mma16<T, true>(L + t0 + ti2 * 16 + o * LD, LD, L + t0 + tj2 * 16 + o * LD, LD, acc2);
#pragma unroll
   for (int v = 0; v < 4; ++v) {
      const int i = t0 + ti * 16 + Mfma<T>::row(lane, v), j = t0 + tj * 16 + (lane & 15);
      const int i2 = t0 + ti2 * 16 + Mfma<T>::row(lane, v), j2 = t0 + tj2 * 16 + (lane & 15);
      if (i >= j) L[i + j * LD] -= acc[v];
      if (two && i2 >= j2) L[i2 + j2 * LD] -= acc2[v];
   }
}

// Blocked Cholesky of the padded kb16 x kb16 block in LDS in 16-column
// panels, each factored whole by wave_panel16 (up to 3 waves). Look-ahead:
// after panel s the next panel's column of tiles is updated first; the waves
// then factor panel s + 1 while the others (and they, once done) apply the
// rest of panel s's update, handed out through an LDS counter. Two barriers
// per panel. All threads call; returns 0 or the 1-based column of the first
// non-positive pivot.
template <typename T, int LD>
__device__ __forceinline__ int lds_potrf_panels(T* L, int kb16, int* sh_bad) {
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int nbk = kb16 / 16;
   __shared__ int next_tile;
   // round s + 1 factors panel s + 1 (at t0); from the second round on it is
   // preceded by panel s's update of that panel's column and overlapped by
   // the rest of panel s's update (one copy of the panel code: it runs once
   // per diagonal block, fetched cold)
   for (int s = -1; s + 1 < nbk; ++s) {
      const int o = s * 16, t0 = o + 16, nt = (kb16 - t0) / 16;
      if (s >= 0) {
         if (*sh_bad) return *sh_bad;
         // the next panel's column of tiles
         for (int ti = wave; ti < nt; ti += 4) panel_update2<T, LD>(L, o, t0, ti, 0, ti, 0, false);
         if (threadIdx.x == 0) next_tile = 0;
         __syncthreads();
      }
      if (wave < panel_waves(kb16, t0)) {
         const int bad = wave_panel16<T, LD>(L, t0, kb16, wave);
         if (wave == 0 && lane == 0) *sh_bad = bad ? t0 + bad : 0;
      }
      if (s >= 0) {
         // the rest of panel s's update: the lower triangle at t0 + 16, two tiles at a time
         const int nr = nt - 1, nrt = nr * (nr + 1) / 2, t1 = t0 + 16;
         for (;;) {
            int tile = 0;
            if (lane == 0) tile = atomicAdd(&next_tile, 2);
            tile = __shfl(tile, 0);
            if (tile >= nrt) break;
            int ti, tj;
            packed_tile(tile, ti, tj);
            // the next tile in the same packed order: next column, or the next row's first
            const int ti2 = tj + 1 <= ti ? ti : ti + 1, tj2 = tj + 1 <= ti ? tj + 1 : 0;
            panel_update2<T, LD>(L, o, t1, ti, tj, ti2, tj2, tile + 1 < nrt);
         }
      }
      __syncthreads();
   }
   return *sh_bad;
}

// Blocked in-place inverse of the padded lower triangular kb16 x kb16 block
// (block columns right to left; Tmp holds kNB x 16 elements, ld kNB).
template <typename T, int LD>
__device__ void lds_trtri_blocked(T* L, T* Tmp, int kb16) {
   typedef typename Mfma<T>::acc_t acc_t;
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const int nbk = kb16 / 16;
   for (int b = wave; b < nbk; b += 4) wave_trtri16<T, LD>(L + b * 16 + b * 16 * LD);
   __syncthreads();
   for (int j = nbk - 2; j >= 0; --j) {
      // T_i = sum_{k=j+1..i} X_ik L_kj for the block rows below j
      for (int i = j + 1 + wave; i < nbk; i += 4) {
         acc_t acc = acc_t{0, 0, 0, 0};
         for (int k = j + 1; k <= i; ++k)
            acc = lds_mma16<T, false>(L + i * 16 + k * 16 * LD, LD, L + k * 16 + j * 16 * LD, LD, acc);
#pragma unroll
         for (int v = 0; v < 4; ++v)
            Tmp[(i * 16 + Mfma<T>::row(lane, v)) + (lane & 15) * kNB] = acc[v];
      }
      __syncthreads();
      // X_ij = -T_i X_jj
      for (int i = j + 1 + wave; i < nbk; i += 4) {
         acc_t acc = acc_t{0, 0, 0, 0};
         acc = lds_mma16<T, false>(Tmp + i * 16, kNB, L + j * 16 + j * 16 * LD, LD, acc);
#pragma unroll
         for (int v = 0; v < 4; ++v)
            L[(i * 16 + Mfma<T>::row(lane, v)) + (j * 16 + (lane & 15)) * LD] = -acc[v];
      }
      __syncthreads();
   }
}

// ---------------------------------------------------------------------------
// Diagonal block of the large-front path: POTRF of the kb x kb block at
// (k0, k0), L written back, its inverse written to dinv (ld = kb) for the
// TRSM-by-GEMM and the solves. With do_potrf == 0 only the inverse of the
// already factored block is formed (small fronts and the strip-TRSM blocks,
// for the solves); with NOINV only the factor (the strip TRSM below does not
// use the inverse, which a later launch forms for the solves).
// ---------------------------------------------------------------------------
// The kb16 x kb16 diagonal block into LDS (ld kNB + 1): the lower triangle,
// identity beyond kb, zero above. The block is rows row0 .. row0 + kb of the
// panel buffer `rs` (ld elements per column, `bytes` long). Thread t reads
// row t % 128 of every other column (t / 128, t / 128 + 2, ...): one load
// per element, a wave's load covers 64 consecutive rows of a column, 128
// bytes of loads in flight per thread. The code is short on purpose: the
// kernels that call this run once per diagonal block, right after other
// kernels, so every instruction they execute is fetched cold.
template <typename T>
__device__ __forceinline__ void load_diag_lds(__amdgpu_buffer_rsrc_t rs, uint32_t bytes, int row0, int ld, int kb, int kb16, T* L) {
   constexpr int LD = kNB + 1;
   constexpr int NU = 128 / sizeof(T);
   constexpr uint32_t ES = sizeof(T);
   const int i = threadIdx.x & (kNB - 1), jt = threadIdx.x >> 7;
   if (i >= kb16) return;
   // rows at or beyond kb read from the end of the buffer (zero)
   const uint32_t vrow = i < kb ? (uint32_t)(row0 + i) * ES : bytes;
   for (int u0 = 0; 2 * u0 < kb16; u0 += NU) {
      T v[NU];
#pragma unroll
      for (int u = 0; u < NU; ++u) {
         const int j = jt + 2 * (u0 + u);
         v[u] = buf_load(rs, j < kb ? vrow + (uint32_t)j * (uint32_t)ld * ES : bytes, (T*)nullptr);
      }
#pragma unroll
      for (int u = 0; u < NU; ++u) {
         const int j = jt + 2 * (u0 + u);
         if (j < kb16) L[i + j * LD] = (i < kb && j < kb) ? (i >= j ? v[u] : T(0)) : T(i == j ? 1 : 0);
      }
   }
}

// Buffer over columns k0 .. k0 + kb of a front (the block's panel).
template <typename T>
__device__ inline __amdgpu_buffer_rsrc_t panel_rsrc(T* arena, const NodeDesc& nd, int k0, int kb, uint32_t& bytes) {
   bytes = (uint32_t)((int64_t)kb * nd.ld * (int64_t)sizeof(T));
   return buf_rsrc(arena + nd.off + (int64_t)k0 * nd.ld, bytes);
}

template <typename T, bool NOINV>
__device__ void diag_block(const DiagTask& t, const NodeDesc& nd, T* __restrict__ arena, T* __restrict__ dinv,
                           T* __restrict__ dinvT, int* __restrict__ info, int do_potrf, T* L, T* Tmp,
                           int* sh_bad) {
   constexpr int LD = kNB + 1;
   const int kb = t.kb;
   const int kb16 = (kb + 15) & ~15;
   uint32_t pbytes;
   const auto rs = panel_rsrc(arena, nd, t.k0, kb, pbytes);
   load_diag_lds<T>(rs, pbytes, t.k0, nd.ld, kb, kb16, L);
   __syncthreads();
   if (do_potrf) {
      const int bad = lds_potrf_panels<T, LD>(L, kb16, sh_bad);
      if (bad) {
         if (threadIdx.x == 0) info[t.node] = t.k0 + bad;
         return;
      }
      // the factor back to the front, in the layout of the load: 16 LDS
      // reads, then 16 stores, per batch (a loop of single iterations waits
      // on every LDS read before its store); masked stores go past the end
      {
         constexpr uint32_t ES = sizeof(T);
         const int i = threadIdx.x & (kNB - 1), jt = threadIdx.x >> 7;
         const uint32_t vrow = (uint32_t)(t.k0 + i) * ES;
         for (int u0 = 0; 2 * u0 < kb16; u0 += 16) {
            T v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) v[u] = L[i + min(jt + 2 * (u0 + u), kNB - 1) * LD];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
               const int j = jt + 2 * (u0 + u);
               const bool ok = i < kb && j < kb && i >= j;
               buf_store(rs, ok ? vrow + (uint32_t)j * (uint32_t)nd.ld * ES : pbytes, v[u]);
            }
         }
      }
      if (NOINV) return;
   }
   lds_trtri_blocked<T, LD>(L, Tmp, kb16);
   T* out = dinv + nd.dinv_off + (int64_t)(t.k0 / kNB) * kNB * kNB;
   T* outT = dinvT + nd.dinv_off + (int64_t)(t.k0 / kNB) * kNB * kNB;
   if (kb == kNB) {
      // row index from the low bits, stores coalesced per column (ld kb =
      // kNB); LDS reads batched ahead of the stores as above
      for (int e0 = 0; e0 < kNB * kNB; e0 += 8 * kThreads) {
         T v[8], w[8];
#pragma unroll
         for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * kThreads + threadIdx.x, i = e & (kNB - 1), j = e >> 7;
            v[u] = L[i + j * LD];
            w[u] = L[j + i * LD];
         }
#pragma unroll
         for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * kThreads + threadIdx.x, i = e & (kNB - 1), j = e >> 7;
            out[e] = (i >= j) ? v[u] : T(0);
            outT[e] = (j >= i) ? w[u] : T(0);
         }
      }
   } else {
      for (int e = threadIdx.x; e < kb * kb; e += kThreads) {
         const int i = e % kb, j = e / kb;
         out[i + j * kb] = (i >= j) ? L[i + j * LD] : T(0);
         outT[i + j * kb] = (j >= i) ? L[j + i * LD] : T(0);
      }
   }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_diag(const DiagTask* __restrict__ tasks,
                                                   const NodeDesc* __restrict__ nodes,
                                                   T* __restrict__ arena, T* __restrict__ dinv,
                                                   T* __restrict__ dinvT, int* __restrict__ info,
                                                   int do_potrf) {
   constexpr int LD = kNB + 1;
   __shared__ T L[kNB * LD];
   __shared__ T Tmp[kNB * 16];
   __shared__ int sh_bad;
   // the diagonal blocks are the factorization's critical path: their waves
   // win instruction issue over the look-ahead updates sharing the CU
   __builtin_amdgcn_s_setprio(3);
   const DiagTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   diag_block<T, false>(t, nd, arena, dinv, dinvT, info, do_potrf, L, Tmp, &sh_bad);
}

// Factor only (the blocks whose panel the strip TRSM below solves).
template <typename T>
__device__ __forceinline__ void diag_factor_task(const DiagTask* __restrict__ tasks, const NodeDesc* __restrict__ nodes,
                                                 T* __restrict__ arena, int* __restrict__ info) {
   constexpr int LD = kNB + 1;
   __shared__ T L[kNB * LD];
   __shared__ int sh_bad;
   __builtin_amdgcn_s_setprio(3);  // the factorization's critical path
   const DiagTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   diag_block<T, true>(t, nd, arena, nullptr, nullptr, info, 1, L, nullptr, &sh_bad);
}
template <typename T>
__global__ __launch_bounds__(kThreads) void k_diag_factor(const DiagTask* __restrict__ tasks,
                                                          const NodeDesc* __restrict__ nodes,
                                                          T* __restrict__ arena, int* __restrict__ info) {
   diag_factor_task<T>(tasks, nodes, arena, info);
}
// FP32, within 104 VGPRs: with its 66 KB of LDS it then fits on a CU beside
// two workgroups of the FP32 Schur update (2 x 200 registers per SIMD lane,
// 2 x 33 KB), so the critical path's diagonal block starts at once instead of
// waiting for a CU to drain
static __global__ __launch_bounds__(kThreads) __attribute__((amdgpu_num_vgpr(104))) void k_diag_factor_f32(
    const DiagTask* __restrict__ tasks, const NodeDesc* __restrict__ nodes, float* __restrict__ arena,
    int* __restrict__ info) {
   diag_factor_task<float>(tasks, nodes, arena, info);
}

// ---------------------------------------------------------------------------
// Panel TRSM of the large-front path by strips, after k_diag_factor: one
// workgroup per 64 rows [r0, r0 + 64) below the diagonal block computes
// X = A21 L11^{-T} by blocked substitution over 16-column blocks b,
// X_b = (A_b - sum_{k<b} X_k L_bk^T) inv(L_bb)^T, with the 16x16 inverses
// formed here from L11 (by wave_trtri16, as k_diag forms them). The chain
// DIAG -> TRSM thus does not wait for the inverse of the whole block (formed
// for the solves by the factorization's last launch).
// Each wave owns 16 rows and keeps its 16x16 blocks in registers in the
// layout of the MFMA output taken transposed: item v of lane l holds
// M[l % 16][row(l, v)]. A product R = P Q (P in that layout, Q from LDS) is
// the sum over s of mfma(Q[row(l, s)][l % 16], item s of P) -- the MFMA's
// k index (s, l / 16) stands for row(l, s) -- and comes out in the same
// layout, so the chain of products needs no shuffles. Loads and stores go
// through a buffer over the block's kb panel columns (32-bit offsets).
// ---------------------------------------------------------------------------
struct StripTask { int node, k0, kb, r0; };

template <typename T>
__global__ __launch_bounds__(kThreads) void k_strip(const StripTask* __restrict__ tasks,
                                                    const NodeDesc* __restrict__ nodes, T* __restrict__ arena) {
   constexpr int LD = kNB + 1;
   typedef typename Mfma<T>::acc_t acc_t;
   __shared__ T L[kNB * LD];
   __builtin_amdgcn_s_setprio(3);  // the factorization's critical path
   const StripTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const int kb = t.kb, kb16 = (kb + 15) & ~15, nbk = kb16 / 16;
   const int ld = nd.ld;
   uint32_t pbytes;
   const auto rs = panel_rsrc(arena, nd, t.k0, kb, pbytes);
   const int i = t.r0 + wave * 16 + (lane & 15);  // this lane's row
   const bool rowok = i < nd.m;
   auto off = [&](int c) {  // byte offset of (row i, panel column c), or past the panel
      return (rowok && c < kb) ? (uint32_t)(i + c * ld) * (uint32_t)sizeof(T) : pbytes;
   };
   // A21 rows of the wave, 16-column blocks in the register layout
   T x[8][4];
#pragma unroll
   for (int b = 0; b < 8; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) x[b][v] = buf_load(rs, off(b * 16 + Mfma<T>::row(lane, v)), (T*)nullptr);
   // L11 (factored by the launch before) into LDS, the 16x16 diagonal inverses in place
   load_diag_lds<T>(rs, pbytes, t.k0, ld, kb, kb16, L);
   __syncthreads();
   for (int b = wave; b < nbk; b += 4) wave_trtri16<T, LD>(L + b * 16 + b * 16 * LD);
   __syncthreads();
   const int r = lane & 15;
#pragma unroll
   for (int b = 0; b < 8; ++b) {
      if (b < nbk) {
         acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
         for (int k = 0; k < b; ++k)
#pragma unroll
            for (int s = 0; s < 4; ++s)
               acc = Mfma<T>::op(L[(b * 16 + r) + (k * 16 + Mfma<T>::row(lane, s)) * LD], x[k][s], acc);
         T tb[4];
#pragma unroll
         for (int v = 0; v < 4; ++v) tb[v] = x[b][v] - acc[v];
         acc = acc_t{0, 0, 0, 0};
#pragma unroll
         for (int s = 0; s < 4; ++s)
            acc = Mfma<T>::op(L[(b * 16 + r) + (b * 16 + Mfma<T>::row(lane, s)) * LD], tb[s], acc);
#pragma unroll
         for (int v = 0; v < 4; ++v) x[b][v] = acc[v];
      }
   }
#pragma unroll
   for (int b = 0; b < 8; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) buf_store(rs, off(b * 16 + Mfma<T>::row(lane, v)), x[b][v]);
}

// ---------------------------------------------------------------------------
// Schur-complement update of the large-front path: C -= A * B^T on one
// 128x128 tile (A, B: K panel columns of the front, rows of the tile's row and
// column blocks). The product is accumulated from zero and subtracted from C
// once at the end (accumulating onto C would round every step at C's
// magnitude). K is streamed in chunks of 16 through a
// double-buffered LDS with the next chunk prefetched into registers (16-byte
// loads: front rows are 8-aligned and leading dimensions a multiple of 8), one
// barrier per chunk. Columns below `pad0` of the tile are left untouched
// (tiles may start a few rows above the trailing block to stay aligned; those
// rows fall in the upper triangle).
// The launch is padded to a multiple of 8 workgroups and workgroup b runs
// task (b % 8) * (G / 8) + b / 8: workgroups are dealt round-robin to the 8
// XCDs, so each XCD gets a contiguous run of tiles sharing A row blocks.
// ---------------------------------------------------------------------------
template <typename T> struct V16;
template <> struct V16<double> {
   typedef double vt __attribute__((ext_vector_type(2)));
   static constexpr int N = 2;
};
template <> struct V16<float> {
   typedef float vt __attribute__((ext_vector_type(4)));
   static constexpr int N = 4;
};

__device__ inline int xcd_task(int b, int G) { return (b & 7) * (G >> 3) + (b >> 3); }

// MODE 0: Schur update C -= A B^T (B in the arena). MODE 2: the same with B
// from `bsrc` (the L D workspace of the indefinite path). MODE 1: panel TRSM
// C = A B^T with B = inv(L_kk) from `bsrc`, C == A in place (the tile spans
// all kb columns, so each workgroup owns its rows and has read all of A
// before it writes). Tasks whose offsets or leading dimensions are not
// multiples of the vector width take scalar loads.
// TB x TBN tiles: 128 x 128; 64 x 64 for the short Schur-update launches on
// the critical path and 64 x 128 for the TRSM (whose tile must span all
// kb columns), so those launches spread over more workgroups.
// One tile of a GEMM task: C (ldc, element type TC) -= A B^T, or C = A B^T
// for MODE 1; A and B of type T (leading dimensions t.lda, t.ldb), products
// on T's MFMA. `vec`: A, B and their leading dimensions allow vector loads.
template <typename T, int MODE, int TB, int TBN, typename TC, int NT = kThreads>
__device__ __forceinline__ void gemm_tile(const GemmTask& t, const T* __restrict__ a, const T* __restrict__ b,
                                          TC* __restrict__ c, int ldc, bool vec) {
   // K staged per barrier: FP64 four-wave tiles 32 columns (longer MFMA chains
   // per stage), eight-wave ones 16 (half the LDS); FP32 16, or 32 for the
   // eight-wave 64-row tiles (so that each thread still moves whole vectors)
   constexpr int BM = TB, BN = TBN, PAD = 4;
   constexpr int BK = sizeof(T) == 8 ? (NT == kThreads ? 32 : 16) : (NT > kThreads && TB == 64 ? 32 : 16);
   constexpr int WGN = NT / 128;  // waves: 2 along the rows, NT / 128 along the columns
   constexpr int WM = BM / 2, WN = BN / WGN, FM = WM / 16, FN = WN / 16;
   typedef typename V16<T>::vt vt;
   constexpr int VN = V16<T>::N;
   constexpr int AV = BM * BK / VN / NT;
   constexpr int BV = BN * BK / VN / NT;
   typedef typename Mfma<T>::acc_t acc_t;
   __shared__ T As[2][BK][BM + PAD];
   __shared__ T Bs[2][BK][BN + PAD];
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int wm = wave / WGN, wn = wave % WGN;
   const int ld = t.lda, ldb = t.ldb;
   acc_t acc[FM][FN];
#pragma unroll
   for (int x = 0; x < FM; ++x)
#pragma unroll
      for (int y = 0; y < FN; ++y) acc[x][y] = acc_t{0, 0, 0, 0};
   vt ra[AV], rb[BV];
   auto gload = [&](int k0) {
      if (vec) {
#pragma unroll
         for (int u = 0; u < AV; ++u) {
            const int v = tid + u * NT;
            const int kk = v / (BM / VN), row = (v % (BM / VN)) * VN;
            ra[u] = *live_or_zero(k0 + kk < t.K && row < t.M,
                                  reinterpret_cast<const vt*>(a + row + (int64_t)(k0 + kk) * ld));
         }
#pragma unroll
         for (int u = 0; u < BV; ++u) {
            const int v = tid + u * NT;
            const int kk = v / (BN / VN), row = (v % (BN / VN)) * VN;
            rb[u] = *live_or_zero(k0 + kk < t.K && row < t.N,
                                  reinterpret_cast<const vt*>(b + row + (int64_t)(k0 + kk) * ldb));
         }
      } else {
#pragma unroll
         for (int u = 0; u < AV; ++u) {
            const int v = tid + u * NT;
            const int kk = v / (BM / VN), row = (v % (BM / VN)) * VN;
#pragma unroll
            for (int e = 0; e < VN; ++e)
               ra[u][e] = *live_or_zero(k0 + kk < t.K && row + e < t.M, a + row + e + (int64_t)(k0 + kk) * ld);
         }
#pragma unroll
         for (int u = 0; u < BV; ++u) {
            const int v = tid + u * NT;
            const int kk = v / (BN / VN), row = (v % (BN / VN)) * VN;
#pragma unroll
            for (int e = 0; e < VN; ++e)
               rb[u][e] = *live_or_zero(k0 + kk < t.K && row + e < t.N, b + row + e + (int64_t)(k0 + kk) * ldb);
         }
      }
   };
   const int nk = (t.K + BK - 1) / BK;
   gload(0);
   for (int kc = 0; kc < nk; ++kc) {
      const int buf = kc & 1;
#pragma unroll
      for (int u = 0; u < AV; ++u) {
         const int v = tid + u * NT;
         const int kk = v / (BM / VN), row = (v % (BM / VN)) * VN;
#pragma unroll
         for (int e = 0; e < VN; ++e) As[buf][kk][row + e] = ra[u][e];
      }
#pragma unroll
      for (int u = 0; u < BV; ++u) {
         const int v = tid + u * NT;
         const int kk = v / (BN / VN), row = (v % (BN / VN)) * VN;
#pragma unroll
         for (int e = 0; e < VN; ++e) Bs[buf][kk][row + e] = rb[u][e];
      }
      __syncthreads();
      if (kc + 1 < nk) gload((kc + 1) * BK);
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
         T af[FM], bf[FN];
#pragma unroll
         for (int x = 0; x < FM; ++x) af[x] = As[buf][kk + (lane >> 4)][wm * WM + x * 16 + (lane & 15)];
#pragma unroll
         for (int y = 0; y < FN; ++y) bf[y] = Bs[buf][kk + (lane >> 4)][wn * WN + y * 16 + (lane & 15)];
#pragma unroll
         for (int x = 0; x < FM; ++x)
#pragma unroll
            for (int y = 0; y < FN; ++y) acc[x][y] = Mfma<T>::op(bf[y], af[x], acc[x][y]);
      }
   }
   // acc[x][y] holds the transposed 16x16 tile: item v of lane l is
   // C[i = l%16][j = row(l, v)], so 16 lanes cover 16 consecutive rows of a
   // column. The C values of one y are all loaded before any is stored.
#pragma unroll
   for (int y = 0; y < FN; ++y) {
      TC cv[FM][4];
      if (MODE != 1) {
#pragma unroll
         for (int x = 0; x < FM; ++x)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
               const int i = wm * WM + x * 16 + (lane & 15);
               const int j = wn * WN + y * 16 + Mfma<T>::row(lane, v);
               const bool ok = i < t.M && j < t.N && j >= t.pad0;
               cv[x][v] = *live_or_zero(ok, c + i + (int64_t)j * ldc);
            }
      }
#pragma unroll
      for (int x = 0; x < FM; ++x)
#pragma unroll
         for (int v = 0; v < 4; ++v) {
            const int i = wm * WM + x * 16 + (lane & 15);
            const int j = wn * WN + y * 16 + Mfma<T>::row(lane, v);
            if (i < t.M && j < t.N && j >= t.pad0)
               c[i + (int64_t)j * ldc] = (MODE == 1) ? (TC)acc[x][y][v] : cv[x][v] - (TC)acc[x][y][v];
         }
   }
}

// NT threads per workgroup: 256 (2 x 2 waves), or 512 (2 x 4 waves, 64 x 32
// per wave) for the far updates: FP64's 135 KB of LDS allow one workgroup
// per CU, and eight waves then keep two per SIMD; FP32 goes from two to four
// waves per SIMD (measured at 80^3: FP64 factor 3.7 ms, FP32 0.4 ms faster).
template <typename T, int MODE, int TB = 128, int TBN = TB, int NT = kThreads>
__global__ __launch_bounds__(NT) void k_update(const GemmTask* __restrict__ tasks,
                                               T* __restrict__ arena,
                                               const T* __restrict__ bsrc) {
   // the panel TRSM and the next pair's columns (64-row tiles) are on the
   // critical path; the far look-ahead updates (128 x 128) are not
   if (TB == 64) __builtin_amdgcn_s_setprio(2);
   const GemmTask t = tasks[xcd_task(blockIdx.x, gridDim.x)];
   if (t.M <= 0) return;
   constexpr int VN = V16<T>::N;
   const bool vec = ((t.a_off | t.b_off | (int64_t)t.lda | (int64_t)t.ldb) % VN) == 0;
   gemm_tile<T, MODE, TB, TBN, T, NT>(t, arena + t.a_off, (MODE == 0 ? arena : bsrc) + t.b_off, arena + t.c_off, t.lda,
                                  vec);
}

// FP32 far updates within 96 VGPRs: two of these eight-wave workgroups per CU
// (4 waves x 96 registers per SIMD lane) leave room for the FP32 diagonal
// factor kernel (108 registers) beside them.
static __global__ __launch_bounds__(2 * kThreads) __attribute__((amdgpu_num_vgpr(96))) void k_update_far_f32(
    const GemmTask* __restrict__ tasks, float* __restrict__ arena) {
   const GemmTask t = tasks[xcd_task(blockIdx.x, gridDim.x)];
   if (t.M <= 0) return;
   const bool vec = ((t.a_off | t.b_off | (int64_t)t.lda | (int64_t)t.ldb) % 4) == 0;
   gemm_tile<float, 0, 128, 128, float, 2 * kThreads>(t, arena + t.a_off, arena + t.b_off, arena + t.c_off, t.lda,
                                                     vec);
}

// Contribution-block update of the indefinite path's FP32 tier: CB -= L21
// (L21 D)^T with both operands rounded to FP32 in the workspace (offsets in
// floats), FP32 MFMA accumulation, the FP64 contribution block in the arena
// (ldc = t.ldc).
static __global__ __launch_bounds__(2 * kThreads) void k_update_cb32(const GemmTask* __restrict__ tasks,
                                                              const float* __restrict__ ws,
                                                              double* __restrict__ arena) {
   const GemmTask t = tasks[xcd_task(blockIdx.x, gridDim.x)];
   if (t.M <= 0) return;
   const bool vec = ((t.a_off | t.b_off | (int64_t)t.lda | (int64_t)t.ldb) % V16<float>::N) == 0;
   gemm_tile<float, 2, 128, 128, double, 2 * kThreads>(t, ws + t.a_off, ws + t.b_off, arena + t.c_off, t.ldc, vec);
}

// ---------------------------------------------------------------------------
// BF16 Schur update for wide fronts in the mixed-precision tier: FP32 front
// storage, the K chunk (32 columns) of A and B rounded to BF16 (round to
// nearest even) as it is staged into LDS in k-contiguous rows, products on the
// 16x16x32 BF16 MFMA accumulated in FP32, subtracted from C at the end. Lane l
// of the MFMA reads row l%16, k = 8*(l/16)..+7 of each operand (one 16-byte
// LDS read). BF16 keeps FP32's 8-bit exponent, so a power-of-two rescaling of
// a block would not change any rounding; none is applied.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

static __global__ __launch_bounds__(kThreads) void k_update_bf16(const GemmTask* __restrict__ tasks,
                                                          float* __restrict__ arena) {
   constexpr int BM = 128, BN = 128, BK = 32, PADK = 8;
   constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
   typedef float f4 __attribute__((ext_vector_type(4)));
   constexpr int AV = BM * BK / 4 / kThreads;  // float4 vectors per thread per chunk
   __shared__ __attribute__((aligned(16))) __bf16 As[2][BM][BK + PADK];
   __shared__ __attribute__((aligned(16))) __bf16 Bs[2][BN][BK + PADK];
   const GemmTask t = tasks[xcd_task(blockIdx.x, gridDim.x)];
   if (t.M <= 0) return;
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int wm = wave >> 1, wn = wave & 1;
   const int ld = t.lda;
   const float* a = arena + t.a_off;
   const float* b = arena + t.b_off;
   float* c = arena + t.c_off;
   v4f acc[FM][FN];
#pragma unroll
   for (int x = 0; x < FM; ++x)
#pragma unroll
      for (int y = 0; y < FN; ++y) acc[x][y] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
   f4 ra[AV], rb[AV];
   auto gload = [&](int k0) {
#pragma unroll
      for (int u = 0; u < AV; ++u) {
         const int v = tid + u * kThreads;
         const int kk = v / (BM / 4), row = (v % (BM / 4)) * 4;
         const int64_t kc = (int64_t)(k0 + kk) * ld;
         ra[u] = *live_or_zero(k0 + kk < t.K && row < t.M, reinterpret_cast<const f4*>(a + row + kc));
         rb[u] = *live_or_zero(k0 + kk < t.K && row < t.N, reinterpret_cast<const f4*>(b + row + kc));
      }
   };
   const int nk = (t.K + BK - 1) / BK;
   gload(0);
   for (int kc = 0; kc < nk; ++kc) {
      const int buf = kc & 1;
#pragma unroll
      for (int u = 0; u < AV; ++u) {
         const int v = tid + u * kThreads;
         const int kk = v / (BM / 4), row = (v % (BM / 4)) * 4;
#pragma unroll
         for (int e = 0; e < 4; ++e) {
            As[buf][row + e][kk] = (__bf16)(ra[u][e]);
            Bs[buf][row + e][kk] = (__bf16)(rb[u][e]);
         }
      }
      __syncthreads();
      if (kc + 1 < nk) gload((kc + 1) * BK);
      bf16x8 af[FM], bf[FN];
#pragma unroll
      for (int x = 0; x < FM; ++x)
         af[x] = *reinterpret_cast<const bf16x8*>(&As[buf][wm * WM + x * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
      for (int y = 0; y < FN; ++y)
         bf[y] = *reinterpret_cast<const bf16x8*>(&Bs[buf][wn * WN + y * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
      for (int x = 0; x < FM; ++x)
#pragma unroll
         for (int y = 0; y < FN; ++y)
            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[y], af[x], acc[x][y], 0, 0, 0);
   }
   // transposed tiles as in k_update: item v of lane l is C[l%16][4(l/16)+v]
#pragma unroll
   for (int y = 0; y < FN; ++y) {
      float cv[FM][4];
#pragma unroll
      for (int x = 0; x < FM; ++x)
#pragma unroll
         for (int v = 0; v < 4; ++v) {
            const int i = wm * WM + x * 16 + (lane & 15);
            const int j = wn * WN + y * 16 + 4 * (lane >> 4) + v;
            cv[x][v] = *live_or_zero(i < t.M && j < t.N && j >= t.pad0, c + i + (int64_t)j * ld);
         }
#pragma unroll
      for (int x = 0; x < FM; ++x)
#pragma unroll
         for (int v = 0; v < 4; ++v) {
            const int i = wm * WM + x * 16 + (lane & 15);
            const int j = wn * WN + y * 16 + 4 * (lane >> 4) + v;
            if (i < t.M && j < t.N && j >= t.pad0) c[i + (int64_t)j * ld] = cv[x][v] - acc[x][y][v];
         }
   }
}

// ---------------------------------------------------------------------------
// One workgroup factorizes a whole front (partial Cholesky over its first n
// columns, Schur complement into the CB), right-looking in panels of kPB
// columns. Per panel: the diagonal block in LDS (the blocked Cholesky and
// inverse of the large-front diagonal kernel), the rows below as one product
// with that inverse (L21 = A21 inv(L11)^T, 16-row strips on MFMA), then the
// trailing lower triangle -= L21 L21^T on MFMA in 32 x 32 tiles with the
// panel read through L1/L2: every trailing entry is read and written once
// per panel, n / kPB passes in all.
// Restates factor_front_posdef's arithmetic (reference src/factor.hxx:19-148,
// src/kernels/factor.hxx:34-263) with a different blocking.
// ---------------------------------------------------------------------------
constexpr int kPB = 64;

template <typename T>
__device__ void front_potrf_block(const NodeDesc& nd, T* __restrict__ F, int* info_slot) {
   constexpr int LD = kPB + 1;
   __shared__ T D[kPB * LD];
   __shared__ T Tmp[kNB * 16];  // scratch of lds_trtri_blocked (ld kNB)
   __shared__ int sh_bad;
   typedef typename Mfma<T>::acc_t acc_t;
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int m = nd.m, n = nd.n, ld = nd.ld;
   for (int k0 = 0; k0 < n; k0 += kPB) {
      const int kb = min(kPB, n - k0), kb16 = (kb + 15) & ~15;
      T* blk = F + k0 + (int64_t)k0 * ld;
      __syncthreads();  // the previous panel's update is complete
      // diagonal block, padded to kb16 with the identity, upper triangle zero
      for (int e = tid; e < kb16 * kb16; e += kThreads) {
         const int i = e % kb16, j = e / kb16;
         const bool live = i < kb && j < kb && i >= j;
         const T v = *live_or_zero(live, blk + i + (int64_t)j * ld);
         D[i + j * LD] = live ? v : ((i == j) ? T(1) : T(0));
      }
      __syncthreads();
      const int bad = lds_potrf_panels<T, LD>(D, kb16, &sh_bad);
      if (bad) {
         if (tid == 0) *info_slot = k0 + bad;
         return;
      }
      for (int e = tid; e < kb * kb; e += kThreads) {
         const int i = e % kb, j = e / kb;
         if (i >= j) blk[i + (int64_t)j * ld] = D[i + j * LD];
      }
      __syncthreads();
      lds_trtri_blocked<T, LD>(D, Tmp, kb16);  // D = inv(L11), lower
      const int r0 = k0 + kb, R = m - r0;
      if (R <= 0) continue;
      // L21 = A21 inv(L11)^T in place. Operands: x = inv(L11)[c][k] (LDS),
      // y = A21[row][k]; item v of lane l is L21[rs + l%16][16 ct + row(l, v)].
      // All K operands of a strip or tile are loaded before its first product:
      // one L2 round trip per tile instead of one per k-step.
      constexpr int KS = kPB / 4;  // k-steps of a full panel
      const int nct = kb16 / 16, nks = kb16 / 4;
      for (int rs = r0 + wave * 16; rs < m; rs += 64) {
         acc_t acc[kPB / 16];
#pragma unroll
         for (int ct = 0; ct < kPB / 16; ++ct) acc[ct] = acc_t{0, 0, 0, 0};
         const int row = rs + (lane & 15);
         T y[KS];
#pragma unroll
         for (int q = 0; q < KS; ++q) {
            const int k = 4 * q + (lane >> 4);
            y[q] = *live_or_zero(row < m && k < kb, F + row + (int64_t)(k0 + k) * ld);
         }
#pragma unroll
         for (int q = 0; q < KS; ++q) {
            if (q >= nks) break;
            const int k = 4 * q + (lane >> 4);
#pragma unroll
            for (int ct = 0; ct < kPB / 16; ++ct)
               if (ct < nct) acc[ct] = Mfma<T>::op(D[(ct * 16 + (lane & 15)) + k * LD], y[q], acc[ct]);
         }
#pragma unroll
         for (int ct = 0; ct < kPB / 16; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
               const int c = ct * 16 + Mfma<T>::row(lane, v);
               if (ct < nct && row < m && c < kb) F[row + (int64_t)(k0 + c) * ld] = acc[ct][v];
            }
      }
      __syncthreads();  // L21 written before the update reads it
      // trailing lower triangle: C[i][j] -= sum_k L21[i][k] L21[j][k] in
      // 32 x 32 tiles (2 x 2 MFMA tiles per wave); operands x = L21[j][k],
      // y = L21[i][k]; item v of lane l is C[i0 + l%16][j0 + row(l, v)].
      const int nt = (R + 31) / 32;
      const int ntiles = nt * (nt + 1) / 2;
      for (int tile = wave; tile < ntiles; tile += 4) {
         int ti = (int)((__builtin_sqrtf(8.0f * tile + 1.0f) - 1.0f) * 0.5f);
         while (ti * (ti + 1) / 2 > tile) --ti;
         while ((ti + 1) * (ti + 2) / 2 <= tile) ++ti;
         const int tj = tile - ti * (ti + 1) / 2;
         const int ri = r0 + ti * 32, rj = r0 + tj * 32;
         acc_t acc[2][2];
#pragma unroll
         for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
         // FP64: the operands in two halves of the panel (register budget)
         constexpr int KC = sizeof(T) == 8 ? KS / 2 : KS;
         for (int q0 = 0; q0 < nks; q0 += KC) {
            T xa[2][KC], yb[2][KC];
#pragma unroll
            for (int q = 0; q < KC; ++q) {
               const int k = 4 * (q0 + q) + (lane >> 4);
#pragma unroll
               for (int u = 0; u < 2; ++u) {
                  const int i = ri + u * 16 + (lane & 15), j = rj + u * 16 + (lane & 15);
                  yb[u][q] = *live_or_zero(i < m && k < kb, F + i + (int64_t)(k0 + k) * ld);
                  xa[u][q] = *live_or_zero(j < m && k < kb, F + j + (int64_t)(k0 + k) * ld);
               }
            }
#pragma unroll
            for (int q = 0; q < KC; ++q) {
               if (q0 + q >= nks) break;
#pragma unroll
               for (int a = 0; a < 2; ++a)
#pragma unroll
                  for (int b = 0; b < 2; ++b) acc[a][b] = Mfma<T>::op(xa[b][q], yb[a][q], acc[a][b]);
            }
         }
         T cv[2][2][4];
#pragma unroll
         for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
               for (int v = 0; v < 4; ++v) {
                  const int i = ri + a * 16 + (lane & 15), j = rj + b * 16 + Mfma<T>::row(lane, v);
                  cv[a][b][v] = *live_or_zero(i < m && j <= i, F + i + (int64_t)j * ld);
               }
#pragma unroll
         for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
               for (int v = 0; v < 4; ++v) {
                  const int i = ri + a * 16 + (lane & 15), j = rj + b * 16 + Mfma<T>::row(lane, v);
                  if (i < m && j <= i) F[i + (int64_t)j * ld] = cv[a][b][v] - acc[a][b][v];
               }
      }
   }
   __syncthreads();
}

// Synthetic symmetric fronts for the batched small-front workload: entry
// (i, j) of front f is uniform in [-1, 1] from a counter-based hash of
// (seed, f, min(i, j), max(i, j)) (splitmix64 finalizer), plus nrow on the
// diagonal; the full square is written (ld = nrow).
__device__ inline uint64_t mix64(uint64_t z) {
   z += 0x9e3779b97f4a7c15ull;
   z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
   z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
   return z ^ (z >> 31);
}
template <typename T>
__global__ __launch_bounds__(kThreads) void k_front_fill(const int* __restrict__ nrow,
                                                         const int* __restrict__ ldv,
                                                         const int64_t* __restrict__ off, T* __restrict__ data,
                                                         uint64_t seed) {
   const int f = blockIdx.y;
   const int m = nrow[f];
   const int64_t ld = ldv ? ldv[f] : m;  // ld >= m; rows m..ld-1 untouched
   T* F = data + off[f];
   const int64_t mm = (int64_t)m * m;
   for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < mm; e += (int64_t)gridDim.x * kThreads) {
      const int i = (int)(e % m), j = (int)(e / m);
      const uint64_t lo = (uint64_t)min(i, j), hi = (uint64_t)max(i, j);
      const uint64_t h = mix64(seed ^ mix64((uint64_t)f * 0x100000001b3ull ^ (hi << 32 | lo)));
      double v = 2.0 * ((double)(h >> 11) * 0x1.0p-53) - 1.0;
      if (i == j) v += (double)m;
      F[i + j * ld] = (T)v;
   }
}

// One workgroup assembles and factorizes a whole front: the lower triangle
// zeroed, the A entries scattered, the children's contribution blocks added
// in list order (one child at a time, so no two writers meet), then the
// partial factorization. Per entry the values arrive in the order of the
// separate k_zero / k_init_a / k_extend_add launches (A value, then the
// children in list order), so the fronts are bitwise the same; the fronts
// of a level take one launch instead of four or more.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_small_asm(const int* __restrict__ node_list,
                                                        const NodeDesc* __restrict__ nodes,
                                                        const int64_t* __restrict__ nlist,
                                                        const int* __restrict__ rlist,
                                                        const double* __restrict__ aval,
                                                        const double* __restrict__ scaling,
                                                        const int* __restrict__ cmap,
                                                        T* __restrict__ arena, int* __restrict__ info) {
   const int node = node_list[blockIdx.x];
   const NodeDesc nd = nodes[node];
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   T* F = arena + nd.off;
   for (int c = wave; c < nd.m; c += 4) {
      T* col = F + (int64_t)c * nd.ld;
      for (int r = c + lane; r < nd.m; r += 64) col[r] = T(0);
   }
   __syncthreads();
   for (int64_t p = threadIdx.x; p < nd.num_a; p += kThreads) {
      const int64_t src = nlist[2 * (nd.amap_off + p)] - 1;
      const int64_t dest = nlist[2 * (nd.amap_off + p) + 1] - 1;
      const int c = (int)(dest / nd.m);
      const int r = (int)(dest - (int64_t)c * nd.m);
      double v = aval[src];
      if (scaling) v = scaling[rlist[nd.rlist_off + r] - 1] * v * scaling[rlist[nd.rlist_off + c] - 1];
      F[(int64_t)c * nd.ld + r] = (T)v;
   }
   for (int ch = nd.first_child; ch != -1;) {
      const NodeDesc c = nodes[ch];
      __syncthreads();  // the previous writer of these entries is done
      const int q = c.m - c.n;
      const int* __restrict__ cm = cmap + c.cmap_off;
      for (int j = wave; j < q; j += 4) {
         T* dst = F + (int64_t)cm[j] * nd.ld;
         const T* src = arena + c.off + (int64_t)(c.n + j) * c.ld + c.n;
         for (int i = j + lane; i < q; i += 64) dst[cm[i]] += src[i];
      }
      ch = c.next_child;
   }
   front_potrf_block<T>(nd, F, info + node);
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
// column-major m x nrhs work vector at solve_off*nrhs (ld = m). Forward:
// gather, then per block [k0, k0+kb): y = inv(L_kk) x_k (in place and into the
// solution), rows below -= L[r, k] y. Backward: gather, own rows -= L21^T x_cb,
// then per block from the last: x_k = inv(L_kk)^T x_k, earlier rows -=
// L[k, j]^T x_k. A task covers rows [r0, r1) of one node (or columns, for the
// transposed products).
// ---------------------------------------------------------------------------

// Forward gather of rows [r0, r1) (at most 256: one row per thread): own
// rows from b, the others zero, then the contribution rows of the children
// that map into the range, children in list order. The child ranges are
// precomputed per task (gch[k0 .. k0+kb) = {child, i0, i1}: rows [i0, i1) of
// the child's contribution block land in [r0, r1), the child map being
// increasing), so no thread searches or walks the child list.
static __global__ __launch_bounds__(kThreads) void k_fwd_gather(const SolveTask* __restrict__ tasks,
                                                         const NodeDesc* __restrict__ nodes,
                                                         const int4* __restrict__ gch,
                                                         const int* __restrict__ rlist,
                                                         const int* __restrict__ cmap,
                                                         const double* __restrict__ b, int ldb,
                                                         double* __restrict__ xloc, int nrhs) {
   const SolveTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   double* xl = xloc + nd.solve_off * nrhs;
   const int m = nd.m;
   const int r = t.r0 + threadIdx.x;
   if (r < t.r1) {
      const int v = (r < nd.n) ? rlist[nd.rlist_off + r] - 1 : 0;
      for (int c0 = 0; c0 < nrhs; c0 += 8) {
         double x[8];
#pragma unroll
         for (int u = 0; u < 8; ++u) x[u] = *live_or_zero(r < nd.n && c0 + u < nrhs, b + v + (int64_t)(c0 + u) * ldb);
#pragma unroll
         for (int u = 0; u < 8; ++u)
            if (c0 + u < nrhs) xl[r + (int64_t)(c0 + u) * m] = x[u];
      }
   }
   for (int k = 0; k < t.kb; ++k) {
      const int4 g = gch[t.k0 + k];
      const NodeDesc cn = nodes[g.x];
      __syncthreads();
      const int i = g.y + threadIdx.x;
      if (i < g.z) {
         const int* cm = cmap + cn.cmap_off;
         const double* cx = xloc + cn.solve_off * nrhs;
         const int pr = cm[i];
         for (int c0 = 0; c0 < nrhs; c0 += 8) {
            double x[8], y[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
               x[u] = *live_or_zero(c0 + u < nrhs, cx + cn.n + i + (int64_t)(c0 + u) * cn.m);
               y[u] = *live_or_zero(c0 + u < nrhs, xl + pr + (int64_t)(c0 + u) * m);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
               if (c0 + u < nrhs) xl[pr + (int64_t)(c0 + u) * m] = y[u] + x[u];
         }
      }
   }
}

// Backward gather of rows [r0, r1) (at most 256) from the current solution.
static __global__ __launch_bounds__(kThreads) void k_bwd_gather(const SolveTask* __restrict__ tasks,
                                                         const NodeDesc* __restrict__ nodes,
                                                         const int* __restrict__ rlist,
                                                         const double* __restrict__ xsol,
                                                         int ldsol, double* __restrict__ xloc,
                                                         int nrhs) {
   const SolveTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   double* xl = xloc + nd.solve_off * nrhs;
   const int m = nd.m;
   const int r = t.r0 + threadIdx.x;
   if (r >= t.r1) return;
   const int v = rlist[nd.rlist_off + r] - 1;
   for (int c0 = 0; c0 < nrhs; c0 += 8) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = *live_or_zero(c0 + u < nrhs, xsol + v + (int64_t)(c0 + u) * ldsol);
#pragma unroll
      for (int u = 0; u < 8; ++u)
         if (c0 + u < nrhs) xl[r + (int64_t)(c0 + u) * m] = x[u];
   }
}

// k index of MFMA step u for lane group kq in the backward solve updates:
// groups of VK = 16 / sizeof(T) consecutive k per lane, so a lane's operand
// for VK steps is one contiguous 16-byte piece of a column of L (the
// forward updates keep k = 4u + kq, contiguous across the rows instead).
template <typename T>
__device__ __forceinline__ int bwd_k(int u, int kq) {
   constexpr int VK = 16 / (int)sizeof(T);
   return (u / VK) * 4 * VK + kq * VK + (u % VK);
}

// One 128-column step of the forward (BWD = false) or backward (BWD = true)
// sweep. Every task needs y = M x_k with M = inv(L_kk) (forward) or
// inv(L_kk)^T (backward, from dinvT), both stored full kb x kb with zeros in
// the other triangle. The operands that do not depend on x (this wave's two
// 16-row tiles of M and its 16-row tile of the update operand) are loaded
// into registers before x is staged, so a step costs about two memory round
// trips. A task with `write` stores y into the solution; forward update
// tasks handle rows [r0, r1) (at most 64: one tile per wave),
// x_r -= L[r, k0:k0+kb] y; backward update tasks handle columns [r0, r1),
// x_j -= L[k0:k0+kb, j]^T y.
template <typename T, bool BWD>
__device__ __forceinline__ void solve_step_dev(const SolveTask& t, const NodeDesc& nd, const T* __restrict__ arena,
                                               const T* __restrict__ dinvs, const int* __restrict__ rlist,
                                               double* __restrict__ xloc, double* __restrict__ xsol, int ldsol,
                                               int nrhs) {
   constexpr int KU = kNB / 4;  // k-steps of the 16x16x4 MFMA over 128
   __shared__ double xs[kNB * 16];
   __shared__ double ys[kNB * 16];
   const int m = nd.m, kb = t.kb;
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const int r = lane & 15, kq = lane >> 4;
   double* xl = xloc + nd.solve_off * nrhs;
   const T* Mk = dinvs + nd.dinv_off + (int64_t)(t.k0 / kNB) * kNB * kNB;
   const T* F = arena + nd.off;
   T mv[2][KU];
#pragma unroll
   for (int q = 0; q < 2; ++q) {
      const int i = (wave + 4 * q) * 16 + r;
#pragma unroll
      for (int u = 0; u < KU; ++u) {
         const int k = 4 * u + kq;
         mv[q][u] = *live_or_zero(i < kb && k < kb, Mk + i + (int64_t)k * kb);
      }
   }
   T av[KU];
   const int ui = t.r0 + wave * 16 + r;  // update row (forward) or column (backward) of this lane
   if (!t.write) {
#pragma unroll
      for (int u = 0; u < KU; ++u) {
         const int k = BWD ? bwd_k<T>(u, kq) : 4 * u + kq;
         const T* p = BWD ? F + t.k0 + k + (int64_t)ui * nd.ld : F + ui + (int64_t)(t.k0 + k) * nd.ld;
         av[u] = *live_or_zero(ui < t.r1 && k < kb, p);
      }
   }
   for (int c0 = 0; c0 < nrhs; c0 += 16) {
      const int nc = min(16, nrhs - c0);
      __syncthreads();
#pragma unroll
      for (int e = threadIdx.x; e < kNB * 16; e += kThreads) {
         const int l = e >> 4, c = e & 15;
         xs[e] = *live_or_zero(l < kb && c < nc, xl + t.k0 + l + (int64_t)(c0 + c) * m);
      }
      // values this task updates, read before the arithmetic (in the row
      // layout of the MFMA that produces the update)
      double cv[4];
      const int ob = t.r0 + wave * 16;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
         const int o = ob + (std::is_same<T, float>::value ? Mfma<float>::row(lane, v) : Mfma<double>::row(lane, v));
         cv[v] = *live_or_zero(!t.write && o < t.r1 && r < nc, xl + o + (int64_t)(c0 + r) * m);
      }
      __syncthreads();
      // y = inv(L_kk) x_k: FP64 MFMA for FP64 factors; FP32 MFMA (twice the
      // rate) for FP32 factors, whose solves feed FP64 refinement
#pragma unroll
      for (int q = 0; q < 2; ++q) {
         const int tile = wave + 4 * q;
         if (tile * 16 >= kb) continue;
         if constexpr (std::is_same<T, float>::value) {
            v4f acc = v4f{0, 0, 0, 0};
#pragma unroll
            for (int u = 0; u < KU; ++u)
               acc = __builtin_amdgcn_mfma_f32_16x16x4f32(mv[q][u], (float)xs[(4 * u + kq) * 16 + r], acc, 0, 0, 0);
#pragma unroll
            for (int v = 0; v < 4; ++v) ys[(tile * 16 + Mfma<float>::row(lane, v)) * 16 + r] = (double)acc[v];
         } else {
            v4d acc = v4d{0, 0, 0, 0};
#pragma unroll
            for (int u = 0; u < KU; ++u)
               acc = __builtin_amdgcn_mfma_f64_16x16x4f64((double)mv[q][u], xs[(4 * u + kq) * 16 + r], acc, 0, 0, 0);
#pragma unroll
            for (int v = 0; v < 4; ++v) ys[(tile * 16 + Mfma<double>::row(lane, v)) * 16 + r] = acc[v];
         }
      }
      __syncthreads();
      if (t.write) {
         for (int e = threadIdx.x; e < kb * nc; e += kThreads) {
            const int i = e % kb, c = e / kb;
            xsol[rlist[nd.rlist_off + t.k0 + i] - 1 + (int64_t)(c0 + c) * ldsol] = ys[i * 16 + c];
         }
         continue;
      }
      if (ob >= t.r1) continue;
      if constexpr (std::is_same<T, float>::value) {
         v4f acc = v4f{0, 0, 0, 0};
#pragma unroll
         for (int u = 0; u < KU; ++u)
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(
                av[u], (float)ys[(BWD ? bwd_k<T>(u, kq) : 4 * u + kq) * 16 + r], acc, 0, 0, 0);
#pragma unroll
         for (int v = 0; v < 4; ++v) {
            const int o = ob + Mfma<float>::row(lane, v);
            if (o < t.r1 && r < nc) xl[o + (int64_t)(c0 + r) * m] = cv[v] - (double)acc[v];
         }
      } else {
         v4d acc = v4d{0, 0, 0, 0};
#pragma unroll
         for (int u = 0; u < KU; ++u)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(
                (double)av[u], ys[(BWD ? bwd_k<T>(u, kq) : 4 * u + kq) * 16 + r], acc, 0, 0, 0);
#pragma unroll
         for (int v = 0; v < 4; ++v) {
            const int o = ob + Mfma<double>::row(lane, v);
            if (o < t.r1 && r < nc) xl[o + (int64_t)(c0 + r) * m] = cv[v] - acc[v];
         }
      }
   }
}


template <typename T, bool BWD>
__global__ __launch_bounds__(kThreads) void k_solve_step(const SolveTask* __restrict__ tasks,
                                                         const NodeDesc* __restrict__ nodes,
                                                         const T* __restrict__ arena,
                                                         const T* __restrict__ dinvs,
                                                         const int* __restrict__ rlist,
                                                         double* __restrict__ xloc,
                                                         double* __restrict__ xsol, int ldsol, int nrhs) {
   const SolveTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   solve_step_dev<T, BWD>(t, nd, arena, dinvs, rlist, xloc, xsol, ldsol, nrhs);
}

// The same step for a launch whose tasks all belong to one front: the front
// and the step come as arguments, so the first loads are the operands
// (workgroup 0 writes the block's solution; workgroup u > 0 updates rows
// (forward) or columns (backward) [r0, r0 + chunk) with r0 = base + (u-1) chunk).
template <typename T, bool BWD>
__global__ __launch_bounds__(kThreads) void k_solve_step1(const NodeDesc nd, int node, int k0, int kb, int base,
                                                          int limit, int chunk, const T* __restrict__ arena,
                                                          const T* __restrict__ dinvs,
                                                          const int* __restrict__ rlist,
                                                          double* __restrict__ xloc, double* __restrict__ xsol,
                                                          int ldsol, int nrhs) {
   SolveTask t{node, k0, kb, 0, 0, 1};
   if (blockIdx.x > 0) {
      t.write = 0;
      t.r0 = base + (int)(blockIdx.x - 1) * chunk;
      t.r1 = min(limit, t.r0 + chunk);
   }
   solve_step_dev<T, BWD>(t, nd, arena, dinvs, rlist, xloc, xsol, ldsol, nrhs);
}

// ---------------------------------------------------------------------------
// Level-wide solves: one launch per level and sweep for the fronts whose
// steps would otherwise take one launch per 128-column block. Workgroup =
// one block of at most 128 rows of one front; it keeps its rows of x in
// registers (FP64, in the MFMA output layout), applies the finished blocks
// of y as they appear and, on a diagonal block, multiplies by the stored
// inverse and publishes its block of y.
//   Forward, diagonal block j (rows [128j, min(n, 128j + 128))):
//     x_j -= L_jk y_k for k = 0 .. j-1 (ascending), y_j = inv(L_jj) x_j.
//   Forward, contribution block (rows at or below n): x -= L_rk y_k for
//     every k -- in the chained launch, after every diagonal block of the
//     level (their work hides behind the chains).
//   Backward, diagonal block j: x_j -= the partial sums of L21^T x2 (the
//     contribution-block product, computed by the launch before, summed in
//     slot order), x_j -= L_ij^T x_i for i = nb-1 .. j+1 (descending),
//     x_j = inv(L_jj)^T x_j.
// Publication: each sweep starts by filling the exchange buffer `ybuf`
// (kNB x 16 values per diagonal block and 16-column chunk of right-hand
// sides) with kYEmpty, a signaling-NaN bit pattern that no arithmetic
// produces. A finished block of y is written there with device-scope
// (write-through) atomic stores; a workgroup that needs it polls its own
// entries with device-scope atomic loads until none holds kYEmpty, so the
// poll and the read of y are the same loads and no flag, cache write-back
// or invalidate is involved. Entries outside the block's rows or the
// chunk's columns are never written and never polled.
// Tasks of a launch are ordered so that every block a workgroup waits for
// is produced by a workgroup of lower index in the same launch (the host
// checks this when it builds the plan): workgroups are dispatched in index
// order, so the lowest-index unfinished workgroup is always resident or
// next to be dispatched, and it waits for nothing unfinished -- the launch
// cannot deadlock. Waits are bounded all the same: a wait that exceeds
// kChainMaxPolls polls (or sees another wait fail) counts an error in `err`
// and goes on, so every workgroup ends; the host reports the error.
// ---------------------------------------------------------------------------
struct ChainTask {
   int node;
   int blk;    // diagonal block index (< nb) or nb for a contribution block
   int nb;     // diagonal blocks of the front
   int r0, r1; // rows of the front this workgroup owns (at most kNB)
   int flag0;  // index of the front's first diagonal block in the exchange buffer
   int cb0;    // backward: first partial-sum slot of the front's contribution-block product
   int cbn;    //   and slots per 64-column chunk (0: no contribution block)
   int g0, gn; // forward: child ranges {child, i0, i1} landing in [r0, r1) (cgch[g0 .. g0+gn))
};

// Forward gather of rows [r0, r0 + 128) of a front into xg (128 x 16, row
// major, rows outside [lo, hi) left zero): b on the pivot rows, zero below,
// then the children's contribution rows in list order -- the order of
// k_fwd_gather, so the sums are the same.
template <int NT>  // threads of the workgroup
__device__ inline void chain_gather(double* xg, const ChainTask& t, const NodeDesc& nd, const NodeDesc* nodes,
                                    const int4* cgch, const int* rlist, const int* cmap, const double* xloc,
                                    const double* b, int ldb, int nrhs, int c0, int nc, int lo, int hi) {
   // row-map loads first, then the values, all in flight together
   constexpr int GU = kNB * 16 / NT;
   int gi[GU];
#pragma unroll
   for (int u = 0; u < GU; ++u) {
      const int e = threadIdx.x + u * NT, l = e >> 4, c = e & 15, row = t.r0 + l;
      gi[u] = (l >= lo && l < hi && row < nd.n && c < nc) ? rlist[nd.rlist_off + row] - 1 : -1;
   }
#pragma unroll
   for (int u = 0; u < GU; ++u) {
      const int e = threadIdx.x + u * NT, c = e & 15;
      xg[e] = gi[u] >= 0 ? b[gi[u] + (int64_t)(c0 + c) * ldb] : 0.0;
   }
   for (int g = t.g0; g < t.g0 + t.gn; ++g) {
      const int4 cg = cgch[g];
      const NodeDesc cn = nodes[cg.x];
      const int* cm = cmap + cn.cmap_off;
      const double* cx = xloc + cn.solve_off * nrhs + cn.n;
      __syncthreads();
      for (int e = threadIdx.x; e < (cg.z - cg.y) * 16; e += NT) {
         const int i = cg.y + (e >> 4), c = e & 15, l = cm[i] - t.r0;
         if (c < nc && l >= lo && l < hi) xg[l * 16 + c] += cx[i + (int64_t)(c0 + c) * cn.m];
      }
   }
   __syncthreads();
}

constexpr int kChainMaxPolls = 1 << 20;
constexpr unsigned long long kYEmpty = 0x7FF4DEADBEEF0000ull;
constexpr int kChainThreads = 512;  // 8 waves: one 16-row tile of the block each


template <typename T, bool BWD>
__global__ __launch_bounds__(kChainThreads) void k_chain(const ChainTask* __restrict__ tasks,
                                                         const NodeDesc* __restrict__ nodes,
                                                         const T* __restrict__ arena, const T* __restrict__ dinvs,
                                                         const int* __restrict__ rlist, double* __restrict__ xloc,
                                                         double* __restrict__ xsol, int ldsol, int nrhs,
                                                         double* __restrict__ ybuf, const double* __restrict__ part,
                                                         int* __restrict__ err, const int4* __restrict__ cgch,
                                                         const int* __restrict__ cmap, const double* __restrict__ b,
                                                         int ldb) {
   constexpr int KU = kNB / 4;                         // k-steps of the 16x16x4 MFMA over 128
   constexpr int YU = kNB * 16 / kChainThreads;        // staged y values per thread
   constexpr uint32_t ES = sizeof(T);
   typedef typename Mfma<T>::acc_t acc_t;
   __shared__ T ys[kNB * 16];  // staged y (rounded to T, as the products use it)
   __shared__ double xg[BWD ? 1 : kNB * 16];  // forward: the block's gathered x
   const ChainTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   const int m = nd.m, n = nd.n, ld = nd.ld, rows = t.r1 - t.r0;
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const int r = lane & 15, kq = lane >> 4;
   const int i = wave * 16 + r;  // block row of this lane's operand loads
   const int nch = (nrhs + 15) / 16;
   double* xl = xloc + nd.solve_off * nrhs;
   // exchange slot of diagonal block b, right-hand-side chunk c
   auto yslot = [&](int b, int c) { return ybuf + ((int64_t)(t.flag0 + b) * nch + c) * (kNB * 16); };
   const bool diag = t.blk < t.nb;
   const int ntile = BWD ? t.nb - 1 - t.blk : min(t.blk, t.nb);
   // the factor columns of the front, [0, n) (the host keeps n * ld * 8 below
   // 2^31); lanes past the block's rows read from the end (zero)
   const uint32_t fbytes = (uint32_t)n * (uint32_t)ld * ES;
   const __amdgpu_buffer_rsrc_t frs = buf_rsrc(arena + nd.off, fbytes);
   // rows of inv(L_jj) for the diagonal step: loaded up front in FP32; in
   // FP64 (twice the registers) after the tile loop
   T mv[KU];
   auto load_inv = [&]() {
      const uint32_t mbytes = (uint32_t)(rows * rows) * ES;
      const __amdgpu_buffer_rsrc_t mrs = buf_rsrc(dinvs + nd.dinv_off + (int64_t)t.blk * kNB * kNB, mbytes);
      const uint32_t lo = i < rows ? (uint32_t)(i + kq * rows) * ES : mbytes;
#pragma unroll
      for (int u = 0; u < KU; ++u) mv[u] = buf_load(mrs, lo + (uint32_t)(4 * u * rows) * ES, (T*)nullptr);
   };
   constexpr bool kInvEarly = sizeof(T) == 4;
   if (kInvEarly && diag) load_inv();
   int sidx[4] = {0, 0, 0, 0};  // solution rows of this lane's rows (diagonal blocks)
   if (diag)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
         const int o = wave * 16 + Mfma<T>::row(lane, v);
         if (o < rows) sidx[v] = rlist[nd.rlist_off + t.r0 + o] - 1;
      }
   // tile s: forward A[i][k] = L[r0 + i][kb0 + k], backward A[i][k] = L[kb0 + k][r0 + i];
   // columns at or beyond n read zero (forward); backward rows at or beyond n
   // read finite entries of L that meet zero rows of y
   T av[KU];
   auto load_tile = [&](int s) {
      const int kb0 = (BWD ? t.nb - 1 - s : s) * kNB;
      if constexpr (BWD) {
         // k = bwd_k(u, kq): 16 contiguous bytes of a column of L per load
         constexpr int VK = 16 / (int)sizeof(T);
         uint32_t lo = (uint32_t)(kb0 + kq * VK + (t.r0 + i) * ld) * ES;
         if (i >= rows) lo = fbytes;
#pragma unroll
         for (int g = 0; g < KU / VK; ++g) {
            const auto raw = __builtin_amdgcn_raw_buffer_load_b128(frs, (int)(lo + (uint32_t)(g * 4 * VK) * ES), 0, 0);
            T piece[VK];
            __builtin_memcpy(piece, &raw, 16);
#pragma unroll
            for (int e = 0; e < VK; ++e) av[g * VK + e] = piece[e];
         }
      } else {
         uint32_t lo = (uint32_t)(t.r0 + i + (kb0 + kq) * ld) * ES;
         const uint32_t step = (uint32_t)(4 * ld) * ES;
         if (i >= rows) lo = fbytes;
#pragma unroll
         for (int u = 0; u < KU; ++u) av[u] = buf_load(frs, lo + (uint32_t)u * step, (T*)nullptr);
      }
   };
   for (int c0 = 0, ci = 0; c0 < nrhs; c0 += 16, ++ci) {
      const int nc = min(16, nrhs - c0);
      // the block's x: backward from the solution (the forward's y on the
      // pivot rows), forward gathered from b and the children
      double xv[4];
      if constexpr (BWD) {
#pragma unroll
         for (int v = 0; v < 4; ++v) {
            const int o = wave * 16 + Mfma<T>::row(lane, v);
            xv[v] = (o < rows && r < nc) ? xsol[sidx[v] + (int64_t)(c0 + r) * ldsol] : 0.0;
         }
      } else {
         if (c0 > 0) __syncthreads();  // the previous chunk is done with xg
         chain_gather<kChainThreads>(xg, t, nd, nodes, cgch, rlist, cmap, xloc, b, ldb, nrhs, c0, nc, 0, rows);
#pragma unroll
         for (int v = 0; v < 4; ++v) xv[v] = xg[(wave * 16 + Mfma<T>::row(lane, v)) * 16 + r];
      }
      if (BWD && t.cbn > 0) {
         // the contribution-block product: partial sums of 64-column chunks
#pragma unroll
         for (int v = 0; v < 4; ++v) {
            const int o = wave * 16 + Mfma<T>::row(lane, v);
            if (o < rows && r < nc) {
               const int col = t.r0 + o;
               const double* pp = part + ((int64_t)(t.cb0 + (col >> 6) * t.cbn) * 64 + (col & 63)) * nrhs + c0 + r;
               const int64_t st = (int64_t)64 * nrhs;
               double sum = 0.0;  // slots summed in order, loads eight at a time
               int k = 0;
               for (; k + 8 <= t.cbn; k += 8) {
                  double pv[8];
#pragma unroll
                  for (int e = 0; e < 8; ++e) pv[e] = pp[(k + e) * st];
#pragma unroll
                  for (int e = 0; e < 8; ++e) sum += pv[e];
               }
               for (; k < t.cbn; ++k) sum += pp[k * st];
               xv[v] -= sum;
            }
         }
      }
      if (ntile > 0) load_tile(0);
      for (int s = 0; s < ntile; ++s) {
         const int kblk = BWD ? t.nb - 1 - s : s, kb0 = kblk * kNB, kbk = min(kNB, n - kb0);
         double yv[YU];
         {
            // poll this thread's entries of block kblk until all are written
            const double* src = yslot(kblk, ci);
            unsigned need = 0;
#pragma unroll
            for (int u = 0; u < YU; ++u) {
               const int e = threadIdx.x + u * kChainThreads, l = e >> 4, c = e & 15;
               yv[u] = 0.0;
               if (l < kbk && c < nc) need |= 1u << u;
            }
            for (int p = 0; need; ++p) {
#pragma unroll
               for (int u = 0; u < YU; ++u)
                  if (need & (1u << u)) {
                     const double v = __hip_atomic_load(src + threadIdx.x + u * kChainThreads, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
                     if (__builtin_bit_cast(unsigned long long, v) != kYEmpty) {
                        yv[u] = v;
                        need &= ~(1u << u);
                     }
                  }
               if (!need) break;
               if (p >= kChainMaxPolls ||
                   ((p & 255) == 255 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
                  __hip_atomic_fetch_add(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                  break;
               }
               __builtin_amdgcn_s_sleep(1);
            }
         }
         __syncthreads();  // every wave is done with the previous ys
#pragma unroll
         for (int u = 0; u < YU; ++u) ys[threadIdx.x + u * kChainThreads] = (T)yv[u];
         __syncthreads();
         acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
         for (int u = 0; u < KU; ++u)
            acc = Mfma<T>::op(av[u], ys[(BWD ? bwd_k<T>(u, kq) : 4 * u + kq) * 16 + r], acc);
#pragma unroll
         for (int v = 0; v < 4; ++v) xv[v] -= (double)acc[v];
         if (s + 1 < ntile) load_tile(s + 1);
      }
      if (!diag) {
#pragma unroll
         for (int v = 0; v < 4; ++v) {
            const int o = wave * 16 + Mfma<T>::row(lane, v);
            if (o < rows && r < nc) xl[t.r0 + o + (int64_t)(c0 + r) * m] = xv[v];
         }
         continue;
      }
      if (!kInvEarly) load_inv();
      __syncthreads();  // every wave is done with ys
#pragma unroll
      for (int v = 0; v < 4; ++v) ys[(wave * 16 + Mfma<T>::row(lane, v)) * 16 + r] = (T)xv[v];
      __syncthreads();
      acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
      for (int u = 0; u < KU; ++u) acc = Mfma<T>::op(mv[u], ys[(4 * u + kq) * 16 + r], acc);
      double* dst = yslot(t.blk, ci);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
         const int o = wave * 16 + Mfma<T>::row(lane, v);
         if (o < rows && r < nc)
            __hip_atomic_store(dst + o * 16 + r, (double)acc[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      // the front's solve vector and the solution (read by later launches)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
         const int o = wave * 16 + Mfma<T>::row(lane, v);
         if (o < rows && r < nc) {
            xl[t.r0 + o + (int64_t)(c0 + r) * m] = (double)acc[v];
            xsol[sidx[v] + (int64_t)(c0 + r) * ldsol] = (double)acc[v];
         }
      }
   }
}

static __global__ void k_fill_y(unsigned long long* p, int64_t n) {
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) p[i] = kYEmpty;
}

// acc += A(16 rows x K) B(K x 16) for small fronts: A from global memory
// (A[i][k] at A[i + k*lda], or A[k + i*lda] when TR; rows >= rvalid and k >=
// K read as zero), B row-major in LDS; loads in batches of eight k-steps.
// Few registers, so that many workgroups share a CU.
template <typename T, bool TR>
__device__ inline typename Mfma<T>::acc_t mma_small(const T* __restrict__ A, int64_t lda, int rvalid, int K,
                                                    const double* Bs, typename Mfma<T>::acc_t acc) {
   const int lane = threadIdx.x & 63;
   const int r = lane & 15, kq = lane >> 4;
   for (int k0 = 0; k0 < K; k0 += 32) {
      T raw[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
         const int kk = k0 + 4 * u + kq;
         const T* p = TR ? A + kk + (int64_t)r * lda : A + r + (int64_t)kk * lda;
         raw[u] = *live_or_zero(r < rvalid && kk < K, p);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
         const int kk = k0 + 4 * u + kq;
         acc = Mfma<T>::op(raw[u], (T)((kk < K) ? Bs[kk * 16 + r] : 0.0), acc);
      }
   }
   return acc;
}

// One workgroup solves a whole small front (one diagonal block, n <= 128,
// m <= 256) in a sweep. Forward: y = inv(L11) x[0:n] to the solution and
// x[n:m] -= L21 y. Backward: x[0:n] -= L21^T x[n:m], then inv(L11)^T x[0:n]
// to the solution. The leaf levels hold thousands of such fronts: this
// kernel keeps its register use low to run several workgroups per CU.
template <typename T, bool BWD>
__global__ __launch_bounds__(kThreads) void k_solve_small(const int* __restrict__ list,
                                                          const NodeDesc* __restrict__ nodes,
                                                          const T* __restrict__ arena,
                                                          const T* __restrict__ dinvs,
                                                          const int* __restrict__ rlist,
                                                          double* __restrict__ xloc,
                                                          double* __restrict__ xsol, int ldsol, int nrhs,
                                                          const double* __restrict__ b, int ldb,
                                                          const int* __restrict__ cmap, int mrows) {
   // LDS sized by the launch to its largest front: x (mrows x 16, mrows =
   // the largest m rounded up to 16, at most 2 kNB) then y (kNB x 16 at
   // most, the largest n); small leaf fronts then fit many workgroups per CU
   extern __shared__ double sdyn[];
   double* xs = sdyn;
   double* ys = sdyn + mrows * 16;
   const NodeDesc nd = nodes[list[blockIdx.x]];
   const int m = nd.m, n = nd.n, q = m - n;
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const int r = lane & 15;
   double* xl = xloc + nd.solve_off * nrhs;
   const T* Mk = dinvs + nd.dinv_off;
   const T* F = arena + nd.off;
   for (int c0 = 0; c0 < nrhs; c0 += 16) {
      const int nc = min(16, nrhs - c0);
      __syncthreads();
      // the front's x, gathered here. Backward: every row from the solution.
      // Forward: b on the pivot rows, zero below, then the contribution
      // rows of the children added in list order.
      // all row-map loads, then all value loads, in flight together
      constexpr int SU = 2 * kNB * 16 / kThreads;
      const int lim = mrows * 16;
      int gi[SU];
#pragma unroll
      for (int u = 0; u < SU; ++u) {
         const int e = threadIdx.x + u * kThreads, l = e >> 4, c = e & 15;
         gi[u] = (e < lim && c < nc && l < (BWD ? m : n)) ? rlist[nd.rlist_off + l] - 1 : -1;
      }
#pragma unroll
      for (int u = 0; u < SU; ++u) {
         const int e = threadIdx.x + u * kThreads, c = e & 15;
         double v = 0.0;
         if (gi[u] >= 0) v = BWD ? xsol[gi[u] + (int64_t)(c0 + c) * ldsol] : b[gi[u] + (int64_t)(c0 + c) * ldb];
         if (e < lim) xs[e] = v;
      }
      if (!BWD)
         for (int ch = nd.first_child; ch != -1;) {
            const NodeDesc cn = nodes[ch];
            const int qc = cn.m - cn.n;
            const int* cm = cmap + cn.cmap_off;
            const double* cx = xloc + cn.solve_off * nrhs + cn.n;
            __syncthreads();
            for (int e = threadIdx.x; e < qc * 16; e += kThreads) {
               const int i = e >> 4, c = e & 15;
               if (c < nc) xs[cm[i] * 16 + c] += cx[i + (int64_t)(c0 + c) * cn.m];
            }
            ch = cn.next_child;
         }
      __syncthreads();
      if (BWD && q > 0) {
         // x[0:n] -= L21^T x[n:m]: output tiles over the n columns
         for (int tile = wave; tile * 16 < n; tile += 4) {
            typename Mfma<T>::acc_t acc = {0, 0, 0, 0};
            acc = mma_small<T, true>(F + n + (int64_t)(tile * 16) * nd.ld, nd.ld, n - tile * 16, q,
                                     xs + n * 16, acc);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
               const int i = tile * 16 + Mfma<T>::row(lane, v);
               if (i < n) xs[i * 16 + r] -= (double)acc[v];
            }
         }
         __syncthreads();
      }
      for (int tile = wave; tile * 16 < n; tile += 4) {
         typename Mfma<T>::acc_t acc = {0, 0, 0, 0};
         acc = mma_small<T, false>(Mk + tile * 16, n, n - tile * 16, n, xs, acc);
#pragma unroll
         for (int v = 0; v < 4; ++v) {
            const int i = tile * 16 + Mfma<T>::row(lane, v);
            if (i < n) ys[i * 16 + r] = (double)acc[v];
         }
      }
      __syncthreads();
      for (int e = threadIdx.x; e < n * nc; e += kThreads) {
         const int i = e % n, c = e / n;
         xsol[rlist[nd.rlist_off + i] - 1 + (int64_t)(c0 + c) * ldsol] = ys[i * 16 + c];
      }
      if (!BWD && q > 0) {
         for (int tile = wave; tile * 16 < q; tile += 4) {
            typename Mfma<T>::acc_t acc = {0, 0, 0, 0};
            acc = mma_small<T, false>(F + n + tile * 16, nd.ld, q - tile * 16, n, ys, acc);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
               const int o = n + tile * 16 + Mfma<T>::row(lane, v);
               if (o < m && r < nc) xl[o + (int64_t)(c0 + r) * m] = xs[o * 16 + r] - (double)acc[v];
            }
         }
      }
   }
}

// Backward contribution-block update split over row chunks: task (node,
// rows [k0, kb) with kb - k0 <= 128, columns [r0, r1) with r1 - r0 <= 64,
// slot `write`) stores part[slot] = L[rows, cols]^T x[rows] (64 x nrhs),
// x[rows] read from the solution through the row map; each wave computes
// one 16-column tile on MFMA with its
// operand prefetched into registers. The reduction below subtracts the
// partial sums of a column chunk in slot order.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_bwd_cbp(const SolveTask* __restrict__ tasks,
                                                      const NodeDesc* __restrict__ nodes,
                                                      const T* __restrict__ arena,
                                                      const double* __restrict__ xsol, int ldsol,
                                                      const int* __restrict__ rlist,
                                                      double* __restrict__ part, int nrhs) {
   constexpr int KU = kNB / 4;
   constexpr int VK = 16 / (int)sizeof(T);
   constexpr uint32_t ES = sizeof(T);
   constexpr int XU = kNB * 16 / kThreads;
   __shared__ double xs[kNB * 16];
   const SolveTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   const int q = t.kb - t.k0;
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const int r = lane & 15, kq = lane >> 4;
   const T* F = arena + nd.off;
   double* out = part + (int64_t)t.write * 64 * nrhs;
   const int j = t.r0 + wave * 16 + r;
   // A[jj][k] = L[k0 + k][j], k in the order bwd_k: a lane's VK consecutive
   // k are 16 contiguous bytes of column j. Rows past the chunk read finite
   // entries of L that meet zero rows of x.
   T av[KU];
   if ((int64_t)nd.n * nd.ld * ES < (int64_t(1) << 31)) {
      const uint32_t fbytes = (uint32_t)nd.n * (uint32_t)nd.ld * ES;
      const __amdgpu_buffer_rsrc_t frs = buf_rsrc(F, fbytes);
      const uint32_t lo = j < t.r1 ? (uint32_t)(t.k0 + kq * VK + j * nd.ld) * ES : fbytes;
#pragma unroll
      for (int g = 0; g < KU / VK; ++g) {
         const auto raw = __builtin_amdgcn_raw_buffer_load_b128(frs, (int)(lo + (uint32_t)(g * 4 * VK) * ES), 0, 0);
         T piece[VK];
         __builtin_memcpy(piece, &raw, 16);
#pragma unroll
         for (int e = 0; e < VK; ++e) av[g * VK + e] = piece[e];
      }
   } else {
#pragma unroll
      for (int u = 0; u < KU; ++u) {
         const int k = bwd_k<T>(u, kq);
         av[u] = *live_or_zero(j < t.r1 && k < q, F + t.k0 + k + (int64_t)j * nd.ld);
      }
   }
   for (int c0 = 0; c0 < nrhs; c0 += 16) {
      const int nc = min(16, nrhs - c0);
      // x of the chunk's rows from the solution through the row map: all
      // row-map loads, then all solution loads, in flight together
      int gi[XU];
      double xv[XU];
#pragma unroll
      for (int u = 0; u < XU; ++u) {
         const int l = (threadIdx.x + u * kThreads) >> 4;
         gi[u] = l < q ? rlist[nd.rlist_off + t.k0 + l] - 1 : -1;
      }
#pragma unroll
      for (int u = 0; u < XU; ++u) {
         const int c = (threadIdx.x + u * kThreads) & 15;
         xv[u] = (gi[u] >= 0 && c < nc) ? xsol[gi[u] + (int64_t)(c0 + c) * ldsol] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < XU; ++u) xs[threadIdx.x + u * kThreads] = xv[u];
      __syncthreads();
      typename Mfma<T>::acc_t acc = {0, 0, 0, 0};
#pragma unroll
      for (int u = 0; u < KU; ++u) acc = Mfma<T>::op(av[u], (T)xs[bwd_k<T>(u, kq) * 16 + r], acc);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
         const int jj = wave * 16 + Mfma<T>::row(lane, v);
         if (r < nc) out[(int64_t)jj * nrhs + c0 + r] = (double)acc[v];
      }
   }
}

// x_j -= sum over slots [k0, k0 + kb) of part[slot][j - r0] for the columns
// [r0, r1) of one node, slots summed in order.
static __global__ __launch_bounds__(kThreads) void k_bwd_cbr(const SolveTask* __restrict__ tasks,
                                                      const NodeDesc* __restrict__ nodes,
                                                      const double* __restrict__ part,
                                                      double* __restrict__ xloc, int nrhs) {
   const SolveTask t = tasks[blockIdx.x];
   const NodeDesc nd = nodes[t.node];
   double* xl = xloc + nd.solve_off * nrhs;
   const int w = t.r1 - t.r0;
   for (int e = threadIdx.x; e < w * nrhs; e += kThreads) {
      const int jj = e % w, c = e / w;
      double s = 0.0;
      for (int k = 0; k < t.kb; ++k) s += part[((int64_t)(t.k0 + k) * 64 + jj) * nrhs + c];
      xl[t.r0 + jj + (int64_t)c * nd.m] -= s;
   }
}

// ---------------------------------------------------------------------------
// FP64 refinement helpers.
// ---------------------------------------------------------------------------
// Refinement step, fused: x += d (both n x nrhs, column-major), a copy of x
// in row-major 16-column chunks (xrm[chunk][i][16], read by k_residual_rm:
// one row's 16 values in one 128-byte line) and the workgroup's max |x| per
// column into part[chunk][blockIdx.x][16] (reduced by k_max_parts; no
// atomics: many workgroups contending for 16 addresses serialize).
// A workgroup walks 64-row x 16-column tiles of one chunk (grid-stride),
// transposed through LDS; wave w owns columns w, w+4, w+8, w+12 of the tile.
static __global__ __launch_bounds__(kThreads) void k_axpy_rm(int n, int nrhs, const double* __restrict__ d,
                                                      double* __restrict__ x, double* __restrict__ xrm,
                                                      double* __restrict__ part) {
   __shared__ double tile[64][17];
   const int ch = blockIdx.y, c0 = ch * 16;
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   double mx[4] = {0.0, 0.0, 0.0, 0.0};
   for (int i0 = blockIdx.x * 64; i0 < n; i0 += gridDim.x * 64) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
         const int c = wave + 4 * k, i = i0 + lane;
         double v = 0.0;
         if (i < n && c0 + c < nrhs) {
            const int64_t idx = i + (int64_t)(c0 + c) * n;
            v = x[idx] + d[idx];
            x[idx] = v;
         }
         tile[lane][c] = v;
         mx[k] = fmax(mx[k], fabs(v));
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 4; ++k) {
         const int e = threadIdx.x + k * kThreads, il = e >> 4, c = e & 15;
         if (i0 + il < n) xrm[((int64_t)ch * n + i0 + il) * 16 + c] = tile[il][c];
      }
      __syncthreads();
   }
#pragma unroll
   for (int k = 0; k < 4; ++k) {
      double m = mx[k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
      if (lane == 0) part[((int64_t)ch * gridDim.x + blockIdx.x) * 16 + wave + 4 * k] = m;
   }
}

// r = b - A x with A in CSR (both triangles, pivot order) and x row-major
// (xrm, 16-column chunks); four lanes per row, four right-hand sides each
// (one 32-byte piece of the row's 128-byte line of x per entry), entries in
// row order; rows grid-stride, the workgroup's max |r| per column into
// part[chunk][blockIdx.x][16] as in k_axpy_rm.
static __global__ __launch_bounds__(kThreads) void k_residual_rm(int n, const int64_t* __restrict__ rowptr,
                                                          const int* __restrict__ col,
                                                          const double* __restrict__ val,
                                                          const double* __restrict__ xrm,
                                                          const double* __restrict__ b, double* __restrict__ r,
                                                          int nrhs, double* __restrict__ part) {
   __shared__ double wmax[4][16];
   const int q = threadIdx.x & 3, ch = blockIdx.y, cq = ch * 16 + 4 * q;
   const double* xc = xrm + (int64_t)ch * n * 16 + 4 * q;
   double mx[4] = {0.0, 0.0, 0.0, 0.0};
   for (int i = (blockIdx.x * kThreads + threadIdx.x) >> 2; i < n; i += (gridDim.x * kThreads) >> 2) {
      double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int e = 0; e < 4; ++e)
         if (cq + e < nrhs) s[e] = b[i + (int64_t)(cq + e) * n];
      const int64_t p0 = rowptr[i], p1 = rowptr[i + 1];
      for (int64_t p = p0; p < p1; ++p) {
         const double v = val[p];
         const v4d xv = *reinterpret_cast<const v4d*>(xc + (int64_t)col[p] * 16);
#pragma unroll
         for (int e = 0; e < 4; ++e) s[e] -= v * xv[e];
      }
#pragma unroll
      for (int e = 0; e < 4; ++e)
         if (cq + e < nrhs) {
            r[i + (int64_t)(cq + e) * n] = s[e];
            mx[e] = fmax(mx[e], fabs(s[e]));
         }
   }
   // lanes of equal q hold the same four columns
#pragma unroll
   for (int e = 0; e < 4; ++e) {
      double m = mx[e];
#pragma unroll
      for (int o = 32; o >= 4; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
      if ((threadIdx.x & 63) < 4) wmax[threadIdx.x >> 6][4 * q + e] = m;
   }
   __syncthreads();
   if (threadIdx.x < 16)
      part[((int64_t)ch * gridDim.x + blockIdx.x) * 16 + threadIdx.x] =
          fmax(fmax(wmax[0][threadIdx.x], wmax[1][threadIdx.x]), fmax(wmax[2][threadIdx.x], wmax[3][threadIdx.x]));
}

// out[c] = max over the nparts partial maxima of column c (part[chunk][p][16],
// c = 16 chunk + column), as the ordered bits of a non-negative double.
// One workgroup per column.
static __global__ __launch_bounds__(kThreads) void k_max_parts(const double* __restrict__ part, int nparts,
                                                        int nrhs, unsigned long long* __restrict__ out) {
   const int c = blockIdx.x, ch = c >> 4, cc = c & 15;
   double m = 0.0;
   for (int p = threadIdx.x; p < nparts; p += kThreads) m = fmax(m, part[((int64_t)ch * nparts + p) * 16 + cc]);
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
   __shared__ double wm[4];
   if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
   __syncthreads();
   if (threadIdx.x == 0 && c < nrhs)
      out[c] = (unsigned long long)__double_as_longlong(fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3])));
}

// Per-column max |v| into out[c] (as ordered bits of a non-negative double).
static __global__ __launch_bounds__(kThreads) void k_colmax(int n, const double* __restrict__ v, int nrhs,
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

// y = A x (both triangles, pivot order), one row per thread.
static __global__ __launch_bounds__(kThreads) void k_spmv(int n, const int64_t* __restrict__ rowptr,
                                                   const int* __restrict__ col,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ x,
                                                   double* __restrict__ y, int nrhs) {
   const int i = blockIdx.x * kThreads + threadIdx.x;
   if (i >= n) return;
   const int64_t p0 = rowptr[i], p1 = rowptr[i + 1];
   for (int c0 = 0; c0 < nrhs; c0 += 16) {  // 16 right-hand sides per pass
      const int nc = min(16, nrhs - c0);
      double s[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) s[c] = 0.0;
      for (int64_t p = p0; p < p1; ++p) {
         const double v = val[p];
         const double* xj = x + col[p] + (int64_t)c0 * n;
#pragma unroll
         for (int c = 0; c < 16; ++c) s[c] += v * xj[(int64_t)min(c, nc - 1) * n];
      }
#pragma unroll
      for (int c = 0; c < 16; ++c)
         if (c < nc) y[i + (int64_t)(c0 + c) * n] = s[c];
   }
}

// Column dot products <x_c, y_c>, reproducible: a fixed grid of kDotBlocks
// blocks per column writes one partial each (fixed thread-to-element map,
// fixed tree), then k_dot_final adds the partials in index order.
constexpr int kDotBlocks = 256;
// grid (kDotBlocks, nrhs, nvec): <x_c, (y + z*ystride)_c> for vector z.
static __global__ __launch_bounds__(kThreads) void k_dot_partial(int n, const double* __restrict__ x,
                                                          const double* __restrict__ ybase,
                                                          int64_t ystride,
                                                          double* __restrict__ part) {
   const int c = blockIdx.y;
   const double* y = ybase + (int64_t)blockIdx.z * ystride;
   double s = 0.0;
   for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += kDotBlocks * kThreads)
      s += x[i + (int64_t)c * n] * y[i + (int64_t)c * n];
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
   __shared__ double ws[4];
   if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
   __syncthreads();
   if (threadIdx.x == 0)
      part[((int64_t)blockIdx.z * gridDim.y + c) * kDotBlocks + blockIdx.x] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

// count = nvec * nrhs sums of kDotBlocks partials, in index order.
static __global__ void k_dot_final(const double* __restrict__ part, int count, double* __restrict__ out) {
   const int c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= count) return;
   double s = 0.0;
   for (int b = 0; b < kDotBlocks; ++b) s += part[c * kDotBlocks + b];
   out[c] = s;
}

// y_c += a_c x_c with per-column coefficients in device memory.
static __global__ __launch_bounds__(kThreads) void k_axpy_cols(int n, int nrhs, const double* __restrict__ a,
                                                        const double* __restrict__ x,
                                                        double* __restrict__ y) {
   const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (e >= (int64_t)n * nrhs) return;
   y[e] += a[e / n] * x[e];
}

// W_c -= sum_{z < nvec} h[z*nrhs + c] V_z,c (the Gram-Schmidt update).
static __global__ __launch_bounds__(kThreads) void k_multi_axpy(int n, int nrhs, int nvec,
                                                         const double* __restrict__ h,
                                                         const double* __restrict__ V,
                                                         int64_t vstride, double* __restrict__ W) {
   const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (e >= (int64_t)n * nrhs) return;
   const int c = (int)(e / n);
   double s = W[e];
   for (int z = 0; z < nvec; ++z) s -= h[z * nrhs + c] * V[z * vstride + e];
   W[e] = s;
}

// y_c = a_c x_c.
static __global__ __launch_bounds__(kThreads) void k_scale_cols(int n, int nrhs, const double* __restrict__ a,
                                                         const double* __restrict__ x,
                                                         double* __restrict__ y) {
   const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (e >= (int64_t)n * nrhs) return;
   y[e] = a[e / n] * x[e];
}

static __global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
   int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) p[i] = v;
}

static __global__ __launch_bounds__(kThreads) void k_axpy(int64_t n, double a, const double* __restrict__ x,
                                                   double* __restrict__ y) {
   int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (i < n) y[i] += a * x[i];
}

static __global__ __launch_bounds__(kThreads) void k_copy(int64_t n, const double* __restrict__ x,
                                                   double* __restrict__ y) {
   for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
      y[i] = x[i];
}

// Grid-stride: one dispatch covers at most 2^32 work-items, so callers cap
// the grid for large n.
static __global__ __launch_bounds__(kThreads) void k_fill(int64_t n, double v, double* __restrict__ y) {
   for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
      y[i] = v;
}

// Caller's order -> pivot order, optionally scaled (the solve driver's
// x2(i,r) = x(invp(i),r) * scaling(i), src/spldlt_factorize_mod.F90:951-966):
// out[i + c n] = x[invp[i]-1 + c ldx] (* s[i]).
static __global__ __launch_bounds__(kThreads) void k_perm_in(int n, int nrhs, const int* __restrict__ invp,
                                                      const double* __restrict__ s,
                                                      const double* __restrict__ x, int64_t ldx,
                                                      double* __restrict__ out) {
   const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (e >= (int64_t)n * nrhs) return;
   const int c = (int)(e / n), i = (int)(e - (int64_t)c * n);
   double v = x[(int64_t)(invp[i] - 1) + c * ldx];
   if (s) v *= s[i];
   out[e] = v;
}

// Pivot order -> caller's order, optionally scaled (:1039-1053):
// out[invp[i]-1 + c ldo] = x[i + c n] (* s[i]).
static __global__ __launch_bounds__(kThreads) void k_perm_out(int n, int nrhs, const int* __restrict__ invp,
                                                       const double* __restrict__ s,
                                                       const double* __restrict__ x,
                                                       double* __restrict__ out, int64_t ldo) {
   const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (e >= (int64_t)n * nrhs) return;
   const int c = (int)(e / n), i = (int)(e - (int64_t)c * n);
   double v = x[e];
   if (s) v *= s[i];
   out[(int64_t)(invp[i] - 1) + c * ldo] = v;
}

// y[i + c n] = s[i] x[i + c n] (rows scaled; y may alias x).
static __global__ __launch_bounds__(kThreads) void k_scale_rows(int n, int nrhs, const double* __restrict__ s,
                                                         const double* x, double* y) {
   const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (e >= (int64_t)n * nrhs) return;
   y[e] = s[e % n] * x[e];
}

// ||A||_inf with full symmetric row sums (tests/common.hxx:1003-1050): the
// CSR holds both triangles, so a row's sum is |A| summed over the CSR row.
// Grid-stride over rows; each workgroup leaves its maximum in part[blockIdx.x].
static __global__ __launch_bounds__(kThreads) void k_csr_rowabs(int n, const int64_t* __restrict__ rowptr,
                                                         const double* __restrict__ val,
                                                         double* __restrict__ part) {
   double m = 0.0;
   for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
      double s = 0.0;
      for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) s += fabs(val[p]);
      m = fmax(m, s);
   }
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
   __shared__ double wm[4];
   if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
   __syncthreads();
   if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
}

// Norm equilibration (MC77-like, infinity norm; Ruiz 2001), one sweep's row
// maxima of |S A S| over the full CSR in pivot order: each entry's scaled
// value is formed as the reference's driver loop forms it for the stored
// lower-triangle entry (row r >= column c in the caller's order):
// (s_r * |a|) * s_c, so that both triangles round alike.
static __global__ __launch_bounds__(kThreads) void k_equil_rowmax(int n, const int64_t* __restrict__ rowptr,
                                                           const int* __restrict__ col,
                                                           const double* __restrict__ val,
                                                           const int* __restrict__ invp,
                                                           const double* __restrict__ s,
                                                           double* __restrict__ rmax) {
   const int i = blockIdx.x * kThreads + threadIdx.x;
   if (i >= n) return;
   const int oi = invp[i];
   const double si = s[i];
   double m = 0.0;
   for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const int j = col[p];
      const double a = fabs(val[p]), sj = s[j];
      const double v = oi >= invp[j] ? (si * a) * sj : (sj * a) * si;
      m = fmax(m, v);
   }
   rmax[i] = m;
}

// s_i /= sqrt(rmax_i) where rmax_i > 0; per-workgroup max |1 - rmax_i| into
// part[blockIdx.x] (the convergence test).
static __global__ __launch_bounds__(kThreads) void k_equil_update(int n, const double* __restrict__ rmax,
                                                           double* __restrict__ s, double* __restrict__ part) {
   const int i = blockIdx.x * kThreads + threadIdx.x;
   double dev = 0.0;
   if (i < n) {
      const double r = rmax[i];
      if (r > 0.0) s[i] = s[i] / sqrt(r);
      dev = fabs(1.0 - r);
   }
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) dev = fmax(dev, __shfl_xor(dev, o, 64));
   __shared__ double wm[4];
   if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = dev;
   __syncthreads();
   if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
}

// out[0] = max of part[0..cnt) (one workgroup).
static __global__ __launch_bounds__(kThreads) void k_max_reduce(const double* __restrict__ part, int cnt,
                                                         double* __restrict__ out) {
   double m = 0.0;
   for (int p = threadIdx.x; p < cnt; p += kThreads) m = fmax(m, part[p]);
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
   __shared__ double wm[4];
   if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
   __syncthreads();
   if (threadIdx.x == 0) out[0] = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
}

// CSR values gathered from the cleaned input values.
static __global__ __launch_bounds__(kThreads) void k_gather_vals(int64_t nnz, const int64_t* __restrict__ src,
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
