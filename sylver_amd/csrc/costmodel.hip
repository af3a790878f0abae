// Learned cost model of the front factorization: an 8 -> 64 -> 64 -> 2 MLP
// (ReLU, FP32) that predicts log2 of a front's factorization time for the
// two kernel variants (one-workgroup kernel, tiled path) from the front's
// shape features. Trained on the device by the kernels below (forward,
// mean-squared-error backward, Adam), with the autotuner's timings as
// targets; inference proposes each front's variant. The reference has no
// counterpart: its scheduler decides by static rules.
//
// Parameter layout (floats): W1[8][64], b1[64], W2[64][64], b2[64],
// W3[64][2], b3[2] (row index = input unit).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "kernels.hpp"
#include "numeric.hpp"
#include "numeric_impl.hpp"

namespace sylver_amd {

namespace {

constexpr int I = kMlpIn, H = kMlpHid, O = kMlpOut;
constexpr int oW1 = 0, oB1 = oW1 + I * H, oW2 = oB1 + H, oB2 = oW2 + H * H, oW3 = oB2 + H, oB3 = oW3 + H * O;
static_assert(oB3 + O == kMlpParams, "parameter layout");

// Layer with ReLU: out[b][j] = max(0, bias[j] + sum_k in[b][k] W[k][j]).
template <int K, int N, bool RELU>
__global__ __launch_bounds__(kThreads) void k_mlp_layer(const float* __restrict__ in, const float* __restrict__ P,
                                                        int ow, int ob, float* __restrict__ out, int B) {
   const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (e >= (int64_t)B * N) return;
   const int b = (int)(e / N), j = (int)(e % N);
   float z = P[ob + j];
   for (int k = 0; k < K; ++k) z = __builtin_fmaf(in[(int64_t)b * K + k], P[ow + k * N + j], z);
   out[e] = RELU ? fmaxf(z, 0.0f) : z;
}

// dY = 2 (Y - T) / (B O): gradient of the loss sum (Y - T)^2 / (B O).
__global__ __launch_bounds__(kThreads) void k_mlp_dy(const float* __restrict__ Y, const float* __restrict__ T,
                                                     float* __restrict__ dY, int B) {
   const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (e >= (int64_t)B * O) return;
   dY[e] = 2.0f * (Y[e] - T[e]) / (float)(B * O);
}

// Loss sum_e (Y - T)^2 / (B O), one workgroup, fixed order; added to *acc.
__global__ __launch_bounds__(kThreads) void k_mlp_loss(const float* __restrict__ Y, const float* __restrict__ T,
                                                       int B, float* __restrict__ acc) {
   __shared__ double part[kThreads];
   double s = 0.0;
   for (int64_t e = threadIdx.x; e < (int64_t)B * O; e += kThreads) {
      const double d = (double)Y[e] - (double)T[e];
      s += d * d;
   }
   part[threadIdx.x] = s;
   __syncthreads();
   if (threadIdx.x == 0) {
      double t = 0.0;
      for (int k = 0; k < kThreads; ++k) t += part[k];
      *acc += (float)(t / ((double)B * O));
   }
}

// Backward through a layer: dIn[b][k] = (sum_j dOut[b][j] W[k][j]) * (act[b][k] > 0).
template <int K, int N>
__global__ __launch_bounds__(kThreads) void k_mlp_back(const float* __restrict__ dOut, const float* __restrict__ P,
                                                       int ow, const float* __restrict__ act,
                                                       float* __restrict__ dIn, int B) {
   const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
   if (e >= (int64_t)B * K) return;
   const int b = (int)(e / K), k = (int)(e % K);
   float s = 0.0f;
   for (int j = 0; j < N; ++j) s = __builtin_fmaf(dOut[(int64_t)b * N + j], P[ow + k * N + j], s);
   dIn[e] = act[e] > 0.0f ? s : 0.0f;
}

// Parameter gradients, one thread per parameter, batch summed in order.
__global__ __launch_bounds__(kThreads) void k_mlp_pgrad(const float* __restrict__ X, const float* __restrict__ H1,
                                                        const float* __restrict__ H2, const float* __restrict__ dH1,
                                                        const float* __restrict__ dH2, const float* __restrict__ dY,
                                                        float* __restrict__ G, int B) {
   const int p = blockIdx.x * kThreads + threadIdx.x;
   if (p >= kMlpParams) return;
   const float *a = nullptr, *d = nullptr;
   int ka = 0, kd = 0, ia = 0, id = 0;  // a[b][ia] * d[b][id], row lengths ka, kd; a == nullptr: bias
   if (p < oB1) {
      a = X, ka = I, ia = (p - oW1) / H, d = dH1, kd = H, id = (p - oW1) % H;
   } else if (p < oW2) {
      d = dH1, kd = H, id = p - oB1;
   } else if (p < oB2) {
      a = H1, ka = H, ia = (p - oW2) / H, d = dH2, kd = H, id = (p - oW2) % H;
   } else if (p < oW3) {
      d = dH2, kd = H, id = p - oB2;
   } else if (p < oB3) {
      a = H2, ka = H, ia = (p - oW3) / O, d = dY, kd = O, id = (p - oW3) % O;
   } else {
      d = dY, kd = O, id = p - oB3;
   }
   float s = 0.0f;
   for (int b = 0; b < B; ++b) {
      const float dv = d[(int64_t)b * kd + id];
      s = a ? __builtin_fmaf(a[(int64_t)b * ka + ia], dv, s) : s + dv;
   }
   G[p] = s;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8), step t >= 1.
__global__ __launch_bounds__(kThreads) void k_adam(float* __restrict__ P, const float* __restrict__ G,
                                                   float* __restrict__ M, float* __restrict__ V, float lr, int t) {
   const int p = blockIdx.x * kThreads + threadIdx.x;
   if (p >= kMlpParams) return;
   const float b1 = 0.9f, b2 = 0.999f;
   const float g = G[p];
   const float m = b1 * M[p] + (1.0f - b1) * g, v = b2 * V[p] + (1.0f - b2) * g * g;
   M[p] = m;
   V[p] = v;
   const float mh = m / (1.0f - powf(b1, (float)t)), vh = v / (1.0f - powf(b2, (float)t));
   P[p] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kThreads - 1) / kThreads); }

struct MlpBuffers {
   float *P = nullptr, *X = nullptr, *T = nullptr, *H1 = nullptr, *H2 = nullptr, *Y = nullptr;
   float *dY = nullptr, *dH1 = nullptr, *dH2 = nullptr, *G = nullptr, *M = nullptr, *V = nullptr, *loss = nullptr;
   explicit MlpBuffers(int B) {
      P = dalloc<float>(kMlpParams);
      X = dalloc<float>((int64_t)B * I);
      T = dalloc<float>((int64_t)B * O);
      H1 = dalloc<float>((int64_t)B * H);
      H2 = dalloc<float>((int64_t)B * H);
      Y = dalloc<float>((int64_t)B * O);
      dY = dalloc<float>((int64_t)B * O);
      dH1 = dalloc<float>((int64_t)B * H);
      dH2 = dalloc<float>((int64_t)B * H);
      G = dalloc<float>(kMlpParams);
      M = dalloc<float>(kMlpParams);
      V = dalloc<float>(kMlpParams);
      loss = dalloc<float>(1024);
   }
   ~MlpBuffers() {
      for (float* p : {P, X, T, H1, H2, Y, dY, dH1, dH2, G, M, V, loss}) dfree(p);
   }
};

// Forward of B rows of X already on the device (X, H1, H2, Y of `w`).
void forward(MlpBuffers& w, const float* X, int B, hipStream_t s) {
   k_mlp_layer<I, H, true><<<grid_for((int64_t)B * H), kThreads, 0, s>>>(X, w.P, oW1, oB1, w.H1, B);
   k_mlp_layer<H, H, true><<<grid_for((int64_t)B * H), kThreads, 0, s>>>(w.H1, w.P, oW2, oB2, w.H2, B);
   k_mlp_layer<H, O, false><<<grid_for((int64_t)B * O), kThreads, 0, s>>>(w.H2, w.P, oW3, oB3, w.Y, B);
}

// Backward of the mean squared error against T; gradients into w.G.
void backward(MlpBuffers& w, const float* X, const float* T, int B, hipStream_t s) {
   k_mlp_dy<<<grid_for((int64_t)B * O), kThreads, 0, s>>>(w.Y, T, w.dY, B);
   k_mlp_back<H, O><<<grid_for((int64_t)B * H), kThreads, 0, s>>>(w.dY, w.P, oW3, w.H2, w.dH2, B);
   k_mlp_back<H, H><<<grid_for((int64_t)B * H), kThreads, 0, s>>>(w.dH2, w.P, oW2, w.H1, w.dH1, B);
   k_mlp_pgrad<<<grid_for(kMlpParams), kThreads, 0, s>>>(X, w.H1, w.H2, w.dH1, w.dH2, w.dY, w.G, B);
}

uint64_t splitmix(uint64_t& x) {
   uint64_t z = (x += 0x9e3779b97f4a7c15ull);
   z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
   z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
   return z ^ (z >> 31);
}

}  // namespace

// Glorot-uniform weights from a splitmix64 stream, zero biases (host).
int mlp_init(float* params, unsigned long long seed) {
   if (!params) return SYLVER_ERROR_CALL_SEQUENCE;
   uint64_t x = seed;
   auto fill = [&](int off, int fan_in, int fan_out) {
      const double a = std::sqrt(6.0 / (fan_in + fan_out));
      for (int k = 0; k < fan_in * fan_out; ++k)
         params[off + k] = (float)(a * (2.0 * ((splitmix(x) >> 11) * 0x1.0p-53) - 1.0));
   };
   std::fill(params, params + kMlpParams, 0.0f);
   fill(oW1, I, H);
   fill(oW2, H, H);
   fill(oW3, H, O);
   return 0;
}

int mlp_grad(const float* params, const float* X, const float* Y, int B, float* grads, float* loss) {
   if (!params || !X || !Y || B < 1 || !grads) return SYLVER_ERROR_CALL_SEQUENCE;
   if (device_count() <= 0) return SYLVER_ERROR_CUDA_UNKNOWN;
   try {
      MlpBuffers w(B);
      HIPCHECK(hipMemcpy(w.P, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(w.X, X, (size_t)B * I * sizeof(float), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(w.T, Y, (size_t)B * O * sizeof(float), hipMemcpyHostToDevice));
      HIPCHECK(hipMemset(w.loss, 0, sizeof(float)));
      forward(w, w.X, B, nullptr);
      k_mlp_loss<<<1, kThreads>>>(w.Y, w.T, B, w.loss);
      backward(w, w.X, w.T, B, nullptr);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipMemcpy(grads, w.G, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
      if (loss) HIPCHECK(hipMemcpy(loss, w.loss, sizeof(float), hipMemcpyDeviceToHost));
      return 0;
   } catch (const HipError&) {
      return SYLVER_ERROR_CUDA_UNKNOWN;
   }
}

// Minibatch Adam on the device; the sample order of every epoch is a seeded
// shuffle; loss_hist[e] = mean minibatch loss of epoch e.
int mlp_train(float* params, const float* X, const float* Y, int N, int epochs, int batch, float lr,
              unsigned long long seed, float* loss_hist) {
   if (!params || !X || !Y || N < 1 || epochs < 1 || batch < 1 || epochs > 1024) return SYLVER_ERROR_CALL_SEQUENCE;
   if (device_count() <= 0) return SYLVER_ERROR_CUDA_UNKNOWN;
   try {
      const int B = std::min(batch, N);
      MlpBuffers w(N);
      HIPCHECK(hipMemcpy(w.P, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
      HIPCHECK(hipMemset(w.M, 0, kMlpParams * sizeof(float)));
      HIPCHECK(hipMemset(w.V, 0, kMlpParams * sizeof(float)));
      HIPCHECK(hipMemset(w.loss, 0, 1024 * sizeof(float)));
      std::vector<int> perm(N);
      std::vector<float> xs((size_t)N * I), ys((size_t)N * O);
      uint64_t st = seed;
      int t = 0;
      std::vector<int> nb(epochs, 0);
      for (int e = 0; e < epochs; ++e) {
         for (int i = 0; i < N; ++i) perm[i] = i;
         for (int i = N - 1; i > 0; --i) std::swap(perm[i], perm[splitmix(st) % (uint64_t)(i + 1)]);
         for (int i = 0; i < N; ++i) {
            std::copy(X + (size_t)perm[i] * I, X + (size_t)perm[i] * I + I, xs.begin() + (size_t)i * I);
            std::copy(Y + (size_t)perm[i] * O, Y + (size_t)perm[i] * O + O, ys.begin() + (size_t)i * O);
         }
         HIPCHECK(hipMemcpy(w.X, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice));
         HIPCHECK(hipMemcpy(w.T, ys.data(), ys.size() * sizeof(float), hipMemcpyHostToDevice));
         for (int b0 = 0; b0 < N; b0 += B) {
            const int nbt = std::min(B, N - b0);
            const float* xb = w.X + (size_t)b0 * I;
            const float* tb = w.T + (size_t)b0 * O;
            forward(w, xb, nbt, nullptr);
            k_mlp_loss<<<1, kThreads>>>(w.Y, tb, nbt, w.loss + e);
            backward(w, xb, tb, nbt, nullptr);
            k_adam<<<grid_for(kMlpParams), kThreads>>>(w.P, w.G, w.M, w.V, lr, ++t);
            ++nb[e];
         }
      }
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipMemcpy(params, w.P, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
      if (loss_hist) {
         HIPCHECK(hipMemcpy(loss_hist, w.loss, epochs * sizeof(float), hipMemcpyDeviceToHost));
         for (int e = 0; e < epochs; ++e) loss_hist[e] /= (float)nb[e];
      }
      return 0;
   } catch (const HipError&) {
      return SYLVER_ERROR_CUDA_UNKNOWN;
   }
}

int mlp_predict(const float* params, const float* X, int N, float* out) {
   if (!params || !X || N < 1 || !out) return SYLVER_ERROR_CALL_SEQUENCE;
   if (device_count() <= 0) return SYLVER_ERROR_CUDA_UNKNOWN;
   try {
      MlpBuffers w(N);
      HIPCHECK(hipMemcpy(w.P, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(w.X, X, (size_t)N * I * sizeof(float), hipMemcpyHostToDevice));
      forward(w, w.X, N, nullptr);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipMemcpy(out, w.Y, (size_t)N * O * sizeof(float), hipMemcpyDeviceToHost));
      return 0;
   } catch (const HipError&) {
      return SYLVER_ERROR_CUDA_UNKNOWN;
   }
}

}  // namespace sylver_amd
