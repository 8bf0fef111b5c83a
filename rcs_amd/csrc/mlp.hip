// Learned ACA rank predictor (see mlp.hpp): forward, backward, gradient reduction and
// Adam kernels. One thread carries one sample through the network; the weights sit in LDS
// as floats converted from their BF16 copy, so every multiply-add runs in FP32 on BF16
// weight values. Gradients are reduced over the batch by one thread per parameter
// (deterministic sums, no atomics).
#include "mlp.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace rcs {

namespace {

__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ inline uint16_t f_to_bf16(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z) {  // Phi(z) + z phi(z)
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__global__ void mlp_to_bf16_kernel(const float* p, uint16_t* w, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = f_to_bf16(p[i]);
}

// Stage the BF16 weights in LDS as floats (all threads of the block).
__device__ inline void load_weights(const uint16_t* w, float* s) {
  for (int i = threadIdx.x; i < MLP_NP; i += blockDim.x) s[i] = bf16_to_f(w[i]);
  __syncthreads();
}

__device__ inline void standardise(const float* X, const float* norm, int smp, float* xn) {
#pragma unroll
  for (int k = 0; k < MLP_IN; ++k) xn[k] = (X[(size_t)smp * MLP_IN + k] - norm[k]) / norm[MLP_IN + k];
}

__global__ void __launch_bounds__(256) mlp_forward_kernel(const uint16_t* w, const float* norm, const float* X, int n,
                                                          float* y) {
  __shared__ float s[MLP_NP];
  load_weights(w, s);
  const int smp = blockIdx.x * blockDim.x + threadIdx.x;
  if (smp >= n) return;
  float xn[MLP_IN], h1[MLP_H];
  standardise(X, norm, smp, xn);
#pragma unroll
  for (int i = 0; i < MLP_H; ++i) {
    float z = s[MLP_B1 + i];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) z += s[MLP_W1 + i * MLP_IN + k] * xn[k];
    h1[i] = gelu(z);
  }
  float out = s[MLP_B3];
  for (int i = 0; i < MLP_H; ++i) {
    float z = s[MLP_B2 + i];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) z += s[MLP_W2 + i * MLP_H + j] * h1[j];
    out += s[MLP_W3 + i] * gelu(z);
  }
  y[smp] = out;
}

// Per-sample backward pass of loss = mean (y - t)^2: stores the layer inputs and the
// pre-activation gradients that the parameter gradients are built from.
__global__ void __launch_bounds__(256) mlp_backward_kernel(const uint16_t* w, const float* norm, const float* X,
                                                           const float* t, int n, float* H1, float* H2, float* DZ1,
                                                           float* DZ2, float* DY, float* E2) {
  __shared__ float s[MLP_NP];
  load_weights(w, s);
  const int smp = blockIdx.x * blockDim.x + threadIdx.x;
  if (smp >= n) return;
  // register arrays are only indexed by unrolled loop counters (no scratch); the second
  // layer is evaluated twice (once for the output, once for its gradient)
  float xn[MLP_IN], h1[MLP_H], dh1[MLP_H];
  standardise(X, norm, smp, xn);
  float* h1o = H1 + (size_t)smp * MLP_H;
  float* h2o = H2 + (size_t)smp * MLP_H;
#pragma unroll
  for (int i = 0; i < MLP_H; ++i) {
    float z = s[MLP_B1 + i];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) z += s[MLP_W1 + i * MLP_IN + k] * xn[k];
    h1[i] = gelu(z);
    h1o[i] = h1[i];
    dh1[i] = 0.0f;
  }
  float out = s[MLP_B3];
  for (int i = 0; i < MLP_H; ++i) {
    float z = s[MLP_B2 + i];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) z += s[MLP_W2 + i * MLP_H + j] * h1[j];
    const float h2 = gelu(z);
    h2o[i] = h2;
    out += s[MLP_W3 + i] * h2;
  }
  const float err = out - t[smp];
  const float dy = 2.0f * err / (float)n;
  DY[smp] = dy;
  E2[smp] = err * err;
  float* dz2o = DZ2 + (size_t)smp * MLP_H;
  for (int i = 0; i < MLP_H; ++i) {
    float z = s[MLP_B2 + i];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) z += s[MLP_W2 + i * MLP_H + j] * h1[j];
    const float dz2 = dy * s[MLP_W3 + i] * gelu_d(z);
    dz2o[i] = dz2;
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) dh1[j] += s[MLP_W2 + i * MLP_H + j] * dz2;
  }
  float* dz1o = DZ1 + (size_t)smp * MLP_H;
#pragma unroll
  for (int j = 0; j < MLP_H; ++j) {
    float z = s[MLP_B1 + j];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) z += s[MLP_W1 + j * MLP_IN + k] * xn[k];
    dz1o[j] = dh1[j] * gelu_d(z);
  }
}

// One thread per parameter (plus one for the loss): sums the batch contributions.
__global__ void __launch_bounds__(256) mlp_reduce_kernel(const float* norm, const float* X, const float* H1,
                                                         const float* H2, const float* DZ1, const float* DZ2,
                                                         const float* DY, const float* E2, int n, float* grad,
                                                         double* loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > MLP_NP) return;
  if (p == MLP_NP) {
    double a = 0;
    for (int s = 0; s < n; ++s) a += E2[s];
    *loss = a / n;
    return;
  }
  float acc = 0;
  if (p < MLP_B1) {
    const int i = p / MLP_IN, k = p % MLP_IN;
    for (int s = 0; s < n; ++s) acc += DZ1[(size_t)s * MLP_H + i] * ((X[(size_t)s * MLP_IN + k] - norm[k]) / norm[MLP_IN + k]);
  } else if (p < MLP_W2) {
    const int i = p - MLP_B1;
    for (int s = 0; s < n; ++s) acc += DZ1[(size_t)s * MLP_H + i];
  } else if (p < MLP_B2) {
    const int i = (p - MLP_W2) / MLP_H, j = (p - MLP_W2) % MLP_H;
    for (int s = 0; s < n; ++s) acc += DZ2[(size_t)s * MLP_H + i] * H1[(size_t)s * MLP_H + j];
  } else if (p < MLP_W3) {
    const int i = p - MLP_B2;
    for (int s = 0; s < n; ++s) acc += DZ2[(size_t)s * MLP_H + i];
  } else if (p < MLP_B3) {
    const int i = p - MLP_W3;
    for (int s = 0; s < n; ++s) acc += DY[s] * H2[(size_t)s * MLP_H + i];
  } else {
    for (int s = 0; s < n; ++s) acc += DY[s];
  }
  grad[p] = acc;
}

__global__ void mlp_adam_kernel(float* p, const float* g, float* m, float* v, int np, float lr, float b1, float b2,
                                float eps, float bc1, float bc2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  m[i] = b1 * m[i] + (1 - b1) * gi;
  v[i] = b2 * v[i] + (1 - b2) * gi * gi;
  p[i] -= lr * (m[i] / bc1) / (sqrtf(v[i] / bc2) + eps);
}

}  // namespace

RankMLP::RankMLP(Ctx& c, unsigned seed) : cx(c) {
  HIPCHK(hipMalloc((void**)&d_p, sizeof(float) * MLP_NP));
  HIPCHK(hipMalloc((void**)&d_w16, sizeof(uint16_t) * MLP_NP));
  HIPCHK(hipMalloc((void**)&d_grad, sizeof(float) * MLP_NP));
  HIPCHK(hipMalloc((void**)&d_m, sizeof(float) * MLP_NP));
  HIPCHK(hipMalloc((void**)&d_v, sizeof(float) * MLP_NP));
  HIPCHK(hipMalloc((void**)&d_norm, sizeof(float) * 2 * MLP_IN));
  // Glorot-uniform weights from a 64-bit LCG, zero biases
  std::vector<float> p(MLP_NP, 0.0f);
  unsigned long long s = 0x9E3779B97F4A7C15ull ^ seed;
  auto uni = [&]() {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (float)((s >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f;
  };
  const float a1 = std::sqrt(6.0f / (MLP_IN + MLP_H)), a2 = std::sqrt(6.0f / (2 * MLP_H)), a3 = std::sqrt(6.0f / (MLP_H + 1));
  for (int i = 0; i < MLP_H * MLP_IN; ++i) p[MLP_W1 + i] = a1 * uni();
  for (int i = 0; i < MLP_H * MLP_H; ++i) p[MLP_W2 + i] = a2 * uni();
  for (int i = 0; i < MLP_H; ++i) p[MLP_W3 + i] = a3 * uni();
  float mean[MLP_IN], scale[MLP_IN];
  for (int k = 0; k < MLP_IN; ++k) { mean[k] = 0; scale[k] = 1; }
  set(p.data(), mean, scale);
}

RankMLP::~RankMLP() {
  for (void* q : {(void*)d_p, (void*)d_w16, (void*)d_grad, (void*)d_m, (void*)d_v, (void*)d_norm})
    if (q) (void)hipFree(q);
}

void RankMLP::sync_bf16() {
  hipLaunchKernelGGL(mlp_to_bf16_kernel, dim3((MLP_NP + 255) / 256), dim3(256), 0, cx.st, (const float*)d_p, d_w16,
                     MLP_NP);
}

void RankMLP::set(const float* params, const float* mean, const float* scale) {
  cx.upload(d_p, params, sizeof(float) * MLP_NP);
  std::memcpy(h_norm, mean, sizeof(float) * MLP_IN);
  std::memcpy(h_norm + MLP_IN, scale, sizeof(float) * MLP_IN);
  cx.upload(d_norm, h_norm, sizeof(float) * 2 * MLP_IN);
  std::vector<float> z(MLP_NP, 0.0f);
  cx.upload(d_m, z.data(), sizeof(float) * MLP_NP);
  cx.upload(d_v, z.data(), sizeof(float) * MLP_NP);
  step_ = 0;
  sync_bf16();
  cx.sync();
}

void RankMLP::get(float* params, float* mean, float* scale) const {
  HIPCHK(hipMemcpyAsync(params, d_p, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost, cx.st));
  HIPCHK(hipStreamSynchronize(cx.st));
  std::memcpy(mean, h_norm, sizeof(float) * MLP_IN);
  std::memcpy(scale, h_norm + MLP_IN, sizeof(float) * MLP_IN);
}

void RankMLP::forward_device(const float* dX, int n, float* dy) {
  if (n <= 0) return;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((n + 255) / 256), dim3(256), 0, cx.st, (const uint16_t*)d_w16,
                     (const float*)d_norm, dX, n, dy);
  HIPCHK(hipGetLastError());
}

void RankMLP::forward_host(const float* X, int n, float* y) {
  if (n <= 0) return;
  float* dX = (float*)cx.pool.alloc(sizeof(float) * (size_t)n * MLP_IN);
  float* dy = (float*)cx.pool.alloc(sizeof(float) * (size_t)n);
  cx.upload(dX, X, sizeof(float) * (size_t)n * MLP_IN);
  forward_device(dX, n, dy);
  HIPCHK(hipMemcpyAsync(y, dy, sizeof(float) * n, hipMemcpyDeviceToHost, cx.st));
  cx.sync();
  cx.pool.free(dX);
  cx.pool.free(dy);
}

double RankMLP::loss_and_grad(const float* dX, const float* dt, int n) {
  float* w = (float*)cx.pool.alloc(sizeof(float) * ((size_t)n * (4 * MLP_H + 2)));
  float *H1 = w, *H2 = H1 + (size_t)n * MLP_H, *DZ1 = H2 + (size_t)n * MLP_H, *DZ2 = DZ1 + (size_t)n * MLP_H;
  float *DY = DZ2 + (size_t)n * MLP_H, *E2 = DY + n;
  hipLaunchKernelGGL(mlp_backward_kernel, dim3((n + 255) / 256), dim3(256), 0, cx.st, (const uint16_t*)d_w16,
                     (const float*)d_norm, dX, dt, n, H1, H2, DZ1, DZ2, DY, E2);
  double* dloss = cx.d_scalars;
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((MLP_NP + 1 + 255) / 256), dim3(256), 0, cx.st, (const float*)d_norm, dX,
                     (const float*)H1, (const float*)H2, (const float*)DZ1, (const float*)DZ2, (const float*)DY,
                     (const float*)E2, n, d_grad, dloss);
  HIPCHK(hipGetLastError());
  cx.pool.free(w);
  return 0.0;
}

double RankMLP::gradient_host(const float* X, const float* t, int n, float* grad) {
  float* dX = (float*)cx.pool.alloc(sizeof(float) * (size_t)n * MLP_IN);
  float* dt = (float*)cx.pool.alloc(sizeof(float) * (size_t)n);
  cx.upload(dX, X, sizeof(float) * (size_t)n * MLP_IN);
  cx.upload(dt, t, sizeof(float) * (size_t)n);
  loss_and_grad(dX, dt, n);
  HIPCHK(hipMemcpyAsync(grad, d_grad, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost, cx.st));
  const double loss = cx.fetch_scalars(1)[0];
  cx.pool.free(dX);
  cx.pool.free(dt);
  return loss;
}

double RankMLP::train_host(const float* X, const float* t, int n, int iters, float lr) {
  if (n <= 0) return 0.0;
  // standardisation from the training inputs
  float mean[MLP_IN], scale[MLP_IN];
  for (int k = 0; k < MLP_IN; ++k) {
    double a = 0, b = 0;
    for (int i = 0; i < n; ++i) a += X[(size_t)i * MLP_IN + k];
    a /= n;
    for (int i = 0; i < n; ++i) b += (X[(size_t)i * MLP_IN + k] - a) * (X[(size_t)i * MLP_IN + k] - a);
    mean[k] = (float)a;
    scale[k] = (float)std::sqrt(b / n);
    if (!(scale[k] > 1e-6f)) scale[k] = 1.0f;
  }
  std::vector<float> p(MLP_NP);
  float m0[MLP_IN], s0[MLP_IN];
  get(p.data(), m0, s0);
  set(p.data(), mean, scale);
  float* dX = (float*)cx.pool.alloc(sizeof(float) * (size_t)n * MLP_IN);
  float* dt = (float*)cx.pool.alloc(sizeof(float) * (size_t)n);
  cx.upload(dX, X, sizeof(float) * (size_t)n * MLP_IN);
  cx.upload(dt, t, sizeof(float) * (size_t)n);
  const float b1 = 0.9f, b2 = 0.999f;
  for (int it = 0; it < iters; ++it) {
    loss_and_grad(dX, dt, n);
    ++step_;
    const float bc1 = 1.0f - std::pow(b1, (float)step_), bc2 = 1.0f - std::pow(b2, (float)step_);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_NP + 255) / 256), dim3(256), 0, cx.st, d_p, (const float*)d_grad, d_m,
                       d_v, MLP_NP, lr, b1, b2, 1e-8f, bc1, bc2);
    sync_bf16();
  }
  loss_and_grad(dX, dt, n);
  const double loss = cx.fetch_scalars(1)[0];
  cx.pool.free(dX);
  cx.pool.free(dt);
  return loss;
}

void RankMLP::predict_ranks(const float* X, int n, int* ranks) {
  std::vector<float> y(n);
  forward_host(X, n, y.data());
  for (int i = 0; i < n; ++i) {  // the output is log(1 + rank); far outside the training data it can be any value
    const double r = std::expm1(std::min(std::max((double)y[i], 0.0), 20.0));
    ranks[i] = r > 0 ? (int)std::min(std::ceil(r), 1e8) : 1;
  }
}

}  // namespace rcs
