// Learned ACA rank predictor: a small MLP (8 inputs, two hidden layers of 64 with GELU,
// one output) whose forward, backward and Adam kernels are hand-written HIP. Weights are
// used in BF16 by the forward pass with FP32 accumulation; the optimiser keeps an FP32
// master copy. Inputs are standardised with per-feature mean and scale, the output is
// log(1 + rank). No counterpart in the reference (its ACA stops adaptively only).
#pragma once
#include <cstdint>

#include "ctx.hpp"

namespace rcs {

constexpr int MLP_IN = 8, MLP_H = 64;
// parameter layout (floats): W1[H][IN], b1[H], W2[H][H], b2[H], W3[H], b3
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_H * MLP_IN, MLP_W2 = MLP_B1 + MLP_H, MLP_B2 = MLP_W2 + MLP_H * MLP_H,
              MLP_W3 = MLP_B2 + MLP_H, MLP_B3 = MLP_W3 + MLP_H, MLP_NP = MLP_B3 + 1;

struct RankMLP {
  explicit RankMLP(Ctx& c, unsigned seed);
  ~RankMLP();
  RankMLP(const RankMLP&) = delete;
  RankMLP& operator=(const RankMLP&) = delete;

  // host access to the FP32 master parameters and the input standardisation
  void get(float* params, float* mean, float* scale) const;
  void set(const float* params, const float* mean, const float* scale);
  // y[i] = network output (log(1 + rank)) for n samples X (n x 8, row-major), device or host
  void forward_host(const float* X, int n, float* y);
  void forward_device(const float* dX, int n, float* dy);
  // loss = mean (y - t)^2 and its gradient with respect to the FP32 parameters (evaluated
  // with the BF16-rounded weights the forward pass uses)
  double gradient_host(const float* X, const float* t, int n, float* grad);
  // full-batch Adam for `iters` steps; sets the standardisation from X first
  double train_host(const float* X, const float* t, int n, int iters, float lr);
  // predicted ranks (rounded up) for host features
  void predict_ranks(const float* X, int n, int* ranks);

 private:
  double loss_and_grad(const float* dX, const float* dt, int n);  // device, into d_grad
  void sync_bf16();
  Ctx& cx;
  float* d_p = nullptr;      // FP32 master
  uint16_t* d_w16 = nullptr; // BF16 copy used by the forward pass
  float* d_grad = nullptr;
  float* d_m = nullptr;
  float* d_v = nullptr;
  float* d_norm = nullptr;   // mean[8], scale[8]
  float h_norm[2 * MLP_IN];
  long long step_ = 0;
};

}  // namespace rcs
