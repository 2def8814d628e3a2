// Learned diagonal CPHF preconditioner: a small MLP (3 -> 64 -> 64 -> 64 -> 1, ReLU, FP32)
// with hand-written forward, backward and Adam kernels.  No reference counterpart: the
// reference preconditions CPHF only by the orbital-energy scaling of its dimensionless form
// (rscf.rs:237-263); this network predicts a per-(a,i) correction of that diagonal.
#pragma once
#include "gemm.hpp"

namespace hh {

class DiagMlp {
 public:
  static constexpr int NIN = 3, H = 64, NPARAM = NIN * H + H + H * H + H + H * H + H + H + 1;
  explicit DiagMlp(uint64_t seed, hipStream_t s);

  // y[n] = f(x[n, :]); x [N, NIN] column-major (feature fastest over samples: x[n + N*f])
  void forward(const float* x, int64_t n, float* y);
  // Weighted MSE loss L = sum_n w_n (f(x_n) - t_n)^2 / sum_n w_n; gradient into `grad` (NPARAM,
  // device) and the loss (host).  Deterministic (fixed-order reductions).
  double loss_and_grad(const float* x, const float* t, const float* w, int64_t n, float* grad);
  // `epochs` full-batch Adam steps (beta 0.9 / 0.999, eps 1e-8); returns the final loss.
  double train(const float* x, const float* t, const float* w, int64_t n, int epochs, float lr);

  void get_params(float* host_out) const;  // NPARAM floats
  void set_params(const float* host_in);
  hipStream_t stream() const { return stream_; }

 private:
  hipStream_t stream_;
  DevBuf p_, g_, m_, v_, act_, del_, ws_, red_;
  int step_ = 0;
};

// Parameter layout (floats): W1 [H, NIN] column-major (W1[j + H*f]), b1 [H], W2 [H, H]
// (W2[k + H*j]: out k, in j), b2 [H], W3 [H, H], b3 [H], W4 [H] (out 1), b4 [1].

}  // namespace hh
