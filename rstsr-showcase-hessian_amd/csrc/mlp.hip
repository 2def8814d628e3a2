// Learned diagonal preconditioner MLP: forward / backward / Adam kernels (see mlp.hpp).
//
// Samples are independent, so forward and backward run one lane per sample with the whole
// network (8641 FP32 parameters, 34.6 KB) staged in LDS and read by broadcast.  Weight and bias
// gradients are sums over samples, computed as FP32 matrix-core GEMMs (activations^T deltas)
// with the library's GEMM, so the reduction order is fixed and results are reproducible.
#include "mlp.hpp"

#include <cmath>

#include "elementwise.hpp"

namespace hh {

namespace {
constexpr int NIN = DiagMlp::NIN, H = DiagMlp::H, NP = DiagMlp::NPARAM;
constexpr int OW1 = 0, OB1 = OW1 + NIN * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H * H,
              OW4 = OB3 + H, OB4 = OW4 + H;
static_assert(OB4 + 1 == NP, "parameter layout");

unsigned nblk(int64_t n, int t = 256) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, t), 65535)); }

__device__ __forceinline__ void stage_params(const float* p, float* s) {
  for (int e = threadIdx.x; e < NP; e += blockDim.x) s[e] = p[e];
  __syncthreads();
}

// act (optional): a1, a2, a3 [N, H] post-ReLU activations, column-major (n + N*j)
__global__ __launch_bounds__(256) void mlp_forward_kernel(const float* __restrict__ prm, const float* __restrict__ x,
                                                          int64_t n, float* __restrict__ y, float* __restrict__ act) {
  __shared__ float s[NP];
  stage_params(prm, s);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float in[NIN];
#pragma unroll
    for (int f = 0; f < NIN; ++f) in[f] = x[i + n * f];
    float h1[H], h2[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
      float v = s[OB1 + j];
#pragma unroll
      for (int f = 0; f < NIN; ++f) v += s[OW1 + j + H * f] * in[f];
      h1[j] = fmaxf(v, 0.f);
    }
    if (act)
#pragma unroll
      for (int j = 0; j < H; ++j) act[i + n * j] = h1[j];
#pragma unroll
    for (int k = 0; k < H; ++k) {
      float v = s[OB2 + k];
#pragma unroll
      for (int j = 0; j < H; ++j) v += s[OW2 + k + H * j] * h1[j];
      h2[k] = fmaxf(v, 0.f);
    }
    if (act)
#pragma unroll
      for (int j = 0; j < H; ++j) act[i + n * (H + j)] = h2[j];
#pragma unroll
    for (int k = 0; k < H; ++k) {
      float v = s[OB3 + k];
#pragma unroll
      for (int j = 0; j < H; ++j) v += s[OW3 + k + H * j] * h2[j];
      h1[k] = fmaxf(v, 0.f);
    }
    if (act)
#pragma unroll
      for (int j = 0; j < H; ++j) act[i + n * (2 * H + j)] = h1[j];
    float out = s[OB4];
#pragma unroll
    for (int k = 0; k < H; ++k) out += s[OW4 + k] * h1[k];
    y[i] = out;
  }
}

// per-sample deltas: dy = 2 w (y - t) / sum_w; d3 = dy W4 . [a3>0]; d2 = W3^T d3 . [a2>0];
// d1 = W2^T d2 . [a1>0].  del: d1, d2, d3 [N, H] and dy [N] (column-major).
__global__ __launch_bounds__(256) void mlp_backward_kernel(const float* __restrict__ prm, const float* __restrict__ y,
                                                           const float* __restrict__ t, const float* __restrict__ w,
                                                           const float* __restrict__ act, int64_t n, float inv_sw,
                                                           float* __restrict__ del) {
  __shared__ float s[NP];
  stage_params(prm, s);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float dy = 2.f * w[i] * (y[i] - t[i]) * inv_sw;
    float d3[H], d2[H];
#pragma unroll
    for (int k = 0; k < H; ++k) d3[k] = act[i + n * (2 * H + k)] > 0.f ? dy * s[OW4 + k] : 0.f;
#pragma unroll
    for (int j = 0; j < H; ++j) {
      float v = 0.f;
#pragma unroll
      for (int k = 0; k < H; ++k) v += s[OW3 + k + H * j] * d3[k];
      d2[j] = act[i + n * (H + j)] > 0.f ? v : 0.f;
    }
#pragma unroll
    for (int j = 0; j < H; ++j) {
      del[i + n * (H + j)] = d2[j];
      del[i + n * (2 * H + j)] = d3[j];
    }
#pragma unroll
    for (int q = 0; q < H; ++q) {
      float v = 0.f;
#pragma unroll
      for (int j = 0; j < H; ++j) v += s[OW2 + j + H * q] * d2[j];
      del[i + n * q] = act[i + n * q] > 0.f ? v : 0.f;
    }
    del[i + n * (3 * H)] = dy;
  }
}

// sum_n w_n (y_n - t_n)^2 and sum_n w_n in FP64, one workgroup, fixed order
__global__ __launch_bounds__(256) void weighted_sse_kernel(const float* y, const float* t, const float* w, int64_t n,
                                                           double* out) {
  __shared__ double a[256], b[256];
  double sa = 0.0, sb = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double d = (double)y[i] - (double)t[i];
    sa += (double)w[i] * d * d;
    sb += (double)w[i];
  }
  a[threadIdx.x] = sa;
  b[threadIdx.x] = sb;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      a[threadIdx.x] += a[threadIdx.x + st];
      b[threadIdx.x] += b[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = a[0];
    out[1] = b[0];
  }
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v, int np, float lr, float b1c, float b2c) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / b1c) / (sqrtf(vi / b2c) + 1e-8f);
  }
}

uint64_t splitmix(uint64_t& st) {
  st += 0x9E3779B97F4A7C15ULL;
  uint64_t z = st;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
}  // namespace

DiagMlp::DiagMlp(uint64_t seed, hipStream_t s) : stream_(s) {
  // He-uniform hidden weights, zero biases, small output layer: the untrained network predicts
  // ~0, i.e. a preconditioner ~exp(0) = 1.
  std::vector<float> h(NP, 0.f);
  uint64_t st = seed;
  auto uni = [&](float bound) { return (float)(((splitmix(st) >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0) * bound; };
  for (int e = 0; e < NIN * H; ++e) h[OW1 + e] = uni(std::sqrt(6.f / NIN));
  for (int e = 0; e < H * H; ++e) h[OW2 + e] = uni(std::sqrt(6.f / H));
  for (int e = 0; e < H * H; ++e) h[OW3 + e] = uni(std::sqrt(6.f / H));
  for (int e = 0; e < H; ++e) h[OW4 + e] = 0.1f * uni(std::sqrt(6.f / H));
  p_.reserve(sizeof(float) * NP);
  g_.reserve(sizeof(float) * NP);
  m_.reserve(sizeof(float) * NP);
  v_.reserve(sizeof(float) * NP);
  HH_HIP(hipMemcpyAsync(p_.ptr, h.data(), sizeof(float) * NP, hipMemcpyHostToDevice, s));
  fill<float>(m_.as<float>(), NP, 0.f, s);
  fill<float>(v_.as<float>(), NP, 0.f, s);
  HH_HIP(hipStreamSynchronize(s));
}

void DiagMlp::get_params(float* out) const {
  HH_HIP(hipMemcpyAsync(out, p_.ptr, sizeof(float) * NP, hipMemcpyDeviceToHost, stream_));
  HH_HIP(hipStreamSynchronize(stream_));
}

void DiagMlp::set_params(const float* in) {
  HH_HIP(hipMemcpyAsync(p_.ptr, in, sizeof(float) * NP, hipMemcpyHostToDevice, stream_));
  HH_HIP(hipStreamSynchronize(stream_));
  fill<float>(m_.as<float>(), NP, 0.f, stream_);
  fill<float>(v_.as<float>(), NP, 0.f, stream_);
  step_ = 0;
}

void DiagMlp::forward(const float* x, int64_t n, float* y) {
  if (n == 0) return;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(nblk(n)), dim3(256), 0, stream_, p_.as<float>(), x, n, y, nullptr);
  HH_HIP(hipGetLastError());
}

double DiagMlp::loss_and_grad(const float* x, const float* t, const float* w, int64_t n, float* grad) {
  HH_REQUIRE(n > 0, "mlp: no samples");
  float* act = static_cast<float*>(act_.reserve(sizeof(float) * n * (3 * H + 1 + 1)));
  float* y = act + 3 * H * n;
  float* ones = y + n;
  float* del = static_cast<float*>(del_.reserve(sizeof(float) * n * (3 * H + 1)));
  double* red = static_cast<double*>(red_.reserve(sizeof(double) * 2));
  fill<float>(ones, n, 1.f, stream_);
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(nblk(n)), dim3(256), 0, stream_, p_.as<float>(), x, n, y, act);
  hipLaunchKernelGGL(weighted_sse_kernel, dim3(1), dim3(256), 0, stream_, y, t, w, n, red);
  HH_HIP(hipGetLastError());
  double hr[2];
  HH_HIP(hipMemcpyAsync(hr, red, sizeof(hr), hipMemcpyDeviceToHost, stream_));
  HH_HIP(hipStreamSynchronize(stream_));
  HH_REQUIRE(hr[1] > 0.0, "mlp: sample weights sum to zero");
  const float inv_sw = (float)(1.0 / hr[1]);
  hipLaunchKernelGGL(mlp_backward_kernel, dim3(nblk(n)), dim3(256), 0, stream_, p_.as<float>(), y, t, w, act, n,
                     inv_sw, del);
  HH_HIP(hipGetLastError());
  const float* d1 = del;
  const float* d2 = del + n * H;
  const float* d3 = del + 2 * n * H;
  const float* dy = del + 3 * n * H;
  const float* a1 = act;
  const float* a2 = act + n * H;
  const float* a3 = act + 2 * n * H;
  // grad[W] = delta^T activation (FP32 matrix cores), grad[b] = delta^T 1
  auto gT = [&](const float* dlt, int64_t rows, const float* act_in, int64_t cols, float* out) {
    GemmArgs g;
    g.M = rows; g.N = cols; g.K = n;
    g.A = dlt; g.lda = n; g.a_kmajor = true;
    g.B = act_in; g.ldb = n; g.b_kmajor = true;
    g.C = out; g.ldc = rows;
    gemm<float, float>(g, PREC_F32, stream_, ws_);
  };
  gT(d1, H, x, NIN, grad + OW1);
  gT(d1, H, ones, 1, grad + OB1);
  gT(d2, H, a1, H, grad + OW2);
  gT(d2, H, ones, 1, grad + OB2);
  gT(d3, H, a2, H, grad + OW3);
  gT(d3, H, ones, 1, grad + OB3);
  gT(dy, 1, a3, H, grad + OW4);
  gT(dy, 1, ones, 1, grad + OB4);
  return hr[0] / hr[1];
}

double DiagMlp::train(const float* x, const float* t, const float* w, int64_t n, int epochs, float lr) {
  double loss = 0.0;
  for (int e = 0; e < epochs; ++e) {
    loss = loss_and_grad(x, t, w, n, g_.as<float>());
    ++step_;
    const float b1c = 1.f - std::pow(0.9f, (float)step_), b2c = 1.f - std::pow(0.999f, (float)step_);
    hipLaunchKernelGGL(adam_kernel, dim3(nblk(NP)), dim3(256), 0, stream_, p_.as<float>(), g_.as<float>(),
                       m_.as<float>(), v_.as<float>(), NP, lr, b1c, b2c);
    HH_HIP(hipGetLastError());
  }
  HH_HIP(hipStreamSynchronize(stream_));
  return loss;
}

}  // namespace hh
