// Small HBM-streaming kernels shared by the components: triangular pack/unpack, casts, fills.
#include "elementwise.hpp"

namespace hh {

namespace {
constexpr int TPB = 256;
unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, TPB), 65536)); }

// Row index of packed-upper entry tp = nu*(nu+1)/2 + mu (mu <= nu).
__device__ __forceinline__ void tp_to_pair(int64_t tp, int64_t& mu, int64_t& nu) {
  int64_t v = (int64_t)((sqrt(8.0 * (double)tp + 1.0) - 1.0) * 0.5);
  while (v * (v + 1) / 2 > tp) --v;
  while ((v + 1) * (v + 2) / 2 <= tp) ++v;
  nu = v;
  mu = tp - v * (v + 1) / 2;
}
}  // namespace

template <typename TI, typename TO>
__global__ void unpack_triu_kernel(const TI* __restrict__ tp, int64_t nao, int64_t ld_tp, int64_t nmat,
                                   TO* __restrict__ out) {
  // one thread per output element (mu, nu, m): coalesced writes, gathered reads
  const int64_t nn = nao * nao;
  const int64_t total = nn * nmat;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = t / nn, e = t % nn;
    const int64_t mu = e % nao, nu = e / nao;
    const int64_t lo = mu < nu ? mu : nu, hi = mu < nu ? nu : mu;
    out[t] = (TO)tp[hi * (hi + 1) / 2 + lo + m * ld_tp];
  }
}

__global__ void pack_triu_tilde_kernel(const double* __restrict__ dm, int64_t nao, int64_t nmat, int symmetrize,
                                       double* __restrict__ out) {
  const int64_t ntp = nao * (nao + 1) / 2;
  const int64_t total = ntp * nmat;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = t / ntp, tp = t % ntp;
    int64_t mu, nu;
    tp_to_pair(tp, mu, nu);
    const double* d = dm + m * nao * nao;
    double v = d[mu + nu * nao];
    if (symmetrize) v += d[nu + mu * nao];
    out[t] = (mu == nu) ? v : 2.0 * v;
  }
}

template <typename TI, typename TO>
__global__ void cast_kernel(const TI* __restrict__ in, int64_t n, TO* __restrict__ out) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    out[t] = (TO)in[t];
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ out, int64_t n, T v) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    out[t] = v;
}

// 16-byte vector copy, grid-stride; every byte is read once and written once.
__global__ void copy16_kernel(const uint4* __restrict__ src, int64_t n16, uint4* __restrict__ dst) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n16; e += (int64_t)gridDim.x * blockDim.x)
    dst[e] = src[e];
}

template <typename TI, typename TO>
void unpack_triu(const TI* tp, int64_t nao, int64_t ld_tp, int64_t nmat, TO* out, hipStream_t s) {
  const int64_t total = nao * nao * nmat;
  if (total == 0) return;
  hipLaunchKernelGGL((unpack_triu_kernel<TI, TO>), dim3(grid_for(total)), dim3(TPB), 0, s, tp, nao, ld_tp, nmat, out);
  HH_HIP(hipGetLastError());
}

void pack_triu_tilde(const double* dm, int64_t nao, int64_t nmat, bool symmetrize, double* out, hipStream_t s) {
  const int64_t total = nao * (nao + 1) / 2 * nmat;
  if (total == 0) return;
  hipLaunchKernelGGL(pack_triu_tilde_kernel, dim3(grid_for(total)), dim3(TPB), 0, s, dm, nao, nmat,
                     symmetrize ? 1 : 0, out);
  HH_HIP(hipGetLastError());
}

template <typename TI, typename TO>
void cast(const TI* in, int64_t n, TO* out, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL((cast_kernel<TI, TO>), dim3(grid_for(n)), dim3(TPB), 0, s, in, n, out);
  HH_HIP(hipGetLastError());
}

template <typename T>
void fill(T* out, int64_t n, T v, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL((fill_kernel<T>), dim3(grid_for(n)), dim3(TPB), 0, s, out, n, v);
  HH_HIP(hipGetLastError());
}

void copy_bytes(void* dst, const void* src, int64_t bytes, hipStream_t s) {
  if (bytes <= 0) return;
  const bool vec = (reinterpret_cast<uintptr_t>(dst) % 16 == 0) && (reinterpret_cast<uintptr_t>(src) % 16 == 0) &&
                   (bytes % 16 == 0);
  if (!vec) {
    HH_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, s));
    return;
  }
  const int64_t n16 = bytes / 16;
  const int64_t nb = std::min<int64_t>((n16 + 255) / 256, 8192);
  hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)nb), dim3(256), 0, s, static_cast<const uint4*>(src), n16,
                     static_cast<uint4*>(dst));
  HH_HIP(hipGetLastError());
}

template void unpack_triu<double, double>(const double*, int64_t, int64_t, int64_t, double*, hipStream_t);
template void unpack_triu<double, float>(const double*, int64_t, int64_t, int64_t, float*, hipStream_t);
template void cast<double, float>(const double*, int64_t, float*, hipStream_t);
template void cast<float, double>(const float*, int64_t, double*, hipStream_t);
template void fill<double>(double*, int64_t, double, hipStream_t);
template void fill<float>(float*, int64_t, float, hipStream_t);
template void fill<unsigned short>(unsigned short*, int64_t, unsigned short, hipStream_t);

}  // namespace hh
