// Device part of the block Krylov orthogonalisation shared by the single-system solver
// (krylov.hip) and the multi-system batch (cphf_batch.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace hh {

constexpr int KR_GB = 128;           // columns per Gram group
constexpr int KR_GLD = KR_GB + 1;    // LDS leading dimension (odd: row walks hit distinct banks)
constexpr int KR_LDL_T = 1024;       // threads of a block running ldl_drop_block

// LDL^T with the drop rule on a Gram matrix G [m, m] (column-major, lower triangle read),
// m <= KR_GB, run by all KR_LDL_T threads of one block.  Pivot t is kept when its Schur
// complement d_t > lindep (the reference's MGS drop rule, krylov_block.rs:225-253: d_t is the
// squared norm Gram-Schmidt in column order would find for column t); a kept pivot is selected
// when also d_t > sel (sel <= lindep: every kept pivot).  Outputs: Tm [m, m] whose column c
// (c < nsel) is row k_c of L^{-1} for the c-th selected pivot k_c, so V Tm holds the selected
// Gram-Schmidt vectors in order; columns c >= nsel are zero.  dout[c] = d_{k_c} (c < nsel),
// *kout = nsel, and (maxd non-null) *maxd = the largest d_t over the kept pivots (0 if none).
__device__ inline void ldl_drop_block(const double* __restrict__ G, int m, double lindep, double sel,
                                      double* __restrict__ Tm, double* __restrict__ dout, int* __restrict__ kout,
                                      double* __restrict__ maxd) {
  __shared__ double S[KR_GB * KR_GLD];
  __shared__ double dv[KR_GB];
  __shared__ int kidx[KR_GB];
  __shared__ unsigned char kf[KR_GB];
  __shared__ int nsel_s;
  const int tid = threadIdx.x;
  for (int e = tid; e < m * m; e += KR_LDL_T) {
    const int i = e % m, j = e / m;
    if (j <= i) S[i + KR_GLD * j] = G[e];
  }
  __syncthreads();
  for (int t = 0; t < m; ++t) {
    const double d = S[t + KR_GLD * t];  // final: step t touches only rows / columns > t
    const bool keep = d > lindep;
    const int r = m - t - 1;
    if (keep) {
      const double inv = 1.0 / d;
      for (int e = tid; e < r * r; e += KR_LDL_T) {
        const int i = t + 1 + e % r, j = t + 1 + e / r;
        if (j <= i) S[i + KR_GLD * j] -= S[i + KR_GLD * t] * S[j + KR_GLD * t] * inv;
      }
      __syncthreads();
      for (int i = t + 1 + tid; i < m; i += KR_LDL_T) S[i + KR_GLD * t] *= inv;
    } else {
      for (int i = t + 1 + tid; i < m; i += KR_LDL_T) S[i + KR_GLD * t] = 0.0;
    }
    if (tid == 0) {
      dv[t] = d;
      kf[t] = keep ? 1 : 0;
    }
    __syncthreads();
  }
  if (tid == 0) {
    int k = 0;
    double mx = 0.0;
    for (int t = 0; t < m; ++t)
      if (kf[t]) {
        mx = dv[t] > mx ? dv[t] : mx;
        if (dv[t] > sel) kidx[k++] = t;
      }
    nsel_s = k;
    *kout = k;
    if (maxd) *maxd = mx;
  }
  __syncthreads();
  const int nsel = nsel_s;
  // row k of L^{-1}: x = e_k; for q = k .. 1: x[s] -= L[q][s] x[q] (s < q).  One wave per row,
  // x[s] held by lane s % 64 in slot s / 64.
  const int lane = tid & 63, wave = tid >> 6;
  for (int c = wave; c < m; c += KR_LDL_T / 64) {
    double* out = Tm + (int64_t)c * m;
    if (c >= nsel) {
      for (int s2 = lane; s2 < m; s2 += 64) out[s2] = 0.0;
      continue;
    }
    const int k = kidx[c];
    double x0 = (lane == k) ? 1.0 : 0.0;
    double x1 = (lane + 64 == k) ? 1.0 : 0.0;
    for (int q = k; q >= 1; --q) {
      const double xq = __shfl(q < 64 ? x0 : x1, q & 63, 64);
      if (lane < q) x0 -= S[q + KR_GLD * lane] * xq;
      if (lane + 64 < q) x1 -= S[q + KR_GLD * (lane + 64)] * xq;
    }
    if (lane < m) out[lane] = x0;
    if (lane + 64 < m) out[lane + 64] = x1;
    if (lane == 0) dout[c] = dv[k];
  }
}

}  // namespace hh
