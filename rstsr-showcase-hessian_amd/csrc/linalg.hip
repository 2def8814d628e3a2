// Blocked FP64 LU (partial pivoting), Cholesky and triangular solves.
//
// Right-looking blocked algorithms with panel width NB: the panel is factored by one
// workgroup (1024 lanes) in global memory (it stays in L2), the trailing update runs on the
// matrix-core GEMM.  Pivot choice: largest |a|, ties to the smallest row index (deterministic).
#include "linalg.hpp"

namespace hh {

namespace {
constexpr int NB = 64;
constexpr int PT = 1024;  // panel kernel threads

__device__ __forceinline__ void argmax_reduce(double& v, int64_t& idx, double* sv, int64_t* si) {
  // block-wide (PT threads) argmax of |v| (v already absolute), ties -> smaller idx
  const int t = threadIdx.x;
  sv[t] = v;
  si[t] = idx;
  __syncthreads();
  for (int s = PT / 2; s > 0; s >>= 1) {
    if (t < s) {
      const double a = sv[t], b = sv[t + s];
      const int64_t ia = si[t], ib = si[t + s];
      if (b > a || (b == a && ib < ia)) {
        sv[t] = b;
        si[t] = ib;
      }
    }
    __syncthreads();
  }
  v = sv[0];
  idx = si[0];
  __syncthreads();
}

// Unblocked LU of panel columns [k0, k0+nb) over rows [k0, nd); records pivots (global row
// indices) in piv[k0..k0+nb); row swaps applied within the panel only.
__global__ __launch_bounds__(PT) void lu_panel_kernel(double* A, int64_t nd, int64_t k0, int nb, int64_t* piv,
                                                      int* info) {
  __shared__ double sv[PT];
  __shared__ int64_t si[PT];
  __shared__ double prow[NB];
  const int t = threadIdx.x;
  for (int c = 0; c < nb; ++c) {
    const int64_t col = k0 + c;
    double* a = A + col * nd;
    double best = -1.0;
    int64_t bi = nd;
    for (int64_t i = col + t; i < nd; i += PT) {
      const double v = fabs(a[i]);
      if (v > best || (v == best && i < bi)) {
        best = v;
        bi = i;
      }
    }
    argmax_reduce(best, bi, sv, si);
    if (t == 0) {
      piv[col] = bi;
      if (!(best > 0.0) && *info == 0) *info = (int)(col + 1);
    }
    // swap rows col <-> bi inside the panel
    if (bi != col && t < nb) {
      double* x = A + (k0 + t) * nd;
      const double tmp = x[col];
      x[col] = x[bi];
      x[bi] = tmp;
    }
    __syncthreads();
    const double d = a[col];
    const double inv = (d != 0.0) ? 1.0 / d : 0.0;
    for (int64_t i = col + 1 + t; i < nd; i += PT) a[i] *= inv;
    if (t < nb) prow[t] = A[(k0 + t) * nd + col];
    __syncthreads();
    // rank-1 update of the remaining panel columns
    for (int j = c + 1; j < nb; ++j) {
      const double u = prow[j];
      double* x = A + (k0 + j) * nd;
      for (int64_t i = col + 1 + t; i < nd; i += PT) x[i] -= a[i] * u;
    }
    __syncthreads();
  }
}

// Apply the panel's swaps (piv[k0..k0+nb)) to columns outside the panel: trailing H columns
// [k0+nb, nd) and all nrhs columns of G.
__global__ void lu_swap_kernel(double* H, int64_t nd, double* G, int64_t nrhs, int64_t k0, int nb,
                               const int64_t* piv) {
  const int64_t ntrail = nd - (k0 + nb);
  const int64_t ncol = ntrail + nrhs;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < ncol; j += (int64_t)gridDim.x * blockDim.x) {
    double* x = (j < ntrail) ? H + (k0 + nb + j) * nd : G + (j - ntrail) * nd;
    for (int c = 0; c < nb; ++c) {
      const int64_t r = k0 + c, p = piv[r];
      if (p != r) {
        const double tmp = x[r];
        x[r] = x[p];
        x[p] = tmp;
      }
    }
  }
}

// X[k0..k0+nb, j] <- L11^-1 X[..., j] for L11 = unit-lower block of A at (k0, k0) (forward
// substitution), for the trailing H columns and all G columns.
__global__ void lu_trsm_kernel(const double* A, int64_t nd, double* H, double* G, int64_t nrhs, int64_t k0, int nb) {
  __shared__ double L[NB * NB];
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int r = e % nb, c = e / nb;
    L[e] = A[(k0 + r) + (k0 + c) * nd];
  }
  __syncthreads();
  const int64_t ntrail = nd - (k0 + nb);
  const int64_t ncol = ntrail + nrhs;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < ncol; j += (int64_t)gridDim.x * blockDim.x) {
    double* x = ((j < ntrail) ? H + (k0 + nb + j) * nd : G + (j - ntrail) * nd) + k0;
    double y[NB];
    for (int r = 0; r < nb; ++r) y[r] = x[r];
    for (int r = 0; r < nb; ++r) {
      double v = y[r];
      for (int c = 0; c < r; ++c) v -= L[r + c * nb] * y[c];
      y[r] = v;
    }
    for (int r = 0; r < nb; ++r) x[r] = y[r];
  }
}

// Backward substitution for the diagonal block U_pp at (r0, r0) of size nb: X[r0.., j] <- U^-1 X.
__global__ void upper_block_solve_kernel(const double* U, int64_t nd, double* G, int64_t nrhs, int64_t r0, int nb) {
  __shared__ double S[NB * NB];
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int r = e % nb, c = e / nb;
    S[e] = U[(r0 + r) + (r0 + c) * nd];
  }
  __syncthreads();
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nrhs; j += (int64_t)gridDim.x * blockDim.x) {
    double* x = G + j * nd + r0;
    double y[NB];
    for (int r = 0; r < nb; ++r) y[r] = x[r];
    for (int r = nb - 1; r >= 0; --r) {
      double v = y[r];
      for (int c = r + 1; c < nb; ++c) v -= S[r + c * nb] * y[c];
      y[r] = v / S[r + r * nb];
    }
    for (int r = 0; r < nb; ++r) x[r] = y[r];
  }
}

// Cholesky of the diagonal block at (k0, k0) (size nb, lower) by one workgroup, in place.
__global__ __launch_bounds__(256) void chol_block_kernel(double* A, int64_t nd, int64_t k0, int nb, int* info) {
  __shared__ double S[NB * NB];
  const int t = threadIdx.x;
  for (int e = t; e < nb * nb; e += blockDim.x) S[e] = A[(k0 + e % nb) + (k0 + e / nb) * nd];
  __syncthreads();
  for (int c = 0; c < nb; ++c) {
    if (t == 0) {
      const double d = S[c + c * nb];
      if (!(d > 0.0)) {
        if (*info == 0) *info = (int)(k0 + c + 1);
        S[c + c * nb] = 1.0;
      } else {
        S[c + c * nb] = sqrt(d);
      }
    }
    __syncthreads();
    const double dc = S[c + c * nb];
    for (int r = c + 1 + t; r < nb; r += blockDim.x) S[r + c * nb] /= dc;
    __syncthreads();
    for (int e = t; e < nb * nb; e += blockDim.x) {
      const int r = e % nb, j = e / nb;
      if (j > c && r >= j) S[e] -= S[r + c * nb] * S[j + c * nb];
    }
    __syncthreads();
  }
  for (int e = t; e < nb * nb; e += blockDim.x) {
    const int r = e % nb, j = e / nb;
    A[(k0 + r) + (k0 + j) * nd] = (r >= j) ? S[e] : 0.0;
  }
}

// Row-wise forward substitution against a lower block Lb (nb x nb at (k0,k0) of L, ld = nd):
// for each row x of X[:, k0..k0+nb):  x <- x Lb^-T
__global__ void trsm_right_lt_kernel(const double* L, int64_t nd, double* X, int64_t nrow, int64_t ldx, int64_t k0,
                                     int nb) {
  __shared__ double S[NB * NB];
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) S[e] = L[(k0 + e % nb) + (k0 + e / nb) * nd];
  __syncthreads();
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nrow; r += (int64_t)gridDim.x * blockDim.x) {
    double y[NB];
    for (int j = 0; j < nb; ++j) y[j] = X[r + (k0 + j) * ldx];
    for (int j = 0; j < nb; ++j) {
      double v = y[j];
      for (int c = 0; c < j; ++c) v -= y[c] * S[j + c * nb];
      y[j] = v / S[j + j * nb];
    }
    for (int j = 0; j < nb; ++j) X[r + (k0 + j) * ldx] = y[j];
  }
}

// Column-wise forward substitution: X[k0..k0+nb, j] <- Lb^-1 X[k0..k0+nb, j] (non-unit lower)
__global__ void trsm_left_l_kernel(const double* L, int64_t nd, double* X, int64_t ncol, int64_t ldx, int64_t k0,
                                   int nb) {
  __shared__ double S[NB * NB];
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) S[e] = L[(k0 + e % nb) + (k0 + e / nb) * nd];
  __syncthreads();
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < ncol; j += (int64_t)gridDim.x * blockDim.x) {
    double* x = X + j * ldx + k0;
    double y[NB];
    for (int r = 0; r < nb; ++r) y[r] = x[r];
    for (int r = 0; r < nb; ++r) {
      double v = y[r];
      for (int c = 0; c < r; ++c) v -= S[r + c * nb] * y[c];
      y[r] = v / S[r + r * nb];
    }
    for (int r = 0; r < nb; ++r) x[r] = y[r];
  }
}

__global__ void zero_strict_upper_kernel(double* A, int64_t nd) {
  const int64_t total = nd * nd;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x)
    if (e % nd < e / nd) A[e] = 0.0;
}

__global__ void zero_int_kernel(int* p) { *p = 0; }

unsigned blocks_for(int64_t n, int tpb) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, tpb), 4096)); }

int read_info(const int* dinfo, hipStream_t s) {
  int h = 0;
  HH_HIP(hipMemcpyAsync(&h, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  HH_HIP(hipStreamSynchronize(s));
  return h;
}

}  // namespace

void lu_solve(double* H, int64_t nd, double* G, int64_t nrhs, hipStream_t s, DevBuf& ws) {
  if (nd == 0) return;
  DevBuf meta;
  char* m = static_cast<char*>(meta.reserve(sizeof(int64_t) * nd + 256));
  int64_t* piv = reinterpret_cast<int64_t*>(m);
  int* info = reinterpret_cast<int*>(m + sizeof(int64_t) * nd);
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, info);
  for (int64_t k0 = 0; k0 < nd; k0 += NB) {
    const int nb = (int)std::min<int64_t>(NB, nd - k0);
    hipLaunchKernelGGL(lu_panel_kernel, dim3(1), dim3(PT), 0, s, H, nd, k0, nb, piv, info);
    HH_HIP(hipGetLastError());
    const int64_t ntrail = nd - (k0 + nb);
    const int64_t ncol = ntrail + nrhs;
    if (ncol > 0) {
      hipLaunchKernelGGL(lu_swap_kernel, dim3(blocks_for(ncol, 256)), dim3(256), 0, s, H, nd, G, nrhs, k0, nb, piv);
      hipLaunchKernelGGL(lu_trsm_kernel, dim3(blocks_for(ncol, 64)), dim3(64), 0, s, H, nd, H, G, nrhs, k0, nb);
      HH_HIP(hipGetLastError());
    }
    const int64_t mrows = nd - (k0 + nb);
    if (mrows > 0) {
      // A22 -= L21 U12  (H trailing part, then G)
      GemmArgs g;
      g.M = mrows; g.K = nb;
      g.A = H + (k0 + nb) + k0 * nd; g.lda = nd; g.a_kmajor = false;
      g.alpha = -1.0; g.beta = 1.0;
      if (ntrail > 0) {
        g.N = ntrail;
        g.B = H + k0 + (k0 + nb) * nd; g.ldb = nd; g.b_kmajor = true;
        g.C = H + (k0 + nb) + (k0 + nb) * nd; g.ldc = nd;
        gemm<double, double>(g, PREC_F64, s, ws);
      }
      if (nrhs > 0) {
        g.N = nrhs;
        g.B = G + k0; g.ldb = nd; g.b_kmajor = true;
        g.C = G + (k0 + nb); g.ldc = nd;
        gemm<double, double>(g, PREC_F64, s, ws);
      }
    }
  }
  const int bad = read_info(info, s);
  if (bad) throw Error(HH_ERR_INVALID, "lu_solve: singular matrix (zero pivot at column " + std::to_string(bad) + ")");
  // back substitution U X = Y, blocks from the bottom
  const int64_t nblk = ceil_div(nd, NB);
  for (int64_t b = nblk - 1; b >= 0; --b) {
    const int64_t r0 = b * NB;
    const int nb = (int)std::min<int64_t>(NB, nd - r0);
    hipLaunchKernelGGL(upper_block_solve_kernel, dim3(blocks_for(nrhs, 64)), dim3(64), 0, s, H, nd, G, nrhs, r0, nb);
    HH_HIP(hipGetLastError());
    if (r0 > 0) {
      GemmArgs g;  // G[0:r0] -= U[0:r0, r0:r0+nb] X[r0:r0+nb]
      g.M = r0; g.N = nrhs; g.K = nb;
      g.A = H + r0 * nd; g.lda = nd; g.a_kmajor = false;
      g.B = G + r0; g.ldb = nd; g.b_kmajor = true;
      g.C = G; g.ldc = nd;
      g.alpha = -1.0; g.beta = 1.0;
      gemm<double, double>(g, PREC_F64, s, ws);
    }
  }
}

bool cholesky(double* H, int64_t nd, hipStream_t s, DevBuf& ws) {
  if (nd == 0) return true;
  DevBuf meta;
  int* info = static_cast<int*>(meta.reserve(256));
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, info);
  for (int64_t k0 = 0; k0 < nd; k0 += NB) {
    const int nb = (int)std::min<int64_t>(NB, nd - k0);
    hipLaunchKernelGGL(chol_block_kernel, dim3(1), dim3(256), 0, s, H, nd, k0, nb, info);
    HH_HIP(hipGetLastError());
    const int64_t m = nd - (k0 + nb);
    if (m > 0) {
      // L21 = A21 L11^-T
      hipLaunchKernelGGL(trsm_right_lt_kernel, dim3(blocks_for(m, 64)), dim3(64), 0, s, H, nd, H + (k0 + nb), m, nd,
                         k0, nb);
      HH_HIP(hipGetLastError());
      // A22 -= L21 L21^T  (full trailing block; the strict upper part is cleared at the end)
      GemmArgs g;
      g.M = m; g.N = m; g.K = nb;
      g.A = H + (k0 + nb) + k0 * nd; g.lda = nd; g.a_kmajor = false;
      g.B = H + (k0 + nb) + k0 * nd; g.ldb = nd; g.b_kmajor = false;  // B[k, n] = L21[n, k]
      g.C = H + (k0 + nb) + (k0 + nb) * nd; g.ldc = nd;
      g.alpha = -1.0; g.beta = 1.0;
      gemm<double, double>(g, PREC_F64, s, ws);
    }
  }
  hipLaunchKernelGGL(zero_strict_upper_kernel, dim3(blocks_for(nd * nd, 256)), dim3(256), 0, s, H, nd);
  HH_HIP(hipGetLastError());
  const int bad = read_info(info, s);
  return bad == 0;
}

void trsm_lower(const double* L, int64_t nd, double* X, int64_t n2, int64_t ldx, bool side_right, hipStream_t s,
                DevBuf& ws) {
  if (nd == 0 || n2 == 0) return;
  for (int64_t k0 = 0; k0 < nd; k0 += NB) {
    const int nb = (int)std::min<int64_t>(NB, nd - k0);
    const int64_t rest = nd - (k0 + nb);
    if (side_right) {
      // X[:, p] <- X[:, p] Lpp^-T ; X[:, after] -= X[:, p] L[after, p]^T
      hipLaunchKernelGGL(trsm_right_lt_kernel, dim3(blocks_for(n2, 64)), dim3(64), 0, s, L, nd, X, n2, ldx, k0, nb);
      HH_HIP(hipGetLastError());
      if (rest > 0) {
        GemmArgs g;
        g.M = n2; g.N = rest; g.K = nb;
        g.A = X + k0 * ldx; g.lda = ldx; g.a_kmajor = false;
        g.B = L + (k0 + nb) + k0 * nd; g.ldb = nd; g.b_kmajor = false;  // B[c, j] = L[k0+nb+j, k0+c]
        g.C = X + (k0 + nb) * ldx; g.ldc = ldx;
        g.alpha = -1.0; g.beta = 1.0;
        gemm<double, double>(g, PREC_F64, s, ws);
      }
    } else {
      hipLaunchKernelGGL(trsm_left_l_kernel, dim3(blocks_for(n2, 64)), dim3(64), 0, s, L, nd, X, n2, ldx, k0, nb);
      HH_HIP(hipGetLastError());
      if (rest > 0) {
        GemmArgs g;  // X[after, :] -= L[after, p] X[p, :]
        g.M = rest; g.N = n2; g.K = nb;
        g.A = L + (k0 + nb) + k0 * nd; g.lda = nd; g.a_kmajor = false;
        g.B = X + k0; g.ldb = ldx; g.b_kmajor = true;
        g.C = X + (k0 + nb); g.ldc = ldx;
        g.alpha = -1.0; g.beta = 1.0;
        gemm<double, double>(g, PREC_F64, s, ws);
      }
    }
  }
}

}  // namespace hh
