// Blocked FP64 LU (partial pivoting), Cholesky and triangular solves.
//
// Right-looking blocked algorithms with panel width NB: the panel is factored by one
// workgroup (1024 lanes) in global memory (it stays in L2), the trailing update runs on the
// matrix-core GEMM.  Pivot choice: largest |a|, ties to the smallest row index (deterministic).
#include "linalg.hpp"

#include <mutex>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

namespace hh {

namespace {
constexpr int NB = 64;
constexpr int PT = 1024;  // panel kernel threads

// Unblocked LU of panel columns [k0, k0+nb) over rows [k0, nd); records pivots (global row
// indices) in piv[k0..k0+nb); row swaps applied within the panel only.  Per column: the argmax
// (largest |a|, ties to the smaller row) by wave shuffles and one exchange of the 16 wave results
// in LDS; the rank-1 update of the panel's remaining columns row by row, eight columns' loads
// issued together.
__global__ __launch_bounds__(PT) void lu_panel_kernel(double* A, int64_t nd, int64_t k0, int nb, int64_t* piv,
                                                      int* info) {
  constexpr int NW = PT / 64;
  __shared__ double wbest[2][NW];
  __shared__ int64_t widx[2][NW];
  __shared__ double prow[2][NB];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int c = 0; c < nb; ++c) {
    const int par = c & 1;  // double-buffered LDS: no barrier between the last reads and the next writes
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
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int64_t oi = __shfl_xor(bi, o, 64);
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    if (lane == 0) {
      wbest[par][w] = best;
      widx[par][w] = bi;
    }
    __syncthreads();
    best = wbest[par][0];
    bi = widx[par][0];
    for (int x = 1; x < NW; ++x) {
      const double ob = wbest[par][x];
      const int64_t oi = widx[par][x];
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    if (t == 0) {
      piv[col] = bi;
      if (!(best > 0.0) && *info == 0) *info = (int)(col + 1);
    }
    // swap rows col <-> bi inside the panel; thread t also keeps the pivot row's entry of its
    // column for the update
    if (t < nb) {
      double* x = A + (k0 + t) * nd;
      if (bi != col) {
        const double tmp = x[col];
        x[col] = x[bi];
        x[bi] = tmp;
      }
      prow[par][t] = x[col];
    }
    __syncthreads();
    const double d = prow[par][c];
    const double inv = (d != 0.0) ? 1.0 / d : 0.0;
    // scale the column and apply the rank-1 update to the panel's remaining columns, row by row
    // (each row belongs to one thread, so the scaled value is reused without a barrier)
    for (int64_t i = col + 1 + t; i < nd; i += PT) {
      const double ai = a[i] * inv;
      a[i] = ai;
      for (int j0 = c + 1; j0 < nb; j0 += 8) {
        double xv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (j0 + q < nb) xv[q] = A[(k0 + j0 + q) * nd + i];
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (j0 + q < nb) A[(k0 + j0 + q) * nd + i] = xv[q] - ai * prow[par][j0 + q];
      }
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
// substitution), for the trailing H columns and all G columns.  One wave per column, lane r
// holding row r (NB = 64 = the wave size): the right-hand side never leaves registers.
__global__ void lu_trsm_kernel(const double* A, int64_t nd, double* H, double* G, int64_t nrhs, int64_t k0, int nb) {
  __shared__ double L[NB * NB];
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int r = e % nb, c = e / nb;
    L[r + c * NB] = A[(k0 + r) + (k0 + c) * nd];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t ntrail = nd - (k0 + nb);
  const int64_t ncol = ntrail + nrhs;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < ncol; j += nw) {
    double* x = ((j < ntrail) ? H + (k0 + nb + j) * nd : G + (j - ntrail) * nd) + k0;
    double y = lane < nb ? x[lane] : 0.0;
    for (int c = 0; c < nb; ++c) {
      const double yc = __shfl(y, c, 64);
      if (lane > c && lane < nb) y -= L[lane + c * NB] * yc;
    }
    if (lane < nb) x[lane] = y;
  }
}

// Backward substitution for the diagonal block U_pp at (r0, r0) of size nb: X[r0.., j] <- U^-1 X.
// One wave per column (lane r holds row r).
__global__ void upper_block_solve_kernel(const double* U, int64_t nd, double* G, int64_t nrhs, int64_t r0, int nb) {
  __shared__ double S[NB * NB];
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int r = e % nb, c = e / nb;
    S[r + c * NB] = U[(r0 + r) + (r0 + c) * nd];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < nrhs; j += nw) {
    double* x = G + j * nd + r0;
    double y = lane < nb ? x[lane] : 0.0;
    for (int r = nb - 1; r >= 0; --r) {
      if (lane == r) y = y / S[r + r * NB];
      const double yr = __shfl(y, r, 64);
      if (lane < r) y -= S[lane + r * NB] * yr;
    }
    if (lane < nb) x[lane] = y;
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
// for each row x of X[:, k0..k0+nb):  x <- x Lb^-T.  One wave per row (lane j holds entry j).
__global__ void trsm_right_lt_kernel(const double* L, int64_t nd, double* X, int64_t nrow, int64_t ldx, int64_t k0,
                                     int nb) {
  __shared__ double S[NB * NB];
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) S[e % nb + (e / nb) * NB] = L[(k0 + e % nb) + (k0 + e / nb) * nd];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < nrow; r += nw) {
    double y = lane < nb ? X[r + (k0 + lane) * ldx] : 0.0;
    for (int c = 0; c < nb; ++c) {
      if (lane == c) y = y / S[c + c * NB];
      const double yc = __shfl(y, c, 64);
      if (lane > c && lane < nb) y -= yc * S[lane + c * NB];
    }
    if (lane < nb) X[r + (k0 + lane) * ldx] = y;
  }
}

// Column-wise forward substitution: X[k0..k0+nb, j] <- Lb^-1 X[k0..k0+nb, j] (non-unit lower).
// One wave per column (lane r holds row r).
__global__ void trsm_left_l_kernel(const double* L, int64_t nd, double* X, int64_t ncol, int64_t ldx, int64_t k0,
                                   int nb) {
  __shared__ double S[NB * NB];
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) S[e % nb + (e / nb) * NB] = L[(k0 + e % nb) + (k0 + e / nb) * nd];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < ncol; j += nw) {
    double* x = X + j * ldx + k0;
    double y = lane < nb ? x[lane] : 0.0;
    for (int c = 0; c < nb; ++c) {
      if (lane == c) y = y / S[c + c * NB];
      const double yc = __shfl(y, c, 64);
      if (lane > c && lane < nb) y -= S[lane + c * NB] * yc;
    }
    if (lane < nb) x[lane] = y;
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

void lu_solve(double* H, int64_t nd, double* G, int64_t nrhs, hipStream_t s, DevBuf& ws, void* scratch,
              bool check) {
  if (nd == 0) return;
  HH_REQUIRE(check || scratch, "lu_solve: an unchecked solve needs the caller's scratch");
  DevBuf meta;
  // pivots [nd] and the status word: the caller's scratch (no allocation) or a buffer of our own
  char* m = scratch ? static_cast<char*>(scratch) : static_cast<char*>(meta.reserve(sizeof(int64_t) * nd + 256));
  // with caller scratch the trailing products must not split K (a split takes workspace)
  const int gemm_split = scratch ? 1 : 0;
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
      hipLaunchKernelGGL(lu_trsm_kernel, dim3(blocks_for(ncol * 64, 256)), dim3(256), 0, s, H, nd, H, G, nrhs, k0, nb);
      HH_HIP(hipGetLastError());
    }
    const int64_t mrows = nd - (k0 + nb);
    if (mrows > 0) {
      // A22 -= L21 U12  (H trailing part, then G)
      GemmArgs g;
      g.M = mrows; g.K = nb;
      g.A = H + (k0 + nb) + k0 * nd; g.lda = nd; g.a_kmajor = false;
      g.alpha = -1.0; g.beta = 1.0;
      g.ksplit = gemm_split;
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
  if (check) {
    const int bad = read_info(info, s);
    if (bad) throw Error(HH_ERR_INVALID, "lu_solve: singular matrix (zero pivot at column " + std::to_string(bad) + ")");
  }
  // back substitution U X = Y, blocks from the bottom
  const int64_t nblk = ceil_div(nd, NB);
  for (int64_t b = nblk - 1; b >= 0; --b) {
    const int64_t r0 = b * NB;
    const int nb = (int)std::min<int64_t>(NB, nd - r0);
    hipLaunchKernelGGL(upper_block_solve_kernel, dim3(blocks_for(nrhs * 64, 256)), dim3(256), 0, s, H, nd, G, nrhs, r0,
                       nb);
    HH_HIP(hipGetLastError());
    if (r0 > 0) {
      GemmArgs g;  // G[0:r0] -= U[0:r0, r0:r0+nb] X[r0:r0+nb]
      g.M = r0; g.N = nrhs; g.K = nb;
      g.A = H + r0 * nd; g.lda = nd; g.a_kmajor = false;
      g.B = G + r0; g.ldb = nd; g.b_kmajor = true;
      g.C = G; g.ldc = nd;
      g.alpha = -1.0; g.beta = 1.0;
      g.ksplit = gemm_split;
      gemm<double, double>(g, PREC_F64, s, ws);
    }
  }
}

// ---- the allocation-free solve ----
// Pivots and the status word live in this file-scope device array.  Each GPU holds its own copy
// (until the process ends); host code reaches it only through the address that
// hipGetSymbolAddress returns with that device current, kept per device below.
static __device__ __align__(256) char lu_scratch[LU_SCRATCH_BYTES];

void* lu_scratch_ptr() {
  static std::mutex mu;
  static void* ptr[64] = {};
  int dev = 0;
  HH_HIP(hipGetDevice(&dev));
  HH_REQUIRE(dev >= 0 && dev < 64, "lu scratch: device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    HH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(lu_scratch)));
    HH_REQUIRE(p != nullptr, "lu scratch: no device address");
    ptr[dev] = p;
  }
  return ptr[dev];
}

void lu_solve_noalloc(double* H, int64_t nd, double* G, int64_t nrhs, hipStream_t s) {
  HH_REQUIRE(nd <= LU_NOALLOC_MAX_ND, "lu_solve_noalloc: nd " + std::to_string(nd) + " exceeds the scratch limit " +
                                          std::to_string(LU_NOALLOC_MAX_ND));
  if (nd == 0) return;
  void* scratch = lu_scratch_ptr();
  DevBuf unused;  // never reserved: every product runs with ksplit = 1
  lu_solve(H, nd, G, nrhs, s, unused, scratch);
  HH_REQUIRE(unused.ptr == nullptr, "lu_solve_noalloc: internal allocation");
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
      hipLaunchKernelGGL(trsm_right_lt_kernel, dim3(blocks_for(m * 64, 256)), dim3(256), 0, s, H, nd, H + (k0 + nb), m, nd,
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
      hipLaunchKernelGGL(trsm_right_lt_kernel, dim3(blocks_for(n2 * 64, 256)), dim3(256), 0, s, L, nd, X, n2, ldx, k0, nb);
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
      hipLaunchKernelGGL(trsm_left_l_kernel, dim3(blocks_for(n2 * 64, 256)), dim3(256), 0, s, L, nd, X, n2, ldx, k0, nb);
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

void eigh(double* H, int64_t nd, double* w, hipStream_t s, DevBuf& ws) {
  HH_REQUIRE(nd > 0 && nd < (int64_t)1 << 31, "eigh: dimension out of range");
  rocblas_handle handle = nullptr;
  HH_REQUIRE(rocblas_create_handle(&handle) == rocblas_status_success, "eigh: rocblas_create_handle failed");
  struct Guard {
    rocblas_handle h;
    ~Guard() { (void)rocblas_destroy_handle(h); }
  } guard{handle};
  HH_REQUIRE(rocblas_set_stream(handle, s) == rocblas_status_success, "eigh: rocblas_set_stream failed");
  // workspace: the off-diagonal of the tridiagonal form (nd doubles) and the info word
  char* base = static_cast<char*>(ws.reserve(sizeof(double) * nd + 256));
  double* e = reinterpret_cast<double*>(base);
  rocblas_int* info = reinterpret_cast<rocblas_int*>(base + sizeof(double) * nd);
  const rocblas_status st = rocsolver_dsyevd(handle, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)nd, H,
                                             (rocblas_int)nd, w, e, info);
  HH_REQUIRE(st == rocblas_status_success, "eigh: rocsolver_dsyevd returned status " + std::to_string((int)st));
  rocblas_int hinfo = 0;
  HH_HIP(hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, s));
  HH_HIP(hipStreamSynchronize(s));
  if (hinfo != 0) throw Error(HH_ERR_INVALID, "eigh: syevd did not converge (info " + std::to_string(hinfo) + ")");
}

}  // namespace hh
