// Blocked FP64 LU (partial pivoting), Cholesky, triangular solves and the Jacobi eigensolver.
//
// Right-looking blocked algorithms with panel width NB: the panel is factored by one
// workgroup (1024 lanes) in global memory (it stays in L2), the trailing update runs on the
// matrix-core GEMM.  Pivot choice: largest |a|, ties to the smallest row index (deterministic).
#include "linalg.hpp"

#include <algorithm>
#include <mutex>

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

// ---------------------------------------------------------------------------------------------
// Symmetric eigendecomposition by the parallel cyclic Jacobi method (Golub and Van Loan,
// Matrix Computations, sections 8.5.2-8.5.5): each sweep visits every (p, q) pair once in m - 1
// rounds of m / 2 disjoint pairs (round-robin "tournament" ordering, m = n rounded up to even,
// the extra index a dummy); one launch per round applies all of the round's rotations at once,
// A' = J^T A J and V' = V J, out of place.  The rotation of pair (p, q) is the symmetric Schur
// decomposition of [[a_pp, a_pq], [a_pq, a_qq]]: tau = (a_qq - a_pp) / 2 a_pq,
// t = sign(tau) / (|tau| + sqrt(1 + tau^2)), c = 1 / sqrt(1 + t^2), s = t c, with J_pp = J_qq = c,
// J_pq = s, J_qp = -s; then a'_pq = 0, a'_pp = a_pp - t a_pq, a'_qq = a_qq + t a_pq exactly.
// Sweeps stop when the off-diagonal Frobenius norm falls below 1e-15 of the whole (quadratic
// convergence: about one more sweep reaches round-off).

namespace {

struct JRot {
  int k;     // the other index of the pair (-1: unpaired)
  double c;  // J[i, i]
  double s;  // J[k, i]
  double t;  // tangent (diagonal update)
  bool is_p;
};

__device__ __forceinline__ int rr_player(int pos, int r, int m) { return pos == 0 ? 0 : 1 + (pos - 1 + r) % (m - 1); }
__device__ __forceinline__ int rr_pos(int x, int r, int m) {
  if (x == 0) return 0;
  int q = (x - 1 - r) % (m - 1);
  if (q < 0) q += m - 1;
  return 1 + q;
}

__device__ __forceinline__ JRot jacobi_rot(const double* __restrict__ A, int n, int m, int r, int i) {
  JRot o{-1, 1.0, 0.0, 0.0, true};
  const int pos = rr_pos(i, r, m);
  const int other = rr_player(m - 1 - pos, r, m);
  if (other >= n) return o;  // paired with the dummy index
  const bool is_p = pos < m / 2;
  const int p = is_p ? i : other, q = is_p ? other : i;
  const double app = A[p + (int64_t)n * p], aqq = A[q + (int64_t)n * q], apq = A[p + (int64_t)n * q];
  double t = 0.0;
  if (apq != 0.0) {
    const double tau = (aqq - app) / (2.0 * apq);
    if (fabs(tau) > 1e150)
      t = 0.5 / tau;
    else
      t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
  }
  const double c = 1.0 / sqrt(1.0 + t * t), sn = t * c;
  o.k = other;
  o.c = c;
  o.s = is_p ? -sn : sn;
  o.t = t;
  o.is_p = is_p;
  return o;
}

// one round: column j per blockIdx.y, rows i over threads
__global__ __launch_bounds__(256) void jacobi_round_kernel(const double* __restrict__ A, double* __restrict__ An,
                                                           const double* __restrict__ V, double* __restrict__ Vn,
                                                           int n, int m, int r) {
  const int j = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ JRot rj;
  if (threadIdx.x == 0) rj = jacobi_rot(A, n, m, r, j);
  __syncthreads();
  if (i >= n) return;
  const JRot ri = jacobi_rot(A, n, m, r, i);
  const int64_t N = n;
  double a;
  if (ri.k >= 0 && i == j) {
    const double apq = A[(ri.is_p ? i : ri.k) + N * (ri.is_p ? ri.k : i)];
    a = A[i + N * i] + (ri.is_p ? -ri.t : ri.t) * apq;
  } else if (ri.k >= 0 && ri.k == j) {
    a = 0.0;
  } else {
    a = ri.c * rj.c * A[i + N * j];
    if (rj.k >= 0) a += ri.c * rj.s * A[i + N * rj.k];
    if (ri.k >= 0) {
      a += ri.s * rj.c * A[ri.k + N * j];
      if (rj.k >= 0) a += ri.s * rj.s * A[ri.k + N * rj.k];
    }
  }
  An[i + N * j] = a;
  double v = rj.c * V[i + N * j];
  if (rj.k >= 0) v += rj.s * V[i + N * rj.k];
  Vn[i + N * j] = v;
}

// A = the symmetric matrix of H's lower triangle, V = identity
__global__ void jacobi_init_kernel(const double* __restrict__ H, int n, double* __restrict__ A, double* __restrict__ V) {
  const int64_t N = n;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= N * N) return;
  const int i = (int)(e % N), j = (int)(e / N);
  A[e] = i >= j ? H[e] : H[j + N * i];
  V[e] = i == j ? 1.0 : 0.0;
}

// per-block partial sums of (off-diagonal, all) squares
__global__ __launch_bounds__(256) void jacobi_norm_kernel(const double* __restrict__ A, int n, double* __restrict__ part) {
  const int64_t N = n;
  double off = 0.0, all = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * N; e += (int64_t)gridDim.x * blockDim.x) {
    const double x = A[e] * A[e];
    all += x;
    if (e % N != e / N) off += x;
  }
  __shared__ double so[256], sa[256];
  so[threadIdx.x] = off;
  sa[threadIdx.x] = all;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      so[threadIdx.x] += so[threadIdx.x + h];
      sa[threadIdx.x] += sa[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = so[0];
    part[2 * blockIdx.x + 1] = sa[0];
  }
}

__global__ void jacobi_diag_kernel(const double* __restrict__ A, int n, double* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = A[i + (int64_t)n * i];
}

__global__ void jacobi_minus_identity_kernel(double* __restrict__ G, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) G[i + (int64_t)n * i] -= 1.0;
}

hipError_t copy_async(double* dst, const double* src, int64_t n, hipStream_t s) {
  return hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, s);
}

// H[:, k] = V[:, perm[k]], w[k] = d[perm[k]]
__global__ void jacobi_sort_kernel(const double* __restrict__ V, const double* __restrict__ d,
                                   const int* __restrict__ perm, int n, double* __restrict__ H, double* __restrict__ w) {
  const int64_t N = n;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= N * N) return;
  const int i = (int)(e % N), k = (int)(e / N);
  H[e] = V[i + N * perm[k]];
  if (i == 0) w[k] = d[perm[k]];
}

}  // namespace

void eigh(double* H, int64_t nd, double* w, hipStream_t s, DevBuf& ws) {
  HH_REQUIRE(nd > 0 && nd <= 32768, "eigh: dimension out of range (1 .. 32768)");
  const int n = (int)nd;
  const int m = n + (n & 1);
  const int64_t nn = nd * nd;
  constexpr int NPART = 512;
  char* base = static_cast<char*>(ws.reserve(4 * nn * sizeof(double) + (2 * NPART + nd) * sizeof(double) +
                                             nd * sizeof(int) + 1024));
  double* A0 = reinterpret_cast<double*>(base);
  double* A1 = A0 + nn;
  double* V0 = A1 + nn;
  double* V1 = V0 + nn;
  double* part = V1 + nn;
  double* d = part + 2 * NPART;
  int* perm_d = reinterpret_cast<int*>(d + nd);
  const unsigned nb = (unsigned)ceil_div(nn, 256);
  hipLaunchKernelGGL(jacobi_init_kernel, dim3(nb), dim3(256), 0, s, H, n, A0, V0);
  HH_HIP(hipGetLastError());
  std::vector<double> hp(2 * NPART);
  const unsigned npb = (unsigned)std::min<int64_t>(NPART, ceil_div(nn, 256));
  bool done = n == 1;
  for (int sweep = 0; sweep < 60 && !done; ++sweep) {
    for (int r = 0; r < m - 1; ++r) {
      hipLaunchKernelGGL(jacobi_round_kernel, dim3((unsigned)ceil_div(nd, 256), (unsigned)n), dim3(256), 0, s, A0, A1,
                         V0, V1, n, m, r);
      std::swap(A0, A1);
      std::swap(V0, V1);
    }
    HH_HIP(hipGetLastError());
    hipLaunchKernelGGL(jacobi_norm_kernel, dim3(npb), dim3(256), 0, s, A0, n, part);
    HH_HIP(hipMemcpyAsync(hp.data(), part, 2 * npb * sizeof(double), hipMemcpyDeviceToHost, s));
    HH_HIP(hipStreamSynchronize(s));
    double off = 0.0, all = 0.0;
    for (unsigned b = 0; b < npb; ++b) {
      off += hp[2 * b];
      all += hp[2 * b + 1];
    }
    if (!(off == off && all == all)) throw Error(HH_ERR_INVALID, "eigh: non-finite matrix elements");
    done = off <= 1e-30 * all;
  }
  if (!done) throw Error(HH_ERR_INVALID, "eigh: Jacobi sweeps did not converge");
  hipLaunchKernelGGL(jacobi_diag_kernel, dim3((unsigned)ceil_div(nd, 256)), dim3(256), 0, s, A0, n, d);
  std::vector<double> hd(nd);
  HH_HIP(hipMemcpyAsync(hd.data(), d, nd * sizeof(double), hipMemcpyDeviceToHost, s));
  HH_HIP(hipStreamSynchronize(s));
  std::vector<int> perm(nd);
  for (int i = 0; i < n; ++i) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return hd[a] < hd[b]; });
  HH_HIP(hipMemcpyAsync(perm_d, perm.data(), nd * sizeof(int), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(jacobi_sort_kernel, dim3(nb), dim3(256), 0, s, V0, d, perm_d, n, H, w);
  HH_HIP(hipGetLastError());
  // The rotations leave V orthonormal to O(sweeps n eps); one Newton-Schulz step of the polar
  // factor, V <- V - V (V^T V - I) / 2, takes that to O(eps^2) terms: G = V^T V - I in A1,
  // A0 = V - G V / 2, then V = A0.
  DevBuf gws;
  GemmArgs g;
  g.M = g.N = g.K = nd;
  g.A = H;
  g.lda = nd;
  g.a_kmajor = true;
  g.B = H;
  g.ldb = nd;
  g.b_kmajor = true;
  g.C = A1;
  g.ldc = nd;
  gemm<double, double>(g, PREC_F64, s, gws);
  hipLaunchKernelGGL(jacobi_minus_identity_kernel, dim3((unsigned)ceil_div(nd, 256)), dim3(256), 0, s, A1, n);
  HH_HIP(copy_async(A0, H, nn, s));
  GemmArgs h;
  h.M = h.N = h.K = nd;
  h.A = H;
  h.lda = nd;
  h.a_kmajor = false;
  h.B = A1;
  h.ldb = nd;
  h.b_kmajor = true;
  h.C = A0;
  h.ldc = nd;
  h.alpha = -0.5;
  h.beta = 1.0;
  gemm<double, double>(h, PREC_F64, s, gws);
  HH_HIP(copy_async(H, A0, nn, s));
  HH_HIP(hipStreamSynchronize(s));  // perm is a host vector; gws is freed on return
}

}  // namespace hh
