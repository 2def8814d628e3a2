// Dense FP64 linear algebra on the GPU: blocked LU with partial pivoting and triangular solves.
#pragma once
#include "gemm.hpp"

namespace hh {

// Solve H X = G in place: H [nd, nd] (column-major, ld = nd) is overwritten by its LU factors,
// G [nd, nrhs] (ld = nd) by X.  Throws Error(HH_ERR_INVALID) if H is singular (zero pivot).
// Replaces rstsr linalg solve_general (LAPACK getrf/getrs) as used by krylov_block.rs:214.
// `scratch` (optional): device memory of at least 8 nd + 256 bytes for the pivots and the status
// word; with it the solve allocates nothing (ws stays untouched).
// check = false: no host synchronisation; the status word (int, 0 = success, else the 1-based
// column of a zero pivot) is left at scratch + 8 nd for the caller (scratch required).
void lu_solve(double* H, int64_t nd, double* G, int64_t nrhs, hipStream_t s, DevBuf& ws, void* scratch = nullptr,
              bool check = true);

// The same solve without any device allocation: pivots and status in a file-scope device array
// of LU_SCRATCH_BYTES (nd <= LU_NOALLOC_MAX_ND, else Error(HH_ERR_INVALID)).  The array is shared
// by every call on one GPU and keeps the last call's contents: calls that overlap on one GPU race.
constexpr int64_t LU_NOALLOC_MAX_ND = 8192;
constexpr int64_t LU_SCRATCH_BYTES = 8 * LU_NOALLOC_MAX_ND + 256;
void lu_solve_noalloc(double* H, int64_t nd, double* G, int64_t nrhs, hipStream_t s);

// Cholesky H = L L^T in place (lower, column-major, ld = nd); returns false if not positive
// definite (a pivot <= 0 was met).
bool cholesky(double* H, int64_t nd, hipStream_t s, DevBuf& ws);

// In-place triangular solves with the lower Cholesky factor L [nd, nd]:
//   side_right = false:  X <- L^-1 X     (X [nd, ncol], ld = ldx)
//   side_right = true :  X <- X L^-T     (X [nrow, nd], ld = ldx)
void trsm_lower(const double* L, int64_t nd, double* X, int64_t nrow_or_ncol, int64_t ldx, bool side_right,
                hipStream_t s, DevBuf& ws);

// Symmetric eigendecomposition H = V diag(w) V^T in FP64 on the device (rstsr linalg eigh as
// used by decomp_j2c_eig / decomp_j2c_cd, pure_decompose.rs:11-20, 41-45): the lower triangle
// of H [nd, nd] is read, H is overwritten by the orthonormal eigenvectors (columns), w [nd]
// (device) receives the eigenvalues in ascending order.  Parallel cyclic Jacobi (one launch per
// round of n / 2 disjoint rotations, a host read per sweep); workspace 4 nd^2 doubles; throws
// Error(HH_ERR_INVALID) if the sweeps do not converge or the matrix is not finite.
void eigh(double* H, int64_t nd, double* w, hipStream_t s, DevBuf& ws);

}  // namespace hh
