// Block Krylov solver for (I + A) x = b on the GPU (reference: src/util/krylov_block.rs).
#pragma once
#include <functional>

#include "gemm.hpp"

namespace hh {

// A acting on a block: ax[n, nblock] = A x[n, nblock] (device, column-major, ld = n), enqueued
// on the solver's stream.
using Aop = std::function<void(const double* x, int64_t nblock, double* ax)>;

struct KrylovStats {
  int cycles = 0;       // aop calls on trial blocks (the reference's total_cycles)
  int restarts = 0;
  int aop_calls = 0;    // including residual evaluations at restarts / with x0
  double max_innerprod = 0.0;
};

// Same contract as krylov_block.rs:39-197: b [n, nset], optional x0 [n, nset], x out [n, nset].
// Convergence: max ||v||^2 of the new trial block < max(lindep, tol^2).  Hard restart every
// max_space cycles.  Throws Error(HH_ERR_NOT_CONVERGED) where the reference panics; x then
// still holds the last approximation.
void krylov_block(const Aop& aop, const double* b, const double* x0, int64_t n, int64_t nset, double tol,
                  int max_cycle, int max_space, double lindep, double* x, KrylovStats* stats, hipStream_t s);

// Non-normalised Gram-Schmidt of the columns of V [n, m] (V is used as scratch).  Kept columns
// (||v||^2 > lindep after projection) are written in order to R [n, *] starting at column
// `r_off`; their squared norms are appended to `norms`.  Returns the number kept.
int64_t orth_block(double* V, int64_t n, int64_t m, double lindep, double* R, int64_t r_off,
                   std::vector<double>& norms, hipStream_t s, DevBuf& ws);

}  // namespace hh
