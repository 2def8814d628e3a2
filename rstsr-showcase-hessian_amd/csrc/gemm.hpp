// Strided, batched, split-K GEMM on CDNA4 matrix cores (gfx950).
//
//   C[m, n] = alpha * sum_k A[m, k] B[k, n] + beta * C[m, n]
//
// A[m, k] lives at  a_kmajor ? A[m*lda + k] : A[m + k*lda]
// B[k, n] lives at  b_kmajor ? B[k + n*ldb] : B[k*ldb + n]
// C[m, n] lives at  C[m + n*ldc]                      (column-major, like the reference's tensors)
// Two batch dimensions (b0 fastest) with independent element strides for A, B, C.
//
// Arithmetic modes:
//   PREC_F64    : v_mfma_f64_16x16x4f64, FP64 accumulate.
//   PREC_F32    : v_mfma_f32_16x16x4f32, FP32 accumulate.
//   PREC_BF16X3 : operands split x = hi + lo (two bf16), hi*hi + hi*lo + lo*hi on
//                 v_mfma_f32_16x16x32_bf16, FP32 accumulate.
// Split-K partial sums are reduced in FP64 in a fixed order (deterministic).
#pragma once
#include "hh_common.hpp"

namespace hh {

enum Prec : int { PREC_F64 = HH_PREC_FP64, PREC_F32 = HH_PREC_FP32, PREC_BF16X3 = HH_PREC_BF16X3 };

struct GemmArgs {
  int64_t M = 0, N = 0, K = 0;
  const void* A = nullptr;
  int64_t lda = 0;
  bool a_kmajor = false;
  const void* B = nullptr;
  int64_t ldb = 0;
  bool b_kmajor = true;
  void* C = nullptr;
  int64_t ldc = 0;
  double alpha = 1.0, beta = 0.0;
  int64_t nb0 = 1, nb1 = 1;
  int64_t sA0 = 0, sA1 = 0, sB0 = 0, sB1 = 0, sC0 = 0, sC1 = 0;
  int ksplit = 0;  // 0: choose automatically
  // Grouped K ranges: non-null device array [nb1][2]; batch index b1 then sums only over
  // k0 <= k < k1 (k0 = krange[2 b1], k1 = krange[2 b1 + 1], 0 <= k0, k1 <= K) of the operands
  // addressed as for the whole K: products that share A and B and differ in a contiguous K
  // slice (per-atom blocks) run as one launch.  krange_max: the longest slice (sizes the K split,
  // which then splits every slice into chunks of the same length).  No second operand pair.
  const int64_t* krange = nullptr;
  int64_t krange_max = 0;
  // Optional second operand pair concatenated along K: for k >= K1 the product uses
  // A2[m, k - K1] B2[k - K1, n] (same majorness as A / B; K is the total length).  K1 = -1: off.
  int64_t K1 = -1;
  const void* A2 = nullptr;
  const void* B2 = nullptr;
  int64_t lda2 = 0, ldb2 = 0;
  int64_t sA2_0 = 0, sA2_1 = 0, sB2_0 = 0, sB2_1 = 0;
  // Non-null: store C as two u16 arrays, C (bf16 hi) and C_lo (bf16 of the remainder), same
  // indexing; beta must be 0 and ksplit 1.
  void* C_lo = nullptr;
  // Optional timing events recorded on the launch stream: before the main kernel, after it
  // (before any split-K reduction) and after the reduction.
  hipEvent_t ev_begin = nullptr, ev_main_end = nullptr, ev_end = nullptr;
};

// Element types: TIn for A and B (double or float), TOut for C (double or float).
// F64 mode requires TIn = double.  `ws` is grown as needed for split-K partials.
template <typename TIn, typename TOut>
void gemm(const GemmArgs& g, Prec prec, hipStream_t stream, DevBuf& ws);

}  // namespace hh
