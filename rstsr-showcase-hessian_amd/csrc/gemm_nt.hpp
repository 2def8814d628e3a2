// "NT" GEMM on packed, K-contiguous operands (the RI-K hot contraction).
//
//   C[m, n] = alpha * sum_k A[m*lda + k] * B[n*ldb + k] + beta * C[m + n*ldc]     (C column-major)
//
// K must be a multiple of 32 and lda, ldb multiples of 8 (16-byte rows); operands are padded
// by their producers.  With `tiled` set (required for BF16X3) both operands are stored in K blocks of 32:
// A[((k / 32) * M + m) * 32 + k % 32], B[((k / 32) * N + n) * 32 + k % 32] (lda, ldb unused), so
// each K step of a workgroup reads one contiguous run of 64 bytes per row for all its rows.
// (a_rows / b_rows replace M / N as the K-block stride for row ranges of larger tiled arrays.)
// Element formats:
//   NT_F64    : A, B double; v_mfma_f64_16x16x4f64, FP64 accumulate.
//   NT_BF16X3 : A, B given as two bf16 arrays each (hi and lo = x - hi); hi*hi + hi*lo + lo*hi
//               on v_mfma_f32_16x16x32_bf16, FP32 accumulate; split-K partials summed in FP64.
#pragma once
#include "gemm.hpp"

namespace hh {

enum NtFormat : int { NT_F64 = 0, NT_BF16X3 = 1 };

struct NtArgs {
  int64_t M = 0, N = 0, K = 0;
  const void* A = nullptr;     // F64: double*;   BF16X3: hi array (unsigned short*)
  const void* A_lo = nullptr;  // BF16X3 only
  int64_t lda = 0;
  const void* B = nullptr;
  const void* B_lo = nullptr;
  int64_t ldb = 0;
  double* C = nullptr;
  float* C_f32 = nullptr;  // BF16X3: FP32 output instead of C
  int64_t ldc = 0;
  double alpha = 1.0, beta = 0.0;
  bool tiled = false;  // K-block tiled operands (BF16X3: required)
  // Tiled operands that are a row range of a larger tiled array: rows per K block of the array
  // (0 = M for A, N for B); the operand pointer then addresses its first row in K block 0.
  int64_t a_rows = 0, b_rows = 0;
  int ksplit = 0;  // 0: automatic; < 0: balance the last wave of workgroups over the CUs
  // BF16X3, instead of C: the bf16 split of alpha C[m, n] with m = j + w_nocc a, n = P + w_naux i
  // written as the RI-K term-2 operand hi / lo[(kk / 32) M + i + w_nocc a][32], kk = P + w_naux_pad j
  // (K split 1; entries with w_naux <= P < w_naux_pad are not written)
  unsigned short* W_hi = nullptr;
  unsigned short* W_lo = nullptr;
  int64_t w_nocc = 0, w_naux = 0, w_naux_pad = 0;
  hipEvent_t ev_begin = nullptr, ev_main_end = nullptr;
};

// Products whose caller leaves ksplit <= 0 are autotuned: the first product of each shape class
// (format, M, and N and K rounded up to powers of two) times the tile configurations and split
// rules on the stream and keeps the fastest (no timing while the stream is being captured into a
// graph).
void gemm_nt(const NtArgs& a, NtFormat fmt, hipStream_t s, DevBuf& ws);
void gemm_nt_set_autotune(bool on);
// the cached choice for a shape's class (false if none yet): configuration index and split rule
bool gemm_nt_choice(int64_t M, int64_t N, int64_t K, int* cfg, int* ksplit, NtFormat fmt = NT_F64);
// one product with configuration `cfg` (F64: 0 .. 2; BF16X3: 0 .. 2) and the split rule in a.ksplit
void gemm_nt_f64_config(const NtArgs& a, int cfg, hipStream_t s, DevBuf& ws);
void gemm_nt_bf16x3_config(const NtArgs& a, int cfg, hipStream_t s, DevBuf& ws);

// Split the column-major [nrow, ncol] matrix src (leading dimension ld) into bf16 hi / lo arrays
// (hi = bf16(x), lo = bf16(x - hi)) in the K-block tiled operand layout with K = rows padded to
// kpad (a multiple of 32): hi/lo[((r / 32) * ncol + n) * 32 + r % 32], zero for r >= nrow.
template <typename T>
void split_bf16_tiled(const T* src, int64_t ld, int64_t nrow, int64_t ncol, int64_t kpad, unsigned short* hi,
                      unsigned short* lo, hipStream_t s);

// The same split with K along the columns: src column-major [nrow, ncol] (leading dimension ld),
// K = columns padded to kpad: hi/lo[((c / 32) * nrow + r) * 32 + c % 32], zero for c >= ncol.
template <typename T>
void split_bf16_tiled_t(const T* src, int64_t ld, int64_t nrow, int64_t ncol, int64_t kpad, unsigned short* hi,
                        unsigned short* lo, hipStream_t s);

// Measured issue-rate ceiling of one MFMA instruction on this device (TFLOP/s): kind 0 =
// v_mfma_f32_16x16x32_bf16 (the BF16x3 products), 1 = v_mfma_f64_16x16x4f64 (the FP64 products),
// 2 = v_mfma_f32_32x32x16_bf16; `iters` loop trips of 8 (kind 2: 4) independent chains per wave,
// 8 waves per SIMD on every CU.
double mfma_peak_tflops(int kind, int iters, hipStream_t s);

}  // namespace hh
