// Generation of the hot-contraction operand T of the RI-K response (BF16x3 storage).
//
// For the perturbations a = a0 .. a0+na-1 of one chunk and every column c = P + naux*i
// (c < nc = naux*nocc):
//   T[mu, c, al] = sum_j bra[mu, j, a] O[j, c] + sum_k Co[mu, k] W[k + nocc*a, c]
// i.e. T_a = [bra_a | C_o] [O ; W_a] with inner length 2 nocc (reference algebra:
// src/ri_jk/hess_r.rs:145-176, the two K terms bra_a O_P and C_o (bra_a^T X_P) before the
// contraction with B_P).  T is written as a bf16 hi / lo pair in the K-block tiled layout of the
// hot contraction: row n = i + nocc*al, K index k = mu + kpad*P, element at
// th/tl[((k / 32) * nocc*na + n) * 32 + k % 32]; every row mu < kpad is written (zero for mu >= nao).
#pragma once
#include "hh_common.hpp"

namespace hh {

// true when the kernel covers nocc (2 nocc <= 256, the register-resident inner dimension)
bool tgen_supported(int64_t nocc);

// bra [nao, nocc, nprop], Co [nao, nocc], O [nocc, nc], W [ldw, nc] (ldw = nocc*nprop), all float,
// column-major.  kpad >= nao, kpad % 32 == 0.  `ws` holds the packed [bra_a | C_o].
void tgen_bf16x3(const float* bra, const float* Co, const float* O, const float* W, int64_t nao, int64_t nocc,
                 int64_t nprop, int64_t nc, int64_t a0, int64_t na, int64_t kpad, unsigned short* th,
                 unsigned short* tl, DevBuf& ws, hipStream_t s);

}  // namespace hh
