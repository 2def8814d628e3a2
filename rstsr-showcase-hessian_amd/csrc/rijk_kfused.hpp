// Fused BF16x3 RI-K contraction (the first K term with its operand generated in place).
//
// K_a = sum_P B_P [ bra_a O_P + C_o W_aP ]  (reference: src/ri_jk/hess_r.rs:145-183) splits into
//   term 1:  sum_P B_P Q_aP,   Q_aP = bra_a O_P          (2 nao^2 nocc per (a, P): the hot part)
//   term 2:  sum_P X_P W_aP,   X_P = B_P C_o, W_aP = bra_a^T X_P   (nocc-sized products)
// This unit computes term 1 without materialising Q: every workgroup owns all nao rows of the
// output and 128 columns (i, a); per K step (one block of 32 nu and one P) each wave builds its
// 32 x 32 block of Q on the matrix cores from bra (kept in registers for the whole step range)
// and O_P (16-byte loads), splits it into bf16 hi / lo in registers and uses it directly as the B
// operand of the contraction with the 32-row blocks of B_P staged in LDS.
//
// Term 2 operands: X' (X as the K-block tiled bf16 pair with K = P + naux_pad j) and W' (W with
// rows i + nocc a, the same K) for one NT product.
#pragma once
#include "hh_common.hpp"

namespace hh {

// 8 <= nocc <= 80 and nao <= 416: the register budget of one wave per SIMD with all nao rows of
// the output in accumulators.
bool kfused_supported(int64_t nao, int64_t nocc);
inline int64_t kfused_nb(int64_t nao) { return ceil_div(nao, 32); }
inline int64_t kfused_jp(int64_t nocc) { return 16 * ceil_div(nocc, 16); }
inline int64_t kfused_naux_pad(int64_t naux) { return ceil_div(naux, 32) * 32; }
inline int64_t kfused_naux_e(int64_t naux) { return 2 * ceil_div(naux, 2); }
// row blocks of one packed B tile (the kernel instance covering nb blocks)
int kfused_nmb(int64_t nb);
// elements of one array (hi or lo) of the packed B and O operands
int64_t kfused_b_len(int64_t nao, int64_t naux);
int64_t kfused_o_len(int64_t nocc, int64_t naux);

// cderi [nao_tp, naux] (FP64, packed upper) -> hi / lo [nb][naux_e][32 nmb][32] plus one tile:
// row mu, 32 nu of block b in the order of the MFMA operand (see rijk_kfused.hip), zero outside
// nao and for P >= naux.
void kfused_pack_b(const double* cderi, int64_t nao, int64_t naux, unsigned short* hi, unsigned short* lo,
                   hipStream_t s);
// O[j + nocc (P + naux i)] (FP64) -> hi / lo [naux_e + 2][nocc][jp] (O_P is symmetric), zero for
// j >= nocc and P >= naux
void kfused_pack_o(const double* O, int64_t nocc, int64_t naux, unsigned short* hi, unsigned short* lo,
                   hipStream_t s);
// bra [nao, nocc, nprop] (FP64) -> hi / lo [nprop][nb][32][jp], zero outside (nao, nocc)
void kfused_pack_bra(const double* bra, int64_t nao, int64_t nocc, int64_t nprop, unsigned short* hi,
                     unsigned short* lo, hipStream_t s);
// X[mu + nao (P + naux j)] (FP32) -> hi / lo [(kk / 32) nao + mu][32], kk = P + naux_pad j
void kfused_pack_x(const float* X, int64_t nao, int64_t naux, int64_t nocc, unsigned short* hi, unsigned short* lo,
                   hipStream_t s);
// The term-2 operand W' = hi / lo [(kk / 32) R + i + nocc a][32], R = nocc nprop, kk = P + naux_pad j,
// holds W_aP[j, i] = sum_mu bra_a[mu, j] X_P[mu, i]; the W product writes it from its epilogue
// (NtArgs::W_hi) for P < naux, and this zeroes the padding naux <= P < naux_pad.
void kfused_zero_w_pad(int64_t naux, int64_t nocc, int64_t nprop, unsigned short* hi, unsigned short* lo,
                       hipStream_t s);

struct KfusedArgs {
  const unsigned short *bh = nullptr, *bl = nullptr;  // kfused_pack_b
  const unsigned short *oh = nullptr, *ol = nullptr;  // kfused_pack_o
  const unsigned short *rh = nullptr, *rl = nullptr;  // kfused_pack_bra
  int64_t nao = 0, naux = 0, nocc = 0, nprop = 0;
  double alpha = 1.0, beta = 0.0;
  double* out = nullptr;  // [nao, nocc, nprop]: alpha * term 1 + beta * out
  hipEvent_t ev_begin = nullptr, ev_main_end = nullptr;
};
void kfused_term1(const KfusedArgs& a, DevBuf& ws, hipStream_t s);

}  // namespace hh
