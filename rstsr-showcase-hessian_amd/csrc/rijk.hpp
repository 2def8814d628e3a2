// RI-JK response component (RHF nspin = 1, UHF nspin = 2).
//
// Mirrors RHessRIJK / UHessRIJK's response side (src/ri_jk/hess_r.rs:1737-1954,
// src/ri_jk/hess_u.rs:38-316) and the separated core get_rijk_response_bra_separated
// (hess_r.rs:82-186).
#pragma once
#include <memory>

#include "gemm.hpp"
#include "gemm_nt.hpp"

namespace hh {

// Device intermediates of a resident-matrix build (G, X transposed, the unpacked V of both spins,
// the packed O and V triangles); a caller building many matrices keeps one set for all of them.
struct ResidentScratch {
  DevBuf g, xr, vp, o2t, vtp;
};

class Rijk {
 public:
  ~Rijk();
  Rijk(const double* cderi, int64_t nao, int64_t naux, double scale_j, double scale_k, int nspin, hipStream_t s);
  // generate_cderi_with_decomp (hess_r.rs:1725-1735) from device j3c [nao_tp, naux] (packed
  // upper, column-major) and j2c [naux, naux].  policy HH_J2C_CD: j2c = L L^T (decomp_j2c_cd,
  // pure_decompose.rs:28-55, with the eigenvalue-floor refactorisation when threshold > 0),
  // cderi = j3c L^-T; HH_J2C_EIG: cderi = j3c J^-1/2 over eigenvalues >= threshold
  // (decomp_j2c_eig, pure_decompose.rs:11-20).  Throws if the Cholesky factorisation fails with
  // no threshold.
  static std::unique_ptr<Rijk> from_j3c(const double* j3c, const double* j2c, int64_t nao, int64_t naux, int policy,
                                        double threshold, double scale_j, double scale_k, int nspin, hipStream_t s);
  // Lower Cholesky factor L of the metric (HH_J2C_CD handles), or null.
  const double* j2c_l() const { return j2c_l_.ptr ? j2c_l_.as<double>() : nullptr; }
  // Symmetric J^-1/2 (HH_J2C_EIG handles), or null.  Either factor F gives V^-1 = F^-T F^-1
  // (L) or F^T F (J^-1/2) and cderi = j3c F^-T (L) or j3c F (J^-1/2).
  const double* j2c_l_inv() const { return j2c_linv_.ptr ? j2c_linv_.as<double>() : nullptr; }
  // B [nao, nao, naux] (cderi unpacked, FP64)
  const double* bfull() const { return bfull_.as<double>(); }
  // device workspace of the skeleton call (rijk_skeleton.hip), kept between calls so a repeated
  // call allocates nothing; released with the handle
  std::shared_ptr<void>& skeleton_workspace() { return skel_ws_; }

  void prepare_response(const double* const* mo_coeff, const double* const* mo_occ, const int64_t* nmo);
  // resp[s] = scale_j * f_s * J_ao mocc_s + scale_k * K_s  (f = 1 RHF, 0.5 UHF)
  void get_response_bra(const double* const* bra, int64_t nprop, double* const* resp, Prec prec,
                        bool accumulate = false);
  // j_ao (unscaled) and k_bra[s] (unscaled, sign included); either may be null.
  void response_bra_separated(const double* const* bra, int64_t nprop, double* j_ao, double* const* k_bra, Prec prec);

  // pair[m, i] = sum_P ((C^T B_P C_o)[m, i])^2 = (mi|mi) in the RI approximation, for all nmo
  // MOs m of the device coefficients C [nao, nmo] (spin s; needs prepare_response).
  void pair_norms(int s, const double* mo_coeff, int64_t nmo, double* pair);

  int64_t nao() const { return nao_; }
  int64_t naux() const { return naux_; }
  int nspin() const { return nspin_; }
  int64_t nocc(int s) const { return nocc_[s]; }
  bool prepared() const { return prepared_; }
  hipStream_t stream() const { return stream_; }
  double scale_j() const { return scale_j_; }
  double scale_k() const { return scale_k_; }

  // Perturbations per K-contraction batch (0 = choose from the size of the intermediate).
  int64_t k_chunk = 0;
  // BF16x3 K path: 0 = the fused contraction where supported (rijk_kfused.hpp), 2 = always the
  // materialised operand (T generation + NT contraction).
  int k_path = 0;
  // Response algorithm: 0 = per call (the reference's algebra, default); 1 = resident response
  // matrix (RHF and UHF, FP64 and BF16x3; rijk_resident.hip): M [R, R], R = nao sum_s nocc_s,
  // built on the first call after prepare_response and kept, each call one product with M.
  // Falls back to the per-call algorithm when M would exceed resident_budget bytes or for FP32.
  int response_path = 0;
  int64_t resident_budget = 24LL << 30;
  bool resident_fits(Prec prec) const;
  // The resident response matrix in FP64, row-major [R][kpad] (R = nao sum_s nocc_s, the alpha
  // block first; kpad = R rounded up to 32, zeros past R), built now if it is not (needs
  // prepare_response; independent of response_path and of the budget).  `scratch` (optional):
  // the build's intermediates go there and stay for the caller's next build.
  const double* resident_matrix(int64_t* R, int64_t* kpad, ResidentScratch* scratch = nullptr);

  // Stage timing with HIP events on the handle's stream (the reference keeps wall-clock
  // `timing` vectors per component, hess_r.rs:1746).  Indices of the accumulated times (ms):
  enum { T_K_HOT = 0, T_K_TOTAL = 1, T_J_TOTAL = 2, T_RESPONSE = 3, T_COUNT = 4 };
  void set_profiling(bool on);
  // times[T_*] in ms and, at index T_COUNT, the number of hot-kernel launches
  void get_timing(double* out) const;
  void reset_timing();

 private:
  // J: j_ao [nao, nao, nprop] from total dm1 (sum over spins); written unscaled.
  void build_j_ao(const double* const* bra, int64_t nprop, Prec prec, double* j_ao);
  // K: out_s[nao, nocc_s, nprop] = alpha * (K_s without its minus sign) + beta * out_s
  void build_k(int s, const double* bra, int64_t nprop, Prec prec, double alpha, double beta, double* out);
  void ensure_mixed();
  void build_k_fused(int s, const double* bra, int64_t nprop, const unsigned short* wp, double alpha, double beta,
                     double* out);
  void build_resident(Prec prec, ResidentScratch* scratch = nullptr);
  void apply_resident(const double* const* bra, int64_t nprop, double* const* resp, Prec prec, bool accumulate);
  // Packed hot-GEMM operand Bp[u][P][mu_pad] (mu_pad = kpad_ = nao rounded up to 32, zeros
  // beyond nao): FP64 copy and/or bf16 hi/lo split, built once per handle.
  void ensure_packed(Prec prec);
  hipEvent_t ev(int i);
  int mark();                     // record a fresh event now, return its index
  void span(int key, int begin);  // accumulate (begin event .. now) into times_[key]
  void flush_timing();

  bool profiling_ = false;
  std::vector<hipEvent_t> events_;
  std::vector<std::pair<int, std::pair<int, int>>> pending_;  // (key, (begin event, end event))
  int next_event_ = 0;
  double times_[T_COUNT] = {0, 0, 0, 0};
  int64_t hot_launches_ = 0;

  int64_t nao_, naux_, nao_tp_;
  double scale_j_, scale_k_;
  int nspin_;
  hipStream_t stream_;
  bool prepared_ = false;
  bool mixed_ready_ = false;
  bool xt_ready_ = false;
  bool kf_ready_ = false;
  int64_t nocc_[2] = {0, 0};

  DevBuf j2c_l_;      // [naux, naux] lower Cholesky factor of the metric (from_j3c, HH_J2C_CD)
  DevBuf j2c_linv_;   // [naux, naux] J^-1/2 of the metric (from_j3c, HH_J2C_EIG)
  DevBuf cderi_;      // [nao_tp, naux] f64
  DevBuf bfull_;      // [nao, nao, naux] f64 (unpacked)
  std::shared_ptr<void> skel_ws_;
  int64_t kpad_ = 0;
  DevBuf bp64_, bp32_, bp_hi_, bp_lo_;  // packed hot-GEMM A operand
  DevBuf tp_, tp_hi_, tp_lo_;           // packed hot-GEMM B operand (one perturbation chunk)
  DevBuf gp_;                           // [bra_a | C_o] packed for the BF16x3 T generation
  int64_t tp_key_[2] = {0, 0};          // (element size, length) of tp_ when its padding was zeroed
  DevBuf cderi32_;    // f32 copies for the mixed-precision modes
  DevBuf bfull32_;
  DevBuf mocc_[2];    // [nao, nocc] f64
  DevBuf mocc32_[2];
  DevBuf x_[2];       // X[mu, P, i] = (B_P C_o)[mu, i]
  DevBuf x32_[2];
  DevBuf o_[2];       // O[j, P, i] = (C_o^T B_P C_o)[j, i]
  DevBuf o32_[2];
  DevBuf xt_hi_[2], xt_lo_[2];  // X as BF16x3 NT operand: K = mu (padded to kpad), K-block tiled
  // per-call scratch
  DevBuf brat_;                 // bra as BF16x3 NT operand (hi then lo), K-block tiled
  DevBuf w_, t_, bra32_, dm_, dmtp_, itm_, jtp_, jao_, ws_;
  // BF16x3 RI-J operands (built on the first BF16x3 J): cderi split into bf16 hi / lo in the
  // K-block tiled NT layout twice, K = the AO pair (cj, rows P) and K = P (cjt, rows the AO pair);
  // jsp_: the per-call split dm_tp and itm
  DevBuf cj_hi_, cj_lo_, cjt_hi_, cjt_lo_, jsp_;
  // fused BF16x3 K (rijk_kfused.hpp): cderi in the contraction layout (per handle); O and X'
  // (hi then lo, per prepare_response and spin); per call the packed bra, W' and the partials
  DevBuf kb_hi_, kb_lo_, ko_[2], kx_[2], kr_, kw_, kf_ws_;
  // resident response matrix (per prepare_response): BF16x3 tiled hi / lo, FP64 row-major; the
  // per-call split / padded bra
  DevBuf res_hi_, res_lo_, res64_, res_b_;
  bool res_ready_[2] = {false, false};  // FP64, BF16x3
  bool j_bf16_ready_ = false;
};

}  // namespace hh
