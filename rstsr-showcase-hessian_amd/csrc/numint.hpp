// Numerical-integration grid contractions for the CPKS response (reference:
// src/numint_matmul/{nimatmul.rs, pure_eval_rho.rs, pure_xcpot.rs, hess_rks.rs}).
#pragma once
#include <vector>

#include "gemm.hpp"

namespace hh {

class Becke;

// XCDenType (src/xceff/flags.rs:19-24): RHO 0, SIGMA 1, TAU 2.
enum DenType : int { DEN_RHO = 0, DEN_SIGMA = 1, DEN_TAU = 2 };
inline int den_nvar(int d) { return d == DEN_RHO ? 1 : (d == DEN_SIGMA ? 4 : 5); }
inline int den_ncomp(int d) { return d == DEN_RHO ? 1 : 4; }
// AO components the Hessian setup reads (derivative order 2 for RHO, 3 for SIGMA / TAU)
int setup_ao_ncomp(int den);

class Numint {
 public:
  // ao [ng, nao, ncomp] (column-major, AO values and first derivatives, as NIMatmul's cached
  // AO, nimatmul.rs:105-126) and weights [ng] are copied to the handle.
  Numint(const double* ao, int64_t ng, int64_t nao, int64_t ncomp, const double* weights, hipStream_t s);

  // rho1[ng, nvar, nprop] = make_rho_from_one_bra_mult_ket(bra, kets) (pure_eval_rho.rs:216-334)
  void rho1_one_bra_mult_ket(int den, const double* bra, int64_t nocc, const double* kets, int64_t nprop, double* rho1,
                             Prec prec);
  // out[nao, nocc, nprop] = make_rks_fxc_pot_with_eff_bra_trans (pure_xcpot.rs:370-465)
  void fxc_pot_bra_trans(int den, const double* fxc_eff, const double* rho1, int64_t nprop, const double* bra,
                         int64_t nocc, double* out, Prec prec, double alpha = 1.0, bool accumulate = false);
  // resp = 4 fxc_pot_bra_trans(rho1(mocc, mo1_bra)) (get_rks_response_bra, hess_rks.rs:570-605),
  // fused per grid batch so rho1 never leaves the batch.
  void rks_response_bra(int den, const double* fxc_eff, const double* mo1_bra, int64_t nprop, const double* mocc,
                        int64_t nocc, double* out, Prec prec, bool accumulate = false);
  // BF16X3 rks_response_bra: the kets product fused with the pointwise pass (nocc <= 128; the
  // results are bit-identical to the unfused kernels').  Default on.
  bool fused_kets = true;

  // UKS, one array per spin (pure_xcpot.rs:921-1028): out[s][nao, nocc_s, nprop] (+)= alpha *
  // bra[s]-transformed potential of rho1 [ng, nvar, 2, nprop] under fxc_eff [ng, nvar, 2, nvar, 2],
  // wv_s[g, x, a] = w_g sum_{y,t} fxc[g, y, t, x, s] rho1[g, y, t, a]
  void uks_fxc_pot_bra_trans(int den, const double* fxc_eff, const double* rho1, int64_t nprop,
                             const double* const* bra, const int64_t* nocc, double* const* out, Prec prec,
                             double alpha = 1.0, bool accumulate = false);
  // get_uks_response_bra (hess_uks.rs:390-441): rho1 per spin from (mocc_s, mo1_bra_s), then 2x the
  // bra-transformed potential; fused per grid batch like the RKS response
  void uks_response_bra(int den, const double* fxc_eff, const double* const* mo1_bra, int64_t nprop,
                        const double* const* mocc, const int64_t* nocc, double* const* out, Prec prec,
                        bool accumulate = false);

  // The exchange-correlation potential matrix out [nao, nao] = dE_xc / dD of an RKS density from
  // vxc_eff [ng, nvar] (the rho_t derivatives, weights not applied): the Fock contribution of the
  // functional in an SCF.  Needs AO first derivatives for SIGMA / TAU.
  void vxc_matrix(int den, const double* vxc_eff, double* out);
  // sum_g w_g exc_g rho_g (exc [ng] the energy per particle, rho the density column)
  double xc_energy(const double* exc, const double* rho);

  // DFT skeleton-Hessian setup over the grid (hess_rks.rs:71-565; hess_uks.rs:56-384), nspin 1
  // (RKS, dm0 = the closed-shell density) or 2 (UKS, one density per spin).  The AO tensor needs
  // derivatives to order 2 (RHO, 10 components) or 3 (SIGMA / TAU, 20 components).
  // setup_rho: rho [ng, nvar, nspin] for the caller's functional evaluation (get_rho_vxc_fxc).
  void setup_rho(int den, int nspin, const double* const* dm0, double* rho);
  struct SetupOut {
    double* de_fxc = nullptr;                           // [3, 3, natm, natm]
    double* de_vxc_diag[2] = {nullptr, nullptr};        // [3, 3, natm, natm] per spin
    double* de_vxc_off[2] = {nullptr, nullptr};         // [3, 3, natm, natm] per spin
    double* vmat_ip[2] = {nullptr, nullptr};            // [nao, nao, 3] per spin
    double* vmat_deriv1[2] = {nullptr, nullptr};        // [nao, nao, 3, natm] per spin
  };
  // aoslice: host [natm, 2] AO ranges (p0, p1) of the atoms in the list; vxc [ng, nvar, nspin],
  // fxc [ng, nvar, nspin, nvar, nspin] (the transformed kernels, weights not applied).
  void hessian_setup(int den, int nspin, const double* const* dm0, int64_t natm, const int64_t* aoslice,
                     const double* vxc, const double* fxc, const SetupOut& out);

  // The setup with the Becke grid shift (hess_rks_becke.rs:916-1266, hess_uks_becke.rs:501-930):
  // the grid is grouped by generating atom (split host [natm + 1]), every point moves with its
  // atom and its weight w = wquad P_A / Z follows the atoms.  Needs the full atom list.
  struct BeckeIn {
    Becke* becke = nullptr;          // tables of the natm atoms, on this handle's stream
    const double* coords = nullptr;  // device [ng, 3]
    const double* wquad = nullptr;   // device [ng], before the partition
    const int64_t* split = nullptr;  // host [natm + 1]
    int hardness = 3;
    const double* exc = nullptr;     // device [ng], energy per particle (energy density exc * rho)
  };
  struct BeckeOut {
    double* de_xc_skeleton = nullptr;                  // [3, 3, natm, natm]
    double* vmat_deriv1_grid[2] = {nullptr, nullptr};  // [nao, nao, 3, natm] per spin
  };
  // `out` (the grid-fixed keys) is optional per field.
  void hessian_setup_becke(int den, int nspin, const double* const* dm0, int64_t natm, const int64_t* aoslice,
                           const double* vxc, const double* fxc, const BeckeIn& bin, const SetupOut& out,
                           const BeckeOut& bout);

  int64_t ng() const { return ng_; }
  int64_t nao() const { return nao_; }
  hipStream_t stream() const { return stream_; }
  int64_t batch = 0;  // grid points per batch (0 = sized from memory)

 private:
  int64_t pick_batch(int64_t ncol) const;
  int64_t batch_for(int64_t ncol, bool bf) const;
  // BF16X3 grid GEMMs on the tiled NT kernel (FP32 intermediates)
  void ensure_bf16();
  void split_kets(const double* B, int64_t ncol, DevBuf& hi, DevBuf& lo);
  void kets_bf16(int nc, int64_t g0, int64_t gb, const unsigned short* bh, const unsigned short* bl, int64_t ncol,
                 float* out);
  void s_tiled(int nc, int64_t gb, int64_t ncol, DevBuf& hi, DevBuf& lo, unsigned short** h, unsigned short** l);
  void back_bf16(int nc, int64_t g0, int64_t gb, const unsigned short* sh, const unsigned short* sl, int64_t ncol,
                 double* out, double alpha, bool accumulate);
  int64_t ng_, nao_, ncomp_;
  hipStream_t stream_;
  DevBuf ao_, w_, aob_, akb_, aob2_, akb2_, rho_, ws_, ws2_;
  bool bf_ready_ = false;
  int64_t ngp_ = 0, naop_ = 0;
  DevBuf aok_hi_, aok_lo_, aot_hi_, aot_lo_, kh_, kl_, kh2_, kl2_, sh_, sl_, sh2_, sl2_;
};

// LibXC-variable derivatives -> rho_t-based effective kernels, unpolarised (reference:
// src/xceff/xc_deriv.rs:305-428, transform_xc_inner).  rho [ng, >= nvar], xc_val [ng, ncol]
// with columns [exc | 1st (xlen) | 2nd (xlen(xlen+1)/2)], xlen = 2 (SIGMA) / 3 (TAU).
// order 1: out [ng, nvar];  order 2: out [ng, nvar, nvar].  RHO: out = the raw derivative block.
void xceff_transform(int den, int order, const double* rho, int64_t ld_rho, const double* xc_val, int64_t ld_xc,
                     int64_t ng, double* out, hipStream_t s);

// Spin-polarised (xc_deriv.rs:188-243 vxc_unfold_sigma_spin1, 379-424): rho element (g, x, s) at
// g + ld_rho x + ld_spin s; xc_val columns [exc | 1st (xlen) | 2nd (xlen(xlen+1)/2)] in XCFun order
// over (rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb[, tau_a, tau_b]), xlen 2 / 5 / 7.
// order 1: out [ng, nvar, 2];  order 2: out [ng, nvar, 2, nvar, 2].
void xceff_transform_pol(int den, int order, const double* rho, int64_t ld_rho, int64_t ld_spin, const double* xc_val,
                         int64_t ld_xc, int64_t ng, double* out, hipStream_t s);

// Columns of LibXC's combined polarised output [zk | 1st | 2nd] in XCFun order (deriv <= 2;
// xc_deriv.rs:17-88 libxc_to_xcfun_mapping): LibXC groups derivatives by variable kind (rho 2,
// sigma 3, tau 2) and second-derivative blocks by kind pairs (rho,rho), (rho,sigma), (rho,tau),
// (sigma,sigma), (sigma,tau), (tau,tau), row-major inside a block (upper triangle on the diagonal
// blocks); XCFun lists all pairs of the flat variable list as combinations with replacement.
std::vector<int64_t> libxc_to_xcfun_map(int den, int deriv);

}  // namespace hh
