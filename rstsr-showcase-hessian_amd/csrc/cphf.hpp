// Dimensionless CPHF driver for restricted SCF (reference: src/hessian/rscf.rs:104-425).
#pragma once
#include <vector>

#include "krylov.hpp"

namespace hh {

// An electron-interaction component that supplies get_response_bra (the reference's
// RHessElecInteractAPI::get_response_bra, trait_rhess.rs:81-190): out[nao, nocc, nprop]
// (+)= response of bra[nao, nocc, nprop].
struct ElecResponse {
  virtual ~ElecResponse() = default;
  virtual void response_bra(const double* bra, int64_t nprop, double* out, bool accumulate, Prec prec) = 0;
};

struct CphfConfig {
  double level_shift = 0.0;
  double tol = 1e-8;
  int max_cycle = 42;
  int max_space = 14;
  double lindep = 1e-14;
};

struct CphfStats {
  KrylovStats krylov;       // summed over refinement sweeps
  int refinements = 0;      // FP64 residual evaluations after the first mixed solve
  double residual = 0.0;    // final max column 2-norm of rhs - (x + A x), FP64
};

class Cphf {
 public:
  Cphf(const double* mo_coeff, int64_t nao, int64_t nmo, const double* mo_occ, const double* mo_energy,
       const CphfConfig& cfg, hipStream_t s);
  void add_component(ElecResponse* c) { comps_.push_back(c); }

  int64_t nocc() const { return nocc_; }
  int64_t nmo() const { return nmo_; }
  int64_t nao() const { return nao_; }
  const double* mo_coeff() const { return c_.as<double>(); }

  // rscf.rs:170-178
  void dimless_rhs(const double* f1mo, const double* s1mo, int64_t nprop, double* rhs);
  // rscf.rs:210-220
  void response_mo(const double* mo1, int64_t nprop, double* resp, Prec prec);
  // rscf.rs:237-263
  void response_dimless(const double* mo1, int64_t nprop, double* resp, Prec prec);
  // rscf.rs:278-300.  prec != FP64: the Krylov solve runs on the mixed-precision response and
  // is wrapped in FP64 iterative refinement until the FP64 residual satisfies the tolerance.
  // `diag` (optional, device [nmo, nocc]): a diagonal preconditioner applied as
  // x = diag * y, i.e. the solve runs on (I + A) diag y = b (vir rows; occ rows keep 1).
  void solve_dimless(const double* rhs, int64_t nprop, double* mo1, Prec prec, CphfStats* st,
                     const double* diag = nullptr);
  // rscf.rs:328-353
  void finalize(const double* f1mo, const double* s1mo, const double* mo1, int64_t nprop, double* mo1_out,
                double* mo_e1);
  // rscf.rs:371-399: de [3, 3, natm, natm]; inputs [nmo|nocc, nocc, 3, natm]
  void cphf_hess(const double* f1mo, const double* s1mo, const double* mo1, const double* mo_e1, int64_t natm,
                 double* de);

  hipStream_t stream() const { return stream_; }
  bool verbose = false;  // per-sweep diagnostics on stderr (the reference prints per cycle)

  // Learned-preconditioner support (virtual rows only; N = nvir * nocc samples, sample index
  // a + nvir * i).  Features x [N, 3] (float, column-major), standardised over the samples:
  // ln(e_a - e_i), ln(pair[a,i]) (RI pair norm (ai|ai)), ln(rms_p |r[a,i,p]|).
  void precond_features(const double* pair, const double* r, int64_t nprop, float* x);
  // Least-squares diagonal mapping residual r to solution x over the perturbations:
  // D*[a,i] = sum_p r x / sum_p r^2; target t = ln(clip(D*, 0.05, 20)), weight w = sum_p r^2 (normalised).
  void precond_target(const double* r, const double* xs, int64_t nprop, float* t, float* w);
  // diag [nmo, nocc] (double) = exp(clip(y, -3, 3)) on virtual rows, 1 on occupied rows.
  void precond_diag(const float* y, double* diag);

 private:
  int64_t nao_, nmo_, nocc_;
  CphfConfig cfg_;
  hipStream_t stream_;
  std::vector<ElecResponse*> comps_;
  DevBuf c_, eocc_, evir_, mocc_;
  DevBuf ubra_, rbra_, tmp_, ws_, res_, g1_, g2_, g3_, sx_, ydiag_;
};

}  // namespace hh
