// Dimensionless CPHF driver for restricted (one spin) and unrestricted (two spins) SCF
// (reference: src/hessian/rscf.rs:104-425, src/hessian/uscf.rs:78-446).
#pragma once
#include <vector>

#include "krylov.hpp"

namespace hh {

// An electron-interaction component that supplies get_response_bra (the reference's
// RHessElecInteractAPI / UHessElecInteractAPI::get_response_bra, trait_rhess.rs:81-190,
// trait_uhess.rs:57-172): out[s][nao, nocc_s, nprop] (+)= response of bra[s][nao, nocc_s, nprop]
// for every spin s < nspin().
struct ElecResponse {
  virtual ~ElecResponse() = default;
  virtual int nspin() const = 0;
  virtual void response_bra(const double* const* bra, int64_t nprop, double* const* out, bool accumulate,
                            Prec prec) = 0;
  virtual hipStream_t stream() const = 0;  // the stream its work is enqueued on
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
  int fallbacks = 0;        // 1 if a stagnating mixed solve was replaced by the plain FP64 solve
};

class Cphf {
 public:
  // mo_coeff[s] device [nao, nmo_s]; mo_occ[s], mo_energy[s] host [nmo_s] (occupied first)
  Cphf(int nspin, const double* const* mo_coeff, int64_t nao, const int64_t* nmo, const double* const* mo_occ,
       const double* const* mo_energy, const CphfConfig& cfg, hipStream_t s);
  void add_component(ElecResponse* c);

  int nspin() const { return nspin_; }
  int64_t nocc(int s = 0) const { return sp_[s].nocc; }
  int64_t nmo(int s = 0) const { return sp_[s].nmo; }
  int64_t nao() const { return nao_; }
  const double* mo_coeff(int s = 0) const { return sp_[s].c.as<double>(); }
  // orbital energies of spin s on the device: occupied [nocc_s], virtual [nmo_s - nocc_s]
  const double* eocc_dev(int s = 0) const { return sp_[s].eocc.as<double>(); }
  const double* evir_dev(int s = 0) const { return sp_[s].evir.as<double>(); }
  // rows of the stacked vector (sum_s nmo_s nocc_s) and the row offset of spin s's block
  int64_t stacked_rows() const { return n_; }
  int64_t spin_offset(int s) const { return sp_[s].off; }
  const CphfConfig& config() const { return cfg_; }
  size_t ncomponents() const { return comps_.size(); }

  // Per-spin arrays [nmo_s, nocc_s, nprop] (mo_e1: [nocc_s, nocc_s, nprop]); one pointer per spin.
  // rscf.rs:170-178 / uscf.rs:165-174
  void dimless_rhs(const double* const* f1mo, const double* const* s1mo, int64_t nprop, double* const* rhs);
  // rscf.rs:210-220 / uscf.rs:203-218
  void response_mo(const double* const* mo1, int64_t nprop, double* const* resp, Prec prec);
  // rscf.rs:237-263 / uscf.rs:229-265
  void response_dimless(const double* const* mo1, int64_t nprop, double* const* resp, Prec prec);
  // rscf.rs:278-300 / uscf.rs:278-349 (the spin blocks stacked per column, as the reference's
  // pack_flattened).  prec != FP64: Krylov on the mixed-precision response inside FP64 iterative
  // refinement until the FP64 residual satisfies the tolerance.  `diag` (optional, one spin only,
  // device [nmo, nocc]): right diagonal preconditioner x = diag * y on the virtual rows.
  void solve_dimless(const double* const* rhs, int64_t nprop, double* const* mo1, Prec prec, CphfStats* st,
                     const double* diag = nullptr);
  // rscf.rs:328-353 / uscf.rs:369-407
  void finalize(const double* const* f1mo, const double* const* s1mo, const double* const* mo1, int64_t nprop,
                double* const* mo1_out, double* const* mo_e1);
  // rscf.rs:371-399 / uscf.rs:409-440: de [3, 3, natm, natm]; inputs [.., .., 3, natm] per spin
  void cphf_hess(const double* const* f1mo, const double* const* s1mo, const double* const* mo1,
                 const double* const* mo_e1, int64_t natm, double* de);

  hipStream_t stream() const { return stream_; }
  bool verbose = false;  // per-sweep diagnostics on stderr (the reference prints per cycle)
  // Replay the response operator from HIP graphs: an operator call with the same operands (input
  // and output addresses, width, precision) as an earlier one is captured once and replayed
  // afterwards (one graph launch instead of its tens of kernel launches).  Used when the driver
  // runs on a created stream (not the legacy default stream) and every component enqueues on it;
  // graphs are dropped whenever a library buffer is freed, a component is added or a component's
  // state changes (response_state_epoch).  Counters: graphs captured and replayed.
  bool use_graphs = true;
  int64_t graph_captures() const { return n_captures_; }
  int64_t graph_launches() const { return n_launches_; }
  // finalize() calls that took the response of mo1 from the last mixed-precision solve: its final
  // residual holds the response of the returned mo1 (FP64, up to corrections whose response was
  // taken in the reduced precision with an error below 5 % of tol), which finalize needs again.
  // Reused when the mo1 given to finalize is bit-identical to that solve's output and no
  // component was added or re-prepared since.
  int64_t finalize_reuses() const { return n_fin_reuse_; }
  ~Cphf();

  // Learned-preconditioner support (one spin; virtual rows only; N = nvir * nocc samples, sample
  // index a + nvir * i).  Features x [N, 3] (float, column-major), standardised over the samples:
  // ln(e_a - e_i), ln(pair[a,i]) (RI pair norm (ai|ai)), ln(rms_p |r[a,i,p]|).
  void precond_features(const double* pair, const double* r, int64_t nprop, float* x);
  // Least-squares diagonal mapping residual r to solution x over the perturbations:
  // D*[a,i] = sum_p r x / sum_p r^2; target t = ln(clip(D*, 0.05, 20)), weight w = sum_p r^2 (normalised).
  void precond_target(const double* r, const double* xs, int64_t nprop, float* t, float* w);
  // diag [nmo, nocc] (double) = exp(clip(y, -3, 3)) on virtual rows, 1 on occupied rows.
  void precond_diag(const float* y, double* diag);

 private:
  struct Spin {
    int64_t nmo = 0, nocc = 0, off = 0;  // off: row offset of the spin block in the stacked vector
    DevBuf c, eocc, evir, ubra, rbra;
  };
  // The stacked vector [n, nprop], n = sum_s nmo_s nocc_s (one spin: the [nmo, nocc, nprop] array).
  void response_mo_stacked(const double* x, int64_t nprop, double* resp, Prec prec);
  void response_dimless_stacked(const double* x, int64_t nprop, double* resp, Prec prec);
  void dimless_scale(const double* x, int64_t nprop, double* resp);  // in place, per spin
  void solve_stacked(const double* rhs, int64_t nprop, double* x, Prec prec, CphfStats* st, const double* diag);
  const double* stack_in(const double* const* v, int64_t nprop, DevBuf& buf);
  double* stack_out(double* const* v, int64_t nprop, DevBuf& buf);
  void unstack(const double* flat, int64_t nprop, double* const* v);

  // the response operator (x -> resp) through the graph cache
  void aop(const double* x, int64_t nprop, double* resp, Prec prec, const double* diag, double* ytmp);
  void aop_body(const double* x, int64_t nprop, double* resp, Prec prec, const double* diag, double* ytmp);
  struct AopGraph {
    const double* x;
    double* resp;
    int64_t nprop;
    int prec;
    const double* diag;
    double* ytmp;
    int uses;
    hipGraphExec_t exec;
  };
  std::vector<AopGraph> graphs_;
  uint64_t graph_gen_ = 0, graph_state_ = 0;
  int64_t n_captures_ = 0, n_launches_ = 0;
  void drop_graphs();
  DevBuf fin_x_, fin_r_, fin_flag_;
  int64_t fin_nprop_ = -1, n_fin_reuse_ = 0;
  uint64_t fin_state_ = 0;  // component-state count when fin_r_ was taken

  int nspin_;
  int64_t nao_, n_ = 0;
  Spin sp_[2];
  CphfConfig cfg_;
  hipStream_t stream_;
  std::vector<ElecResponse*> comps_;
  DevBuf tmp_, ws_, res_, g1_, g2_, g3_, sx_, ydiag_, st_in_, st_out_, st_aux_;
};

}  // namespace hh
