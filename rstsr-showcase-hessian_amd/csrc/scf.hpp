// Restricted Hartree-Fock with density-fitted Coulomb and exchange on the device: the
// mo_coeff / mo_occ / mo_energy the reference's Hessian reads from a converged SCF
// (RHessSCF::new, src/hessian/rscf.rs:30-95; its tests load them from PySCF, tests/hessian_rhf/
// mod.rs:39-41).  The RI metric and the 3-centre integrals are those of the RI-JK Hessian
// component (RHessRIJK::new_without_cderi, hess_r.rs:1751-1773), so the analytic Hessian is the
// second derivative of exactly this energy.
#pragma once
#include "cint.hpp"

namespace hh {

struct ScfResult {
  double energy = 0.0;      // total energy (electronic + nuclear repulsion)
  double e_nuc = 0.0;       // nuclear repulsion
  double last_delta = 0.0;  // |E_k - E_{k-1}| of the last cycle
  double last_error = 0.0;  // max |X^T (F D S - S D F) X| of the last cycle
  int cycles = 0;
  bool converged = false;
};

class Numint;
// Kohn-Sham terms of a restricted SCF: the functional list (LibXC ids with scales, evaluated by
// xc_eval_eff on the grid of `ni`, whose AO values carry first derivatives for GGA / meta-GGA) and
// the fraction of exact exchange exx that scales K (1 for HF).
struct ScfXc {
  Numint* ni = nullptr;
  int den = 0;  // HH_DEN_RHO / SIGMA / TAU: the strictest family of the list
  int nfunc = 0;
  const int* ids = nullptr;
  const double* scales = nullptr;
  double exx = 1.0;
};

// Closed shell, nocc doubly occupied orbitals (RHF, or RKS with `xc`); converged when
// |dE| < conv_tol and the commutator error < sqrt(conv_tol) (DIIS over up to 8 Fock matrices,
// core-Hamiltonian guess).  mo_coeff [nao, nao] and mo_energy [nao] on the device (orbital
// energies ascending).  Throws Error(HH_ERR_INVALID) for a linearly dependent basis (overlap
// eigenvalue below 1e-10 of the largest) and Error(HH_ERR_NOT_CONVERGED) if max_cycle is reached.
ScfResult rhf_scf(const Cint& mol, const Cint& aux, int nocc, double conv_tol, int max_cycle, double* mo_coeff,
                  double* mo_energy, hipStream_t s, const ScfXc* xc = nullptr);
// Unrestricted (UHF, or UKS with `xc`): nalpha / nbeta singly occupied orbitals per channel,
// mo_coeff[s] [nao, nao] and mo_energy[s] [nao] per channel on the device (the reference's
// UHessSCF inputs, uscf.rs:32-77); DIIS over the stacked channels.
ScfResult uhf_scf(const Cint& mol, const Cint& aux, int nalpha, int nbeta, double conv_tol, int max_cycle,
                  double* const* mo_coeff, double* const* mo_energy, hipStream_t s, const ScfXc* xc = nullptr);

}  // namespace hh
