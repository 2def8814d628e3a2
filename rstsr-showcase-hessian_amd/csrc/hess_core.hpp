// The one-electron parts of the analytic Hessian on the device: nuclear repulsion
// (get_nuc_repl_hess, src/hessian/nuc_repl.rs:14-58), the core-Hamiltonian skeleton term
// (get_hess_hcore with generator_hcore_deriv2, hcore.rs:15-125), the overlap term
// (get_hess_ovlp, ovlp.rs:24-52), the first-derivative matrices the CPHF right-hand side takes
// (generator_hcore_deriv1, hcore.rs:140-177; generator_ovlp_deriv1, ovlp.rs:67-82) and the
// (energy-weighted) density matrices.  All outputs column-major on the device:
// Hessians [3, 3, natm, natm], first derivatives [nao, nao, 3, natm].
#pragma once
#include "cint.hpp"

namespace hh {

void nuc_repl_hess(const Cint& mol, double* out, hipStream_t s);
// dm0 [nao, nao] symmetric
void hcore_hess(const Cint& mol, const double* dm0, double* out, hipStream_t s);
// dme0 [nao, nao] symmetric (energy-weighted density)
void ovlp_hess(const Cint& mol, const double* dme0, double* out, hipStream_t s);
void hcore_deriv1(const Cint& mol, double* out, hipStream_t s);
void ovlp_deriv1(const Cint& mol, double* out, hipStream_t s);
// out [nao, nao] = sum_i w[i] C[:, i] C[:, i]^T over the nmo columns of C [nao, nmo]; w host
void weighted_density(const double* C, int64_t nao, int64_t nmo, const double* w, double* out, hipStream_t s);

}  // namespace hh
