// Exchange-correlation functionals with their first and second derivatives on the device: the
// values LibXC supplies to the reference (xc_func.compute_xc in libxc_eval_inner,
// src/xceff/libxc_wrap.rs:32-82), in LibXC's output layout, for the functionals of the
// reference's tests and configurations: LDA_X (Slater), LDA_C_VWN (VWN5), LDA_C_VWN_RPA,
// GGA_X_B88, GGA_C_LYP, the hybrid HYB_GGA_XC_B3LYP (0.08 Slater + 0.72 B88 + 0.19 VWN_RPA +
// 0.81 LYP; the 0.20 of exact exchange is the RI-K scale of the caller), the meta-GGAs
// MGGA_X_TPSS and MGGA_C_TPSS and the hybrid HYB_MGGA_XC_TPSS0 (0.75 TPSS exchange + TPSS
// correlation; 0.25 exact exchange).  Functional ids are LibXC's.
#pragma once
#include "hh_common.hpp"

namespace hh {

enum XcId : int {
  XC_LDA_X = 1,
  XC_LDA_C_VWN = 7,
  XC_LDA_C_VWN_RPA = 8,
  XC_GGA_X_B88 = 106,
  XC_GGA_C_LYP = 131,
  XC_HYB_GGA_XC_B3LYP = 402,
  XC_MGGA_X_TPSS = 202,
  XC_MGGA_C_TPSS = 231,
  XC_HYB_MGGA_XC_TPSS0 = 396,
};

// true for ids this library evaluates; *family 0 (LDA: rho), 1 (GGA: rho, sigma) or 2
// (meta-GGA: rho, sigma, tau); *exx the fraction of exact exchange the caller adds (B3LYP 0.2,
// TPSS0 0.25, else 0)
bool xc_known(int id, int* family, double* exx);

// Number of output columns: [zk | first derivatives | unique second derivatives] up to `deriv`,
// in LibXC's order (unpolarised: rho, sigma, tau; polarised: rho_a, rho_b, sigma_aa, sigma_ab,
// sigma_bb, tau_a, tau_b; second derivatives v2rho2, v2rhosigma, v2rhotau, v2sigma2,
// v2sigmatau, v2tau2 as LibXC lists them, without the Laplacian columns).
__host__ __device__ int64_t xc_ncol(int family, int nspin, int deriv);

// out[g + ld_out c] (+)= coef * column c of functional `id` at point g.  rho [ng, nvar, nspin]
// with element (g, x, s) at g + ld_rho x + ld_spin s: x = 0 the density, x = 1..3 its gradient
// (GGA, meta-GGA), x = 4 the kinetic energy density tau (meta-GGA).  zk is the energy per particle (LibXC's convention); points with total density below
// 1e-15 give zero.
void xc_eval(int id, int nspin, int deriv, const double* rho, int64_t ld_rho, int64_t ld_spin, int64_t ng, double coef,
             bool accumulate, double* out, int64_t ld_out, hipStream_t s);

// libxc_eval_eff (src/xceff/libxc_wrap.rs:88-122) fused with the accumulation of
// get_rho_vxc_fxc (src/numint_matmul/hess_rks.rs:174-218, hess_uks.rs:56-108): the functional's
// energy per particle and its derivatives with respect to the density components (rho, grad rho, tau)
// of each spin, written from registers (no LibXC-variable intermediate):
//   exc[g] (+)= scale zk;  vxc[g + ng (x + nvl s)];  fxc[g + ng ((x + nvl s) + nspin nvl (y + nvl t))]
// nvl = the component count of the caller's arrays (1 RHO, 4 SIGMA, 5 TAU); a functional with
// fewer components fills the leading slice.  Null outputs are skipped.
void xc_eval_eff(int id, int nspin, int nvl, const double* rho, int64_t ld_rho, int64_t ld_spin, int64_t ng,
                 double scale, bool accumulate, double* exc, double* vxc, double* fxc, hipStream_t s);

}  // namespace hh
