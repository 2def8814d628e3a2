// RI-JK skeleton Hessian and first-derivative Fock terms on the device (reference:
// get_rijk_skeleton_decomposed_separated, src/ri_jk/hess_r.rs:202-1718, consumed by
// RHessRIJK::make_skeleton_hess / get_deriv1_bra, hess_r.rs:1858-1909, and the UHF
// counterparts, hess_u.rs:150-261).
//
// The reference splits the second derivative into ~35 keyed terms.  Here the same quantity is
// organised as the second derivative of E = 1/2 sum_n a_n^T V^-1 a_n (a_n[P] = (P|Pi_n), one
// "density" Pi = D for J, the occupied pairs Pi_ij = c_i c_j^T for K):
//   H = sum_n [ (d^2 a_n) . d_n  -  1/2 d_n^T (d^2 V) d_n  +  Z_n^T V^-1 Z_n ],
//   d_n = V^-1 a_n,  Z_n = d a_n - (d V) d_n.
// The first two terms are linear in M_P = sum_n d_n[P] Pi_n and F = sum_n d_n d_n^T, so J and K
// share one pass over the second-derivative integrals; the third is a product of per-pair
// derivative vectors (matrix cores).  The 3-centre derivative integrals are requested per
// auxiliary range from a caller-supplied generator (the reference's
// generator_hess_intor_j3c_by_aux, cint_handling.rs:80-92), one kind at a time.
#pragma once
#include "rijk.hpp"

namespace hh {

struct SkeletonInput {
  const double* const* mo_coeff;  // per spin, device [nao, nmo_s]
  const int64_t* nmo;
  const double* const* mo_occ;    // per spin, host [nmo_s]
  const double* j2c;              // device [naux, naux] or null (use the handle's factor)
  const double* j2c_ip1;          // device [naux, naux, 3]
  const double* j2c_ipip1;        // device [naux, naux, 3, 3]
  const double* j2c_ip1ip2;       // device [naux, naux, 3, 3]
  hh_int3c_deriv_fn int3c;
  void* ctx;
  int64_t nrange;
  const int64_t* aux_ranges;      // host [nrange, 2]
  int64_t natm;
  const int64_t* aoslice;         // host [natm, 2]
  const int64_t* auxslice;        // host [natm, 2]
};

// de [3, 3, natm, natm] = scale_j de_J - f scale_k sum_s de_K,s  (f = 1/2 RHF, 1 UHF);
// deriv1_bra[s] [nao, nocc_s, 3, natm] = scale_j j1ao mocc_s - f scale_k k1ao_s mocc_s
// (deriv1_bra may be null to skip the first-derivative terms).
void rijk_skeleton(Rijk& r, const SkeletonInput& in, double* de, double* const* deriv1_bra);

}  // namespace hh
