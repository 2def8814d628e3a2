// Gaussian-basis integral and AO-evaluation engine on the device (the parts of libcint the
// reference calls on its Hessian path: eval_gto, nimatmul.rs:94-99; int2c2e / int3c2e and their
// ip1 / ip2 / ipip1 / ipvip1 / ip1ip2 / ipip2 derivatives, cint_handling.rs:20-92,
// hess_r.rs:416, 537-540, 788-789, 1058, 1356, 1731).
//
// The molecule is given in libcint's own format (atm [natm][6], bas [nbas][8], env), so a caller
// of libcint passes the same arrays.  General contractions (nctr > 1) are split into one
// contracted shell per contraction, which keeps libcint's AO order (components fastest, then
// contractions, then shells).  Spherical functions follow libcint: real solid harmonics
// normalised on the unit sphere, m = -l .. l (p as x, y, z), times the radial function with the
// env coefficients as given; Cartesian functions carry the s / p factors of libcint
// (1 / (2 sqrt(pi)), sqrt(3 / (4 pi))) and none for l >= 2.
#pragma once
#include <memory>
#include <vector>

#include "hh_common.hpp"

namespace hh {

struct CShell {
  double x, y, z;  // centre
  int l;
  int nprim;
  int prim_off;  // exponents at prim[prim_off + p], coefficients at prim[prim_off + nprim + p]
  int ao0;       // first AO of this contraction
};

class Cint {
 public:
  static constexpr int LMAX = 5;  // highest angular momentum of a shell
  Cint(const int32_t* atm, int natm, const int32_t* bas, int nbas, const double* env, int64_t nenv, bool cart);
  int64_t nao() const { return nao_; }
  int nbas() const { return (int)bas_ao_.size() - 1; }
  bool cart() const { return cart_; }
  // AO offset of libcint shell i (i <= nbas: nao)
  int64_t ao_loc(int i) const { return bas_ao_[i]; }

  // out [ng, nao, ncomp] (column-major, ng fastest), ncomp = (deriv+1)(deriv+2)(deriv+3)/6 in
  // libcint's component order; coords device [ng][3].  (GTOval_sph / GTOval_cart deriv 0-3)
  void eval_gto(int deriv, const double* coords, int64_t ng, double* out, hipStream_t s) const;

  // Integral kinds (HH_INT_* in hesship.h).  int2c2e of this basis over shells [sh0, sh1) x all
  // (rows of the range): out [n01, nao, comp]; int3c2e over all AO pairs of `mol` and the
  // auxiliary shells [sh0, sh1) of `aux`: out [nao, nao, n01, comp] (the reference's
  // generator_hess_intor_j3c_by_aux layout, cint_handling.rs:80-92, comp 9 as [3, 3] with the
  // later function's direction fastest).
  void int2c2e(int kind, int sh0, int sh1, double* out, hipStream_t s) const;
  static void int3c2e(const Cint& mol, const Cint& aux, int kind, int sh0, int sh1, double* out, hipStream_t s);
  // int3c2e with "s2ij" symmetry: the upper triangle (mu <= nu) packed, out [nao_tp, n01]
  // (nao_tp = nao (nao + 1) / 2, index nu (nu + 1) / 2 + mu): the j3c of the metric
  // decomposition (hess_r.rs:1731).
  static void int3c2e_s2ij(const Cint& mol, const Cint& aux, int sh0, int sh1, double* out, hipStream_t s);

  // One-electron integrals out [nao, nao, comp] (libcint's int1e_ovlp / int1e_kin and their
  // ip / ipip / ipvip derivatives: comp 3 for IP1 (on a), 9 = 3 t + s for IPIP1 (t, s on a) and
  // IPVIP1 (t on a, s on b)); op HH_INT1E_OVLP or HH_INT1E_KIN.
  void int1e(int op, int kind, double* out, hipStream_t s) const;
  // Nuclear attraction: atom = -1: sum_C -Z_C <a| 1 / |r - R_C| |b> (int1e_nuc and its ip / ipip /
  // ipvip derivatives on the functions); atom >= 0: <a| 1 / |r - R_atom| |b> (int1e_rinv with the
  // origin at that nucleus, no charge), as 3-centre Coulomb integrals against unit point charges.
  void int1e_nuc(int kind, int atom, double* out, hipStream_t s) const;
  // <a| 1 / |r - R_C| |b> for every nucleus C: out [nao, nao, natm, comp]
  void int1e_rinv_all(int kind, double* out, hipStream_t s) const;
  // AO range [ao0, ao1) of atom a (aoslice_by_atom; shells must be grouped by atom)
  int64_t atom_ao0(int a) const { return atom_ao_[2 * a]; }
  int64_t atom_ao1(int a) const { return atom_ao_[2 * a + 1]; }
  bool atoms_contiguous() const { return atom_contiguous_; }
  int natm() const { return (int)atom_z_.size(); }
  double atom_charge(int a) const { return atom_z_[a]; }
  const double* atom_xyz(int a) const { return &atom_xyz_[3 * a]; }

  const std::vector<CShell>& cshells() const { return cs_; }

 private:
  const Cint& point_charges() const;
  std::vector<double> atom_z_, atom_xyz_;
  std::vector<int64_t> atom_ao_;
  bool atom_contiguous_ = false;
  mutable std::unique_ptr<Cint> points_;
  std::vector<CShell> cs_;          // contracted shells (host copy)
  std::vector<int> bas_cs_;         // first contracted shell of libcint shell i (i <= nbas)
  std::vector<int64_t> bas_ao_;     // AO offset of libcint shell i
  std::vector<double> prim_h_;
  DevBuf d_cs_, d_prim_;
  int64_t nao_ = 0;
  bool cart_ = false;
  int lmax_ = 0;
};

}  // namespace hh
