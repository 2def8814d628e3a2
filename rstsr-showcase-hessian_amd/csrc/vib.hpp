// Harmonic vibrational analysis of a Cartesian Hessian (host FP64; reference:
// src/hessian/vib.rs).  The Hessian is a [3 natm, 3 natm] matrix: the readout is a few
// milliseconds of CPU work at any molecular size this library treats, so it runs on the host
// next to the caller's copy of the Hessian.
#pragma once
#include <cstdint>
#include <vector>

namespace hh {
namespace vib {

enum Rotor : int { ROTOR_ATOM = 0, ROTOR_LINEAR = 1, ROTOR_REGULAR = 2 };
enum Trv : int { TRV_V = 0, TRV_TR = 1, TRV_NONE = 2 };

// Symmetric eigendecomposition by cyclic Jacobi rotations: a [n, n] column-major (copied),
// eigenvalues ascending in w, eigenvectors as the columns of v.
void eigh(const double* a, int n, std::vector<double>& w, std::vector<double>& v);

// Thin SVD of a [m, k] column-major (m >= k) by one-sided Jacobi: singular values descending in
// s, left singular vectors as the columns of u [m, k].
void svd_thin(const double* a, int m, int k, std::vector<double>& u, std::vector<double>& s);

// Translation (bit 1) / rotation (bit 2) basis [3 natm, n] (vib.rs:59-140).  nrt >= 0 keeps the
// leading nrt singular vectors; nrt < 0 keeps those above 3 natm * max(s) * eps.
std::vector<double> tr_space(const double* mass, const double* geom, int natm, int space, int nrt, int* ncol);

// Rotational constants [3] of a mass-centred geometry (vib.rs:152-197); unit 0 GHz, 1 cm^-1.
void rotation_const(const double* mass, const double* geom_c, int natm, int unit, double* out);
int rotor_type(const double* rc_ghz);

struct Result {
  int ndof = 0;
  std::vector<double> omega, q, w, x, mu, k, dq0, qtp0, xtp0, theta_vib;
  std::vector<int> imag, trv;
  std::vector<int64_t> degeneracy;
};

// vib.rs:336-552.  hess [3 natm, 3 natm] (E_h / a0^2), geom [3, natm] (a0), mass [natm] (u).
Result harmonic_analysis(const double* hess, const double* geom, const double* mass, int natm, bool project_trans,
                         bool project_rot);

}  // namespace vib
}  // namespace hh
