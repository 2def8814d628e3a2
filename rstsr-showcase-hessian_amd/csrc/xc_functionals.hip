// Exchange-correlation functionals on the device (see xc_functionals.hpp).
//
// Each functional is written once, as its energy density e(rho_a, rho_b, sigma_aa, sigma_ab,
// sigma_bb, tau_a, tau_b), over a second-order jet type: a value with its gradient and Hessian
// with respect to the N input variables, propagated through every arithmetic operation
// (forward-mode differentiation).  The unpolarised case evaluates the same expressions at
// rho_a = rho_b = rho/2, sigma_ss' = sigma/4, tau_s = tau/2 with jets in (rho, sigma, tau), so its
// derivatives are those LibXC reports for the total density.  One thread per grid point, FP64.
//
// Formulas (published forms):
//   Slater exchange      e = -3/4 (6/pi)^(1/3) sum_s rho_s^(4/3)
//   B88 (Becke 1988)     e = sum_s [-3/4 (6/pi)^(1/3) rho_s^(4/3) - beta rho_s^(4/3) x_s^2 / (1 + 6 beta x_s asinh x_s)],
//                        x_s = |grad rho_s| / rho_s^(4/3), beta = 0.0042
//   VWN (Vosko, Wilk, Nusair 1980) with x = sqrt(rs), X(x) = x^2 + b x + c, Q = sqrt(4c - b^2):
//     F(A, b, c, x0) = A [ln(x^2/X) + 2b/Q atan(Q/(2x+b)) - b x0/X(x0) (ln((x-x0)^2/X) + 2(b+2x0)/Q atan(Q/(2x+b)))]
//     eps = F_P + F_a f(z)/f''(0) (1 - z^4) + (F_F - F_P) f(z) z^4, f(z) = ((1+z)^(4/3) + (1-z)^(4/3) - 2)/(2^(4/3) - 2)
//     with the fit-V parameters (LDA_C_VWN) or the RPA ones (LDA_C_VWN_RPA) for the paramagnetic,
//     ferromagnetic and spin-stiffness curves
//   LYP in the form of Miehlich, Savin, Stoll and Preuss (1989), no density Laplacian:
//     e = -4a/(1 + d r^-1/3) ra rb / r - a b w {ra rb [2^(11/3) C_F (ra^(8/3) + rb^(8/3))
//         + (47/18 - 7 delta/18) |grad r|^2 - (5/2 - delta/18)(saa + sbb) - (delta - 11)/9 (ra/r saa + rb/r sbb)]
//         - 2/3 r^2 |grad r|^2 + (2/3 r^2 - ra^2) sbb + (2/3 r^2 - rb^2) saa},
//     w = exp(-c r^-1/3) / (1 + d r^-1/3) r^(-11/3), delta = c r^-1/3 + d r^-1/3 / (1 + d r^-1/3),
//     a = 0.04918, b = 0.132, c = 0.2533, d = 0.349, C_F = 3/10 (3 pi^2)^(2/3)
//   TPSS (Tao, Perdew, Staroverov, Scuseria 2003), exchange by spin scaling
//     E_x[rho_a, rho_b] = sum_s e_x(2 rho_s, 4 sigma_ss, 2 tau_s) / 2, e_x(n, s, tau) = e_x^unif F_x,
//     p = s / (4 (3 pi^2)^(2/3) n^(8/3)), z = tau_W / tau (tau_W = s / 8n; tau below tau_W is
//     taken as tau_W), alpha = (tau - tau_W) / tau_unif, q_b = 9/20 (alpha - 1) /
//     sqrt(1 + b alpha (alpha - 1)) + 2p/3,
//     x = {[10/81 + c z^2 / (1 + z^2)^2] p + 146/2025 q_b^2 - 73/405 q_b p sqrt(1/2 (3r/5)^2 + 1/2)
//          + (10/81)^2 p^2 / kappa + 2 sqrt(e) 10/81 (3z/5)^2 + e mu p^3} / (1 + sqrt(e) p)^2
//     (r = z / p = (3 pi^2)^(2/3) n^(5/3) / 2 tau), F_x = 1 + kappa - kappa / (1 + x / kappa);
//     kappa 0.804, b 0.40, c 1.59096, e 1.537, mu 0.21951
//   TPSS correlation: n eps_rev (1 + d eps_rev z^3), d = 2.8, eps_rev = eps_PBE (1 + C z^2)
//     - (1 + C) z^2 sum_s n_s / n eps_s, eps_s = max(eps_PBE(n_s, zeta = 1, sigma_ss), eps_PBE),
//     evaluated as eps_PBE (1 - z^2) + (1 + C) z^2 sum_s n_s / n (eps_PBE - eps_s),
//     C = (0.53 + 0.87 zeta^2 + 0.50 zeta^4 + 2.26 zeta^6) / {1 + xi^2 [(1+zeta)^(-4/3)
//     + (1-zeta)^(-4/3)] / 2}^4, xi^2 = |grad zeta|^2 / (4 (3 pi^2 n)^(2/3))
//   PBE correlation (Perdew, Burke, Ernzerhof 1996): eps_PW92 + gamma phi^3 ln[1 + beta/gamma t^2
//     (1 + A t^2) / (1 + A t^2 + A^2 t^4)], A = beta/gamma / (exp(-eps_PW92 / gamma phi^3) - 1),
//     t^2 = pi sigma / (16 phi^2 k_F n^2), phi = ((1+zeta)^(2/3) + (1-zeta)^(2/3)) / 2,
//     beta 0.06672455060314922, gamma (1 - ln 2) / pi^2
//   PW92 (Perdew, Wang 1992) with LibXC's modified constants:
//     G(rs) = -2A (1 + a1 rs) ln[1 + 1 / (2A (b1 rs^(1/2) + b2 rs + b3 rs^(3/2) + b4 rs^2))],
//     eps = G_P + G_a f(zeta) (1 - zeta^4) / f''(0) + (G_F - G_P) f(zeta) zeta^4
//   TPSS0: 0.75 TPSS exchange + TPSS correlation (the 0.25 of exact exchange is the caller's)
// Thresholds: total density < 1e-15 gives zero; a spin channel below 1e-15 contributes no
// exchange; 1 +- zeta is floored at 2.2e-16 (its power then constant), as LibXC's zeta threshold.
#include "xc_functionals.hpp"

#include <cmath>

namespace hh {
namespace {

constexpr double kDensThr = 1e-15;
constexpr double kZetaThr = 2.220446049250313e-16;
// sigma floor: LibXC's sigma threshold (dens_threshold^(4/3)) squared
constexpr double kSigmaFloor = 1e-40;

__host__ __device__ constexpr int hidx(int i, int j, int n) { return i * n - i * (i - 1) / 2 + (j - i); }

template <int N>
struct Jet {
  static constexpr int NH = N * (N + 1) / 2;
  double v, g[N], h[NH];
};

template <int N>
__host__ __device__ __forceinline__ Jet<N> cst(double c) {
  Jet<N> r;
  r.v = c;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = 0.0;
#pragma unroll
  for (int k = 0; k < Jet<N>::NH; ++k) r.h[k] = 0.0;
  return r;
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> var(double x, int i) {
  Jet<N> r = cst<N>(x);
  r.g[i] = 1.0;
  return r;
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> operator+(const Jet<N>& a, const Jet<N>& b) {
  Jet<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] + b.g[i];
#pragma unroll
  for (int k = 0; k < Jet<N>::NH; ++k) r.h[k] = a.h[k] + b.h[k];
  return r;
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> operator-(const Jet<N>& a, const Jet<N>& b) {
  Jet<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.g[i] - b.g[i];
#pragma unroll
  for (int k = 0; k < Jet<N>::NH; ++k) r.h[k] = a.h[k] - b.h[k];
  return r;
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> operator*(double c, const Jet<N>& a) {
  Jet<N> r;
  r.v = c * a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = c * a.g[i];
#pragma unroll
  for (int k = 0; k < Jet<N>::NH; ++k) r.h[k] = c * a.h[k];
  return r;
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> operator+(double c, const Jet<N>& a) {
  Jet<N> r = a;
  r.v += c;
  return r;
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> operator*(const Jet<N>& a, const Jet<N>& b) {
  Jet<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = a.v * b.g[i] + b.v * a.g[i];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = i; j < N; ++j) {
      const int k = hidx(i, j, N);
      r.h[k] = a.v * b.h[k] + b.v * a.h[k] + a.g[i] * b.g[j] + a.g[j] * b.g[i];
    }
  return r;
}

// f(a) from f, f', f'' at a.v (chain rule to second order)
template <int N>
__host__ __device__ __forceinline__ Jet<N> chain(const Jet<N>& a, double f0, double f1, double f2) {
  Jet<N> r;
  r.v = f0;
#pragma unroll
  for (int i = 0; i < N; ++i) r.g[i] = f1 * a.g[i];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = i; j < N; ++j) {
      const int k = hidx(i, j, N);
      r.h[k] = f1 * a.h[k] + f2 * a.g[i] * a.g[j];
    }
  return r;
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> powc(const Jet<N>& a, double p) {
  const double x = a.v, xp = pow(x, p);
  return chain(a, xp, p * xp / x, p * (p - 1.0) * xp / (x * x));
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> recip(const Jet<N>& a) {
  const double x = a.v, y = 1.0 / x;
  return chain(a, y, -y * y, 2.0 * y * y * y);
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> operator/(const Jet<N>& a, const Jet<N>& b) {
  return a * recip(b);
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> jexp(const Jet<N>& a) {
  const double e = exp(a.v);
  return chain(a, e, e, e);
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> jlog(const Jet<N>& a) {
  const double x = a.v;
  return chain(a, log(x), 1.0 / x, -1.0 / (x * x));
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> jsqrt(const Jet<N>& a) {
  const double s = sqrt(a.v);
  return chain(a, s, 0.5 / s, -0.25 / (s * a.v));
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> jatan(const Jet<N>& a) {
  const double x = a.v, d = 1.0 / (1.0 + x * x);
  return chain(a, atan(x), d, -2.0 * x * d * d);
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> jlog1p(const Jet<N>& a) {
  const double x = a.v, d = 1.0 / (1.0 + x);
  return chain(a, log1p(x), d, -d * d);
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> jexpm1(const Jet<N>& a) {
  const double e = exp(a.v);
  return chain(a, expm1(a.v), e, e);
}

// the larger of two jets by value (the second one on a tie)
template <int N>
__host__ __device__ __forceinline__ Jet<N> jmax(const Jet<N>& a, const Jet<N>& b) {
  return a.v > b.v ? a : b;
}

// g(t) = sqrt(t) asinh(sqrt(t)) (t = x^2 of B88), regular at t = 0 (series below 1e-6)
template <int N>
__host__ __device__ __forceinline__ Jet<N> xasinhx_of_sq(const Jet<N>& a) {
  const double t = a.v;
  if (t < 1e-6) return chain(a, t - t * t / 6.0 + 3.0 * t * t * t / 40.0, 1.0 - t / 3.0 + 9.0 * t * t / 40.0,
                             -1.0 / 3.0 + 9.0 * t / 20.0);
  const double u = sqrt(t), A = asinh(u), R = sqrt(1.0 + t);
  const double g = u * A, g1 = A / (2.0 * u) + 1.0 / (2.0 * R);
  const double g2 = 1.0 / (4.0 * t * R) - A / (4.0 * u * t) - 1.0 / (4.0 * R * R * R);
  return chain(a, g, g1, g2);
}

// (1 + z)^p with the zeta threshold: constant once 1 + z falls below it
template <int N>
__host__ __device__ __forceinline__ Jet<N> opzp(const Jet<N>& opz, double p) {
  if (opz.v <= kZetaThr) return cst<N>(pow(kZetaThr, p));
  return powc(opz, p);
}

template <int N>
__host__ __device__ __forceinline__ Jet<N> opz43(const Jet<N>& opz) {
  return opzp(opz, 4.0 / 3.0);
}

struct VwnSet {
  double A, b, c, x0;
};
// fit V (LDA_C_VWN) and RPA (LDA_C_VWN_RPA) parameters; k: 0 paramagnetic, 1 ferromagnetic,
// 2 spin stiffness
__host__ __device__ __forceinline__ VwnSet vwn_params(bool rpa, int k) {
  const double a_stiff = -1.0 / (6.0 * M_PI * M_PI);
  if (!rpa) {
    if (k == 0) return {0.0310907, 3.72744, 12.9352, -0.10498};
    if (k == 1) return {0.01554535, 7.06042, 18.0578, -0.32500};
    return {a_stiff, 1.13107, 13.0045, -0.0047584};
  }
  if (k == 0) return {0.0310907, 13.0720, 42.7198, -0.409286};
  if (k == 1) return {0.01554535, 20.1231, 101.578, -0.743294};
  return {a_stiff, 1.06835, 11.4813, -0.228344};
}

template <int N>
__host__ __device__ Jet<N> vwn_curve(const VwnSet& p, const Jet<N>& x, const Jet<N>& rs) {
  const double Q = sqrt(4.0 * p.c - p.b * p.b);
  const double Xx0 = p.x0 * p.x0 + p.b * p.x0 + p.c;
  const Jet<N> X = rs + p.b * x + cst<N>(p.c);
  const Jet<N> at = jatan(Q * recip(2.0 * x + cst<N>(p.b)));
  const Jet<N> xm = x + cst<N>(-p.x0);
  const Jet<N> t1 = jlog(rs / X) + (2.0 * p.b / Q) * at;
  const Jet<N> t2 = jlog(xm * xm / X) + (2.0 * (p.b + 2.0 * p.x0) / Q) * at;
  return p.A * (t1 - (p.b * p.x0 / Xx0) * t2);
}

// VWN energy density: rho eps_c(rs, zeta)
template <int N>
__host__ __device__ Jet<N> vwn(bool rpa, const Jet<N>& ra, const Jet<N>& rb) {
  const Jet<N> rho = ra + rb;
  const Jet<N> rs = cbrt(3.0 / (4.0 * M_PI)) * powc(rho, -1.0 / 3.0);
  const Jet<N> x = jsqrt(rs);
  const Jet<N> eP = vwn_curve(vwn_params(rpa, 0), x, rs), eF = vwn_curve(vwn_params(rpa, 1), x, rs),
               eA = vwn_curve(vwn_params(rpa, 2), x, rs);
  const Jet<N> z = (ra - rb) / rho;
  const double fden = pow(2.0, 4.0 / 3.0) - 2.0, fpp = 4.0 / (9.0 * (cbrt(2.0) - 1.0));
  const Jet<N> fz = (1.0 / fden) * (opz43(1.0 + z) + opz43(1.0 + (-1.0) * z) + cst<N>(-2.0));
  const Jet<N> z2 = z * z, z4 = z2 * z2;
  const Jet<N> eps = eP + (1.0 / fpp) * (eA * fz * (1.0 + (-1.0) * z4)) + (eF - eP) * fz * z4;
  return rho * eps;
}

// Slater exchange of one spin channel
template <int N>
__host__ __device__ Jet<N> slater_s(const Jet<N>& rs_) {
  const double cx = 0.75 * cbrt(6.0 / M_PI);
  return (-cx) * powc(rs_, 4.0 / 3.0);
}

// B88 exchange of one spin channel (density rho_s, sigma_ss)
template <int N>
__host__ __device__ Jet<N> b88_s(const Jet<N>& r, const Jet<N>& s) {
  const double beta = 0.0042;
  const Jet<N> r43 = powc(r, 4.0 / 3.0);
  const Jet<N> t = s * powc(r, -8.0 / 3.0);  // x_s^2
  return slater_s(r) - beta * (r43 * t * recip(1.0 + (6.0 * beta) * xasinhx_of_sq(t)));
}

template <int N>
__host__ __device__ Jet<N> lyp(const Jet<N>& ra, const Jet<N>& rb, const Jet<N>& saa, const Jet<N>& sab, const Jet<N>& sbb) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const double CF = 0.3 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
  const Jet<N> rho = ra + rb;
  const Jet<N> rm13 = powc(rho, -1.0 / 3.0);
  const Jet<N> den = 1.0 + d * rm13;
  const Jet<N> iden = recip(den);
  const Jet<N> omega = jexp((-c) * rm13) * iden * powc(rho, -11.0 / 3.0);
  const Jet<N> delta = c * rm13 + d * (rm13 * iden);
  const Jet<N> g = saa + 2.0 * sab + sbb;
  const Jet<N> rab = ra * rb;
  const Jet<N> irho = recip(rho);
  const Jet<N> t1 = (-4.0 * a) * (iden * rab * irho);
  const Jet<N> br = pow(2.0, 11.0 / 3.0) * CF * (powc(ra, 8.0 / 3.0) + powc(rb, 8.0 / 3.0)) +
                    (47.0 / 18.0 + (-7.0 / 18.0) * delta) * g - (2.5 + (-1.0 / 18.0) * delta) * (saa + sbb) -
                    (1.0 / 9.0) * ((delta + cst<N>(-11.0)) * ((ra * irho) * saa + (rb * irho) * sbb));
  const Jet<N> rho2 = rho * rho;
  const Jet<N> brace = rab * br - (2.0 / 3.0) * (rho2 * g) + ((2.0 / 3.0) * rho2 - ra * ra) * sbb +
                       ((2.0 / 3.0) * rho2 - rb * rb) * saa;
  return t1 - (a * b) * (omega * brace);
}

// PW92 G(rs) of one parameter set (A, a1, b1, b2, b3, b4)
template <int N>
__host__ __device__ Jet<N> pw92_g(const Jet<N>& rs, const double* P) {
  const Jet<N> srs = jsqrt(rs);
  const Jet<N> q = (2.0 * P[0]) * (P[2] * srs + P[3] * rs + P[4] * (rs * srs) + P[5] * (rs * rs));
  return (-2.0 * P[0]) * ((1.0 + P[1] * rs) * jlog1p(recip(q)));
}

// PW92 correlation per particle at (rs, zeta); opz, omz = 1 + zeta, 1 - zeta
template <int N>
__host__ __device__ Jet<N> pw92_eps(const Jet<N>& rs, const Jet<N>& z, const Jet<N>& opz, const Jet<N>& omz) {
  // paramagnetic, ferromagnetic, minus the spin stiffness
  const double P0[6] = {0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294};
  const double P1[6] = {0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517};
  const double P2[6] = {0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671};
  const double fpp = 1.709920934161365617563962776245;
  const Jet<N> e0 = pw92_g(rs, P0), e1 = pw92_g(rs, P1);
  const Jet<N> ac = (-1.0) * pw92_g(rs, P2);
  const double fden = pow(2.0, 4.0 / 3.0) - 2.0;
  const Jet<N> f = (1.0 / fden) * (opz43(opz) + opz43(omz) + cst<N>(-2.0));
  const Jet<N> z2 = z * z, z4 = z2 * z2;
  return e0 + (1.0 / fpp) * (ac * f * (1.0 + (-1.0) * z4)) + (e1 - e0) * f * z4;
}

// PBE correlation per particle of density n, polarisation z (opz, omz = 1 + z, 1 - z),
// |grad n|^2 = s
template <int N>
__host__ __device__ Jet<N> pbe_c_eps(const Jet<N>& n, const Jet<N>& z, const Jet<N>& opz, const Jet<N>& omz,
                                     const Jet<N>& s) {
  const double beta = 0.06672455060314922, gamma = (1.0 - M_LN2) / (M_PI * M_PI);
  const Jet<N> rs = cbrt(3.0 / (4.0 * M_PI)) * powc(n, -1.0 / 3.0);
  const Jet<N> ec = pw92_eps(rs, z, opz, omz);
  const Jet<N> phi = 0.5 * (opzp(opz, 2.0 / 3.0) + opzp(omz, 2.0 / 3.0));
  const Jet<N> kf = cbrt(3.0 * M_PI * M_PI) * powc(n, 1.0 / 3.0);
  const Jet<N> phi2 = phi * phi, phi3 = phi2 * phi;
  const Jet<N> t2 = (M_PI / 16.0) * (s * recip(phi2 * kf * (n * n)));
  const Jet<N> A = (beta / gamma) * recip(jexpm1((-1.0 / gamma) * (ec * recip(phi3))));
  const Jet<N> y = A * t2;
  const Jet<N> frac = (1.0 + y) * recip(1.0 + y + y * y);
  return ec + gamma * (phi3 * jlog1p((beta / gamma) * (t2 * frac)));
}

// TPSS exchange energy per volume of a spin-unpolarised density n (|grad n|^2 = s, tau)
template <int N>
__host__ __device__ Jet<N> tpss_x_total(const Jet<N>& n, const Jet<N>& s, const Jet<N>& tau) {
  const double kappa = 0.804, b = 0.40, c = 1.59096, e = 1.537, mu = 0.21951;
  const double K2 = pow(3.0 * M_PI * M_PI, 2.0 / 3.0), se = sqrt(e);
  const Jet<N> exunif = (-0.75 * cbrt(3.0 / M_PI)) * powc(n, 4.0 / 3.0);
  const Jet<N> n53 = powc(n, 5.0 / 3.0);
  const Jet<N> p = (1.0 / (4.0 * K2)) * (s * powc(n, -8.0 / 3.0));
  const Jet<N> tw = 0.125 * (s * recip(n));
  const bool above = tau.v > tw.v;
  const Jet<N> taue = above ? tau : tw;
  const Jet<N> itaue = recip(taue);
  // z = 1 and alpha = 0 exactly (constants) where tau is taken as tau_W
  const Jet<N> z = above ? tw * itaue : cst<N>(1.0);
  const Jet<N> alpha = above ? (1.0 / (0.3 * K2)) * ((taue - tw) * recip(n53)) : cst<N>(0.0);
  const Jet<N> am1 = alpha + cst<N>(-1.0);
  const Jet<N> qb = 0.45 * (am1 * recip(jsqrt(1.0 + b * (alpha * am1)))) + (2.0 / 3.0) * p;
  const Jet<N> r = (0.5 * K2) * (n53 * itaue);
  const Jet<N> z2 = z * z;
  const Jet<N> opz2 = 1.0 + z2;
  const Jet<N> p2 = p * p;
  const Jet<N> num = (cst<N>(10.0 / 81.0) + c * (z2 * recip(opz2 * opz2))) * p + (146.0 / 2025.0) * (qb * qb) -
                     (73.0 / 405.0) * (qb * p * jsqrt(0.18 * (r * r) + cst<N>(0.5))) +
                     ((10.0 / 81.0) * (10.0 / 81.0) / kappa) * p2 + (2.0 * se * 10.0 / 81.0 * 0.36) * z2 +
                     (e * mu) * (p2 * p);
  const Jet<N> den = 1.0 + se * p;
  const Jet<N> x = num * recip(den * den);
  return exunif * (cst<N>(1.0 + kappa) - kappa * recip(1.0 + (1.0 / kappa) * x));
}

// TPSS exchange of one spin channel: e_x(2 rho_s, 4 sigma_ss, 2 tau_s) / 2
template <int N>
__host__ __device__ Jet<N> tpss_x_s(const Jet<N>& r, const Jet<N>& s, const Jet<N>& t) {
  return 0.5 * tpss_x_total(2.0 * r, 4.0 * s, 2.0 * t);
}

template <int N>
__host__ __device__ Jet<N> tpss_c(const Jet<N>& ra, const Jet<N>& rb, const Jet<N>& saa, const Jet<N>& sab,
                                  const Jet<N>& sbb, const Jet<N>& ta, const Jet<N>& tb) {
  const double d = 2.8;
  const Jet<N> n = ra + rb;
  const Jet<N> in = recip(n);
  const Jet<N> z = (ra - rb) * in;
  // 1 +- zeta without cancellation
  const Jet<N> opz = 2.0 * (ra * in), omz = 2.0 * (rb * in);
  const Jet<N> s = saa + 2.0 * sab + sbb;
  const Jet<N> epbe = pbe_c_eps(n, z, opz, omz, s);
  const Jet<N> one = cst<N>(1.0), two = cst<N>(2.0), zero = cst<N>(0.0);
  const Jet<N> eta = jmax(pbe_c_eps(ra, one, two, zero, saa), epbe);
  const Jet<N> etb = jmax(pbe_c_eps(rb, one, two, zero, sbb), epbe);
  const Jet<N> tw = 0.125 * (s * in);
  const Jet<N> tt = ta + tb;
  const Jet<N> zt = tt.v > tw.v ? tw * recip(tt) : cst<N>(1.0);
  const Jet<N> in2 = in * in;
  const Jet<N> gz2 = 4.0 * ((in2 * in2) * (rb * rb * saa - 2.0 * (ra * rb * sab) + ra * ra * sbb));
  const Jet<N> xi2 = (0.25 / pow(3.0 * M_PI * M_PI, 2.0 / 3.0)) * (gz2 * powc(n, -2.0 / 3.0));
  const Jet<N> zz = z * z, z4 = zz * zz;
  const Jet<N> c0 = cst<N>(0.53) + 0.87 * zz + 0.50 * z4 + 2.26 * (z4 * zz);
  const Jet<N> cd = 1.0 + 0.5 * (xi2 * (opzp(opz, -4.0 / 3.0) + opzp(omz, -4.0 / 3.0)));
  const Jet<N> cd2 = cd * cd;
  const Jet<N> C = c0 * recip(cd2 * cd2);
  const Jet<N> zt2 = zt * zt;
  // eps_PBE (1 + C z^2) - (1 + C) z^2 sum_s n_s/n eps_s, written with sum_s n_s/n = 1 so that it
  // is zero exactly (value and derivatives) where z = 1 and eps_s = eps_PBE for both spins
  const Jet<N> erev = epbe * (1.0 + (-1.0) * zt2) + (1.0 + C) * (zt2 * ((ra * in) * (epbe - eta) + (rb * in) * (epbe - etb)));
  return n * (erev * (1.0 + d * (erev * (zt2 * zt))));
}

// energy density of functional `id` (scaled by its weights for the hybrid); only the functionals
// of family FAM are compiled into an instantiation
template <int N, int FAM>
__host__ __device__ Jet<N> energy(int id, const Jet<N>& ra, const Jet<N>& rb, const Jet<N>& saa, const Jet<N>& sab,
                                  const Jet<N>& sbb, const Jet<N>& ta, const Jet<N>& tb, bool on_a, bool on_b) {
  Jet<N> e = cst<N>(0.0);
  auto slater = [&]() {
    Jet<N> x = cst<N>(0.0);
    if (on_a) x = x + slater_s(ra);
    if (on_b) x = x + slater_s(rb);
    return x;
  };
  if constexpr (FAM == 0) {
    switch (id) {
      case XC_LDA_X: e = slater(); break;
      case XC_LDA_C_VWN: e = vwn(false, ra, rb); break;
      case XC_LDA_C_VWN_RPA: e = vwn(true, ra, rb); break;
      default: break;
    }
  } else if constexpr (FAM == 1) {
    auto b88 = [&]() {
      Jet<N> x = cst<N>(0.0);
      if (on_a) x = x + b88_s(ra, saa);
      if (on_b) x = x + b88_s(rb, sbb);
      return x;
    };
    switch (id) {
      case XC_GGA_X_B88: e = b88(); break;
      case XC_GGA_C_LYP: e = lyp(ra, rb, saa, sab, sbb); break;
      case XC_HYB_GGA_XC_B3LYP:
        e = 0.08 * slater() + 0.72 * b88() + 0.19 * vwn(true, ra, rb) + 0.81 * lyp(ra, rb, saa, sab, sbb);
        break;
      default: break;
    }
  } else {
    auto tpssx = [&]() {
      Jet<N> x = cst<N>(0.0);
      if (on_a) x = x + tpss_x_s(ra, saa, ta);
      if (on_b) x = x + tpss_x_s(rb, sbb, tb);
      return x;
    };
    switch (id) {
      case XC_MGGA_X_TPSS: e = tpssx(); break;
      case XC_MGGA_C_TPSS: e = tpss_c(ra, rb, saa, sab, sbb, ta, tb); break;
      case XC_HYB_MGGA_XC_TPSS0: e = 0.75 * tpssx() + tpss_c(ra, rb, saa, sab, sbb, ta, tb); break;
      default: break;
    }
  }
  return e;
}

// FAM: 0 LDA (rho), 1 GGA (rho, sigma), 2 meta-GGA (rho, sigma, tau)
template <int NSPIN, int FAM>
struct XcVars {
  static constexpr int N = NSPIN == 1 ? FAM + 1 : (FAM == 0 ? 2 : (FAM == 1 ? 5 : 7));
};

// The functional's jet at point g over its LibXC variables (unpolarised: rho[, sigma[, tau]];
// polarised: rho_a, rho_b[, sigma_aa, sigma_ab, sigma_bb[, tau_a, tau_b]]).  Returns false below
// the density threshold (LibXC leaves such points zero).  rtot: total density.
template <int NSPIN, int FAM>
__host__ __device__ __forceinline__ bool eval_point(int id, const double* __restrict__ rho, int64_t ld_rho,
                                                    int64_t ld_spin, int64_t g, Jet<XcVars<NSPIN, FAM>::N>& e,
                                                    double& rtot) {
  constexpr int N = XcVars<NSPIN, FAM>::N;
  constexpr bool GGA = FAM >= 1;
  const double r0 = rho[g], r1 = NSPIN == 2 ? rho[g + ld_spin] : 0.0;
  rtot = NSPIN == 2 ? r0 + r1 : r0;
  if (!(rtot >= kDensThr)) return false;
  Jet<N> ra, rb, saa = cst<N>(0.0), sab = cst<N>(0.0), sbb = cst<N>(0.0), ta = cst<N>(0.0), tb = cst<N>(0.0);
  bool on_a = true, on_b = true;
  if (NSPIN == 1) {
    const Jet<N> r = var<N>(r0, 0);
    ra = 0.5 * r;
    rb = ra;
    if (GGA) {
      const double gx = rho[g + ld_rho], gy = rho[g + 2 * ld_rho], gz = rho[g + 3 * ld_rho];
      const double sg = fmax(gx * gx + gy * gy + gz * gz, kSigmaFloor);
      saa = 0.25 * var<N>(sg, 1);
      sab = saa;
      sbb = saa;
    }
    if (FAM == 2) {
      ta = 0.5 * var<N>(rho[g + 4 * ld_rho], 2);
      tb = ta;
    }
  } else {
    // LibXC floors each spin density at the threshold; exchange of a channel at or below it is zero
    on_a = r0 > kDensThr;
    on_b = r1 > kDensThr;
    ra = var<N>(fmax(r0, kDensThr), 0);
    rb = var<N>(fmax(r1, kDensThr), 1);
    if (GGA) {
      double ga[3], gb[3];
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        ga[t] = rho[g + (t + 1) * ld_rho];
        gb[t] = rho[g + (t + 1) * ld_rho + ld_spin];
      }
      const double s_aa = fmax(ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2], kSigmaFloor);
      const double s_bb = fmax(gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2], kSigmaFloor);
      const double s_ab = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
      saa = var<N>(s_aa, 2);
      sab = var<N>(s_ab, 3);
      sbb = var<N>(s_bb, 4);
    }
    if (FAM == 2) {
      ta = var<N>(rho[g + 4 * ld_rho], 5);
      tb = var<N>(rho[g + 4 * ld_rho + ld_spin], 6);
    }
  }
  e = energy<N, FAM>(id, ra, rb, saa, sab, sbb, ta, tb, on_a, on_b);
  return true;
}

// LibXC raw layout: out[g + ld_out c], c over [zk | first | unique second derivatives]
template <int NSPIN, int FAM>
__host__ __device__ __forceinline__ void raw_point(int64_t g, int id, int deriv, const double* __restrict__ rho,
                                                   int64_t ld_rho, int64_t ld_spin, double coef, int accumulate,
                                                   double* __restrict__ out, int64_t ld_out) {
  constexpr int N = XcVars<NSPIN, FAM>::N;
  Jet<N> e;
  double rtot;
  const bool on = eval_point<NSPIN, FAM>(id, rho, ld_rho, ld_spin, g, e, rtot);
  const int ncol = (int)xc_ncol(FAM, NSPIN, deriv);
  auto put = [&](int c, double v) {
    double* o = out + g + ld_out * c;
    *o = accumulate ? *o + coef * v : coef * v;
  };
  if (!on) {
    for (int c = 0; c < ncol; ++c) put(c, 0.0);
    return;
  }
  int c = 0;
  put(c++, e.v / rtot);
  if (deriv >= 1)
#pragma unroll
    for (int i = 0; i < N; ++i) put(c++, e.g[i]);
  if (deriv >= 2) {
    if (NSPIN == 1 || FAM == 0) {
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = i; j < N; ++j) put(c++, e.h[hidx(i, j, N)]);
    } else {
      // LibXC blocks: v2rho2 (aa, ab, bb), v2rhosigma (rho_s x sigma_*), v2rhotau (rho_s x tau_*),
      // v2sigma2 (upper, row-major), v2sigmatau (sigma_* x tau_*), v2tau2 (aa, ab, bb)
      put(c++, e.h[hidx(0, 0, N)]);
      put(c++, e.h[hidx(0, 1, N)]);
      put(c++, e.h[hidx(1, 1, N)]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 2; j < 5; ++j) put(c++, e.h[hidx(i, j, N)]);
      if (FAM == 2)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 5; j < 7; ++j) put(c++, e.h[hidx(i, j, N)]);
#pragma unroll
      for (int i = 2; i < 5; ++i)
#pragma unroll
        for (int j = i; j < 5; ++j) put(c++, e.h[hidx(i, j, N)]);
      if (FAM == 2) {
#pragma unroll
        for (int i = 2; i < 5; ++i)
#pragma unroll
          for (int j = 5; j < 7; ++j) put(c++, e.h[hidx(i, j, N)]);
        put(c++, e.h[hidx(5, 5, N)]);
        put(c++, e.h[hidx(5, 6, N)]);
        put(c++, e.h[hidx(6, 6, N)]);
      }
    }
  }
}

// Effective kernels over the density components (rho, grad rho[, tau]) of each spin, written straight
// from the jet (the chain rule of transform_xc_inner applied in registers): exc[g],
// vxc[g + ng (x + NVL s)], fxc[g + ng ((x + NVL s) + NS NVL (y + NVL t))], NVL the component count
// of the caller's arrays (>= the functional's own: an LDA functional fills the leading slice).
template <int NSPIN, int FAM>
__host__ __device__ __forceinline__ void eff_point(int64_t g, int id, const double* __restrict__ rho, int64_t ld_rho,
                                                   int64_t ld_spin, int64_t ng, int nvl, double scale, int accumulate,
                                                   double* __restrict__ exc, double* __restrict__ vxc,
                                                   double* __restrict__ fxc) {
  constexpr int N = XcVars<NSPIN, FAM>::N;
  constexpr bool GGA = FAM >= 1;
  constexpr int NV = FAM == 0 ? 1 : (FAM == 1 ? 4 : 5);  // density components the functional depends on, per spin
  constexpr int NC = NV * NSPIN;
  Jet<N> e;
  double rtot;
  const bool on = eval_point<NSPIN, FAM>(id, rho, ld_rho, ld_spin, g, e, rtot);
  auto put = [&](double* base, int64_t c, double v) {
    if (!base) return;
    double* o = base + g + ng * c;
    *o = accumulate ? *o + scale * v : scale * v;
  };
  const int ncl = nvl * NSPIN;  // caller's (component, spin) count
  auto oidx = [&](int c) { return (c % NV) + nvl * (c / NV); };
  if (!on) {
    put(exc, 0, 0.0);
    for (int c = 0; c < NC; ++c) put(vxc, oidx(c), 0.0);
    for (int cy = 0; cy < NC; ++cy)
      for (int cx = 0; cx < NC; ++cx) put(fxc, oidx(cx) + (int64_t)ncl * oidx(cy), 0.0);
    return;
  }
  put(exc, 0, e.v / rtot);
  // Jacobian column c = (x, s): LibXC variable k1 with coefficient c1, plus (k2, c2) for the
  // gradient components of a polarised density (sigma_ab depends on both spins' gradients)
  int k1[NC], k2[NC];
  double c1[NC], c2[NC];
#pragma unroll
  for (int sp = 0; sp < NSPIN; ++sp) {
    const int c0 = NV * sp;
    k1[c0] = sp;
    c1[c0] = 1.0;
    k2[c0] = -1;
    c2[c0] = 0.0;
    if (GGA) {
#pragma unroll
      for (int t = 1; t < 4; ++t) {
        const int c = c0 + t;
        const double gs = rho[g + ld_rho * t + (NSPIN == 2 ? ld_spin * sp : 0)];
        if (NSPIN == 1) {
          k1[c] = 1;
          c1[c] = 2.0 * gs;
          k2[c] = -1;
          c2[c] = 0.0;
        } else {
          const double go = rho[g + ld_rho * t + ld_spin * (1 - sp)];
          k1[c] = sp == 0 ? 2 : 4;
          c1[c] = 2.0 * gs;
          k2[c] = 3;
          c2[c] = go;
        }
      }
    }
    if (FAM == 2) {
      // tau_s is a LibXC variable itself
      const int c = c0 + 4;
      k1[c] = NSPIN == 1 ? 2 : 5 + sp;
      c1[c] = 1.0;
      k2[c] = -1;
      c2[c] = 0.0;
    }
  }
  auto H = [&](int p, int q) { return p <= q ? e.h[hidx(p, q, N)] : e.h[hidx(q, p, N)]; };
#pragma unroll
  for (int c = 0; c < NC; ++c) put(vxc, oidx(c), c1[c] * e.g[k1[c]] + (k2[c] >= 0 ? c2[c] * e.g[k2[c]] : 0.0));
  if (!fxc) return;
#pragma unroll
  for (int cy = 0; cy < NC; ++cy)
#pragma unroll
    for (int cx = 0; cx < NC; ++cx) {
      double v = c1[cx] * c1[cy] * H(k1[cx], k1[cy]);
      if (k2[cy] >= 0) v += c1[cx] * c2[cy] * H(k1[cx], k2[cy]);
      if (k2[cx] >= 0) v += c2[cx] * c1[cy] * H(k2[cx], k1[cy]);
      if (k2[cx] >= 0 && k2[cy] >= 0) v += c2[cx] * c2[cy] * H(k2[cx], k2[cy]);
      const int x = cx % NV, y = cy % NV, sx = cx / NV, sy = cy / NV;
      if (GGA && x == y && x >= 1 && x <= 3) {
        // second derivative of sigma in the gradient components
        if (NSPIN == 1)
          v += 2.0 * e.g[1];
        else
          v += sx != sy ? e.g[3] : (sx == 0 ? 2.0 * e.g[2] : 2.0 * e.g[4]);
      }
      put(fxc, oidx(cx) + (int64_t)ncl * oidx(cy), v);
    }
}

template <int NSPIN, int FAM>
__global__ __launch_bounds__(128) void xc_raw_kernel(int id, int deriv, const double* __restrict__ rho, int64_t ld_rho,
                                                     int64_t ld_spin, int64_t ng, double coef, int accumulate,
                                                     double* __restrict__ out, int64_t ld_out) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g < ng) raw_point<NSPIN, FAM>(g, id, deriv, rho, ld_rho, ld_spin, coef, accumulate, out, ld_out);
}

template <int NSPIN, int FAM>
__global__ __launch_bounds__(128) void xc_eff_kernel(int id, const double* __restrict__ rho, int64_t ld_rho,
                                                     int64_t ld_spin, int64_t ng, int nvl, double scale, int accumulate,
                                                     double* __restrict__ exc, double* __restrict__ vxc,
                                                     double* __restrict__ fxc) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g < ng) eff_point<NSPIN, FAM>(g, id, rho, ld_rho, ld_spin, ng, nvl, scale, accumulate, exc, vxc, fxc);
}

}  // namespace

bool xc_known(int id, int* family, double* exx) {
  int f = 0;
  bool ok = true;
  double x = 0.0;
  switch (id) {
    case XC_LDA_X:
    case XC_LDA_C_VWN:
    case XC_LDA_C_VWN_RPA: break;
    case XC_GGA_X_B88:
    case XC_GGA_C_LYP: f = 1; break;
    case XC_HYB_GGA_XC_B3LYP:
      f = 1;
      x = 0.2;
      break;
    case XC_MGGA_X_TPSS:
    case XC_MGGA_C_TPSS: f = 2; break;
    case XC_HYB_MGGA_XC_TPSS0:
      f = 2;
      x = 0.25;
      break;
    default: ok = false;
  }
  if (family) *family = f;
  if (exx) *exx = x;
  return ok;
}

__host__ __device__ int64_t xc_ncol(int family, int nspin, int deriv) {
  const int64_t n = nspin == 1 ? family + 1 : (family == 0 ? 2 : (family == 1 ? 5 : 7));
  int64_t c = 1;
  if (deriv >= 1) c += n;
  if (deriv >= 2) c += n * (n + 1) / 2;
  return c;
}

namespace {
constexpr const char* kUnsupported = "xc: unsupported functional id (LibXC ids 1, 7, 8, 106, 131, 402, 202, 231, 396)";
// components per spin of each family (rho; rho + gradient; rho + gradient + tau)
constexpr int kFamComponents[3] = {1, 4, 5};

template <int NSPIN, int FAM>
void launch_raw(int id, int deriv, const double* rho, int64_t ld_rho, int64_t ld_spin, int64_t ng, double coef, int acc,
                double* out, int64_t ld_out, hipStream_t s) {
  hipLaunchKernelGGL((xc_raw_kernel<NSPIN, FAM>), dim3((unsigned)ceil_div(ng, 128)), dim3(128), 0, s, id, deriv, rho,
                     ld_rho, ld_spin, ng, coef, acc, out, ld_out);
}

template <int NSPIN, int FAM>
void launch_eff(int id, const double* rho, int64_t ld_rho, int64_t ld_spin, int64_t ng, int nvl, double scale, int acc,
                double* exc, double* vxc, double* fxc, hipStream_t s) {
  hipLaunchKernelGGL((xc_eff_kernel<NSPIN, FAM>), dim3((unsigned)ceil_div(ng, 128)), dim3(128), 0, s, id, rho, ld_rho,
                     ld_spin, ng, nvl, scale, acc, exc, vxc, fxc);
}
}  // namespace

void xc_eval(int id, int nspin, int deriv, const double* rho, int64_t ld_rho, int64_t ld_spin, int64_t ng, double coef,
             bool accumulate, double* out, int64_t ld_out, hipStream_t s) {
  int fam = 0;
  HH_REQUIRE(xc_known(id, &fam, nullptr), kUnsupported);
  HH_REQUIRE(nspin == 1 || nspin == 2, "xc: nspin must be 1 or 2");
  HH_REQUIRE(deriv >= 0 && deriv <= 2, "xc: derivative order 0..2");
  HH_REQUIRE(ng >= 0, "xc: negative point count");
  if (ng == 0) return;
  HH_REQUIRE(rho && out, "xc: null argument");
  HH_REQUIRE(ld_out >= ng && ld_rho >= ng && (nspin == 1 || ld_spin >= kFamComponents[fam] * ld_rho),
             "xc: leading dimensions too small");
  const int acc = accumulate ? 1 : 0;
  const int key = 3 * (nspin - 1) + fam;
  switch (key) {
    case 0: launch_raw<1, 0>(id, deriv, rho, ld_rho, ld_spin, ng, coef, acc, out, ld_out, s); break;
    case 1: launch_raw<1, 1>(id, deriv, rho, ld_rho, ld_spin, ng, coef, acc, out, ld_out, s); break;
    case 2: launch_raw<1, 2>(id, deriv, rho, ld_rho, ld_spin, ng, coef, acc, out, ld_out, s); break;
    case 3: launch_raw<2, 0>(id, deriv, rho, ld_rho, ld_spin, ng, coef, acc, out, ld_out, s); break;
    case 4: launch_raw<2, 1>(id, deriv, rho, ld_rho, ld_spin, ng, coef, acc, out, ld_out, s); break;
    default: launch_raw<2, 2>(id, deriv, rho, ld_rho, ld_spin, ng, coef, acc, out, ld_out, s); break;
  }
  HH_HIP(hipGetLastError());
}

void xc_eval_eff(int id, int nspin, int nvl, const double* rho, int64_t ld_rho, int64_t ld_spin, int64_t ng,
                 double scale, bool accumulate, double* exc, double* vxc, double* fxc, hipStream_t s) {
  int fam = 0;
  HH_REQUIRE(xc_known(id, &fam, nullptr), kUnsupported);
  HH_REQUIRE(nspin == 1 || nspin == 2, "xc: nspin must be 1 or 2");
  HH_REQUIRE(nvl == 1 || nvl == 4 || nvl == 5, "xc: output component count must be 1 (RHO), 4 (SIGMA) or 5 (TAU)");
  HH_REQUIRE(fam != 1 || nvl >= 4, "xc: a GGA functional needs density gradients (SIGMA or TAU outputs)");
  HH_REQUIRE(fam != 2 || nvl == 5, "xc: a meta-GGA functional needs tau (TAU outputs)");
  HH_REQUIRE(ng >= 0, "xc: negative point count");
  if (ng == 0) return;
  HH_REQUIRE(rho && (exc || vxc || fxc), "xc: null argument");
  HH_REQUIRE(ld_rho >= ng && (nspin == 1 || ld_spin >= kFamComponents[fam] * ld_rho), "xc: leading dimensions too small");
  const int acc = accumulate ? 1 : 0;
  const int key = 3 * (nspin - 1) + fam;
  switch (key) {
    case 0: launch_eff<1, 0>(id, rho, ld_rho, ld_spin, ng, nvl, scale, acc, exc, vxc, fxc, s); break;
    case 1: launch_eff<1, 1>(id, rho, ld_rho, ld_spin, ng, nvl, scale, acc, exc, vxc, fxc, s); break;
    case 2: launch_eff<1, 2>(id, rho, ld_rho, ld_spin, ng, nvl, scale, acc, exc, vxc, fxc, s); break;
    case 3: launch_eff<2, 0>(id, rho, ld_rho, ld_spin, ng, nvl, scale, acc, exc, vxc, fxc, s); break;
    case 4: launch_eff<2, 1>(id, rho, ld_rho, ld_spin, ng, nvl, scale, acc, exc, vxc, fxc, s); break;
    default: launch_eff<2, 2>(id, rho, ld_rho, ld_spin, ng, nvl, scale, acc, exc, vxc, fxc, s); break;
  }
  HH_HIP(hipGetLastError());
}

}  // namespace hh
