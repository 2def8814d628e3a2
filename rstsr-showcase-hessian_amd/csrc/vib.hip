// Harmonic vibrational analysis on the host (see vib.hpp).  Reference: src/hessian/vib.rs, a
// port of psi4's qcdb/vib.py; the physical constants are the CODATA 2014 values it uses.
#include "vib.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

#include "hh_common.hpp"

namespace hh {
namespace vib {

namespace {
constexpr double NA = 6.022140857e23;
constexpr double HARTREE2J = 4.35974465e-18;
constexpr double CLIGHT = 299792458.0;
constexpr double BOHR2ANG = 0.52917721067;
constexpr double HPLANCK = 6.62607004e-34;
constexpr double KBOLTZ = 1.38064852e-23;
constexpr double AMU2KG = 1.66053904e-27;
constexpr double PI = 3.14159265358979323846;

// cm^-1 per sqrt(E_h / (a0^2 u)) (vib.rs:33-35)
double uconv_cm1() { return std::sqrt(NA * HARTREE2J * 1.0e19) / (2.0 * PI * CLIGHT * BOHR2ANG); }
}  // namespace

void eigh(const double* a_in, int n, std::vector<double>& w, std::vector<double>& v) {
  std::vector<double> a(a_in, a_in + (size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < j; ++i) {
      const double s = 0.5 * (a[i + (size_t)n * j] + a[j + (size_t)n * i]);
      a[i + (size_t)n * j] = a[j + (size_t)n * i] = s;
    }
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) v[i + (size_t)n * i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double x = a[i + (size_t)n * j];
        tot += x * x;
        if (i != j) off += x * x;
      }
    if (off == 0.0 || off <= 1e-32 * tot) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double& apq = a[p + (size_t)n * q];
        const double app = a[p + (size_t)n * p], aqq = a[q + (size_t)n * q];
        if (apq == 0.0) continue;
        const double g = 100.0 * std::fabs(apq);
        if (sweep > 3 && std::fabs(app) + g == std::fabs(app) && std::fabs(aqq) + g == std::fabs(aqq)) {
          apq = 0.0;
          a[q + (size_t)n * p] = 0.0;
          continue;
        }
        // rotation J = [[c, s], [-s, c]] in the (p, q) plane with (J^T A J)_pq = 0
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = a[k + (size_t)n * p], akq = a[k + (size_t)n * q];
          a[k + (size_t)n * p] = c * akp - s * akq;
          a[k + (size_t)n * q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = a[p + (size_t)n * k], aqk = a[q + (size_t)n * k];
          a[p + (size_t)n * k] = c * apk - s * aqk;
          a[q + (size_t)n * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[k + (size_t)n * p], vkq = v[k + (size_t)n * q];
          v[k + (size_t)n * p] = c * vkp - s * vkq;
          v[k + (size_t)n * q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return a[x + (size_t)n * x] < a[y + (size_t)n * y]; });
  w.resize(n);
  std::vector<double> vs((size_t)n * n);
  for (int j = 0; j < n; ++j) {
    w[j] = a[order[j] + (size_t)n * order[j]];
    std::copy(v.begin() + (size_t)n * order[j], v.begin() + (size_t)n * (order[j] + 1), vs.begin() + (size_t)n * j);
  }
  v.swap(vs);
}

void svd_thin(const double* a, int m, int k, std::vector<double>& u, std::vector<double>& s) {
  std::vector<double> b(a, a + (size_t)m * k);
  const double eps = std::numeric_limits<double>::epsilon();
  for (int sweep = 0; sweep < 100; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < k - 1; ++p)
      for (int q = p + 1; q < k; ++q) {
        double al = 0.0, be = 0.0, ga = 0.0;
        for (int i = 0; i < m; ++i) {
          const double x = b[i + (size_t)m * p], y = b[i + (size_t)m * q];
          al += x * x;
          be += y * y;
          ga += x * y;
        }
        if (ga == 0.0 || std::fabs(ga) <= eps * std::sqrt(al * be)) continue;
        // rotate columns p, q so that they become orthogonal
        const double zeta = (be - al) / (2.0 * ga);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
        for (int i = 0; i < m; ++i) {
          const double x = b[i + (size_t)m * p], y = b[i + (size_t)m * q];
          b[i + (size_t)m * p] = c * x - sn * y;
          b[i + (size_t)m * q] = sn * x + c * y;
        }
        rotated = true;
      }
    if (!rotated) break;
  }
  std::vector<double> sv(k);
  for (int j = 0; j < k; ++j) {
    double nn = 0.0;
    for (int i = 0; i < m; ++i) nn += b[i + (size_t)m * j] * b[i + (size_t)m * j];
    sv[j] = std::sqrt(nn);
  }
  std::vector<int> order(k);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return sv[x] > sv[y]; });
  u.assign((size_t)m * k, 0.0);
  s.resize(k);
  for (int j = 0; j < k; ++j) {
    const int o = order[j];
    s[j] = sv[o];
    if (sv[o] > 0.0)
      for (int i = 0; i < m; ++i) u[i + (size_t)m * j] = b[i + (size_t)m * o] / sv[o];
  }
}

std::vector<double> tr_space(const double* mass, const double* geom, int natm, int space, int nrt, int* ncol) {
  const int ndof = 3 * natm;
  std::vector<double> raw;
  int nraw = 0;
  auto push = [&](auto f) {
    raw.resize((size_t)ndof * (nraw + 1));
    for (int a = 0; a < natm; ++a) {
      const double sm = std::sqrt(mass[a]);
      const double x = geom[3 * a], y = geom[3 * a + 1], z = geom[3 * a + 2];
      double d[3];
      f(x, y, z, d);
      for (int c = 0; c < 3; ++c) raw[(size_t)ndof * nraw + 3 * a + c] = sm * d[c];
    }
    ++nraw;
  };
  if (space & 1) {
    push([](double, double, double, double* d) { d[0] = 1; d[1] = 0; d[2] = 0; });
    push([](double, double, double, double* d) { d[0] = 0; d[1] = 1; d[2] = 0; });
    push([](double, double, double, double* d) { d[0] = 0; d[1] = 0; d[2] = 1; });
  }
  if (space & 2) {  // r x e_axis: y uz - z uy, z ux - x uz, x uy - y ux
    push([](double, double y, double z, double* d) { d[0] = 0; d[1] = -z; d[2] = y; });
    push([](double x, double, double z, double* d) { d[0] = z; d[1] = 0; d[2] = -x; });
    push([](double x, double y, double, double* d) { d[0] = -y; d[1] = x; d[2] = 0; });
  }
  if (nraw == 0) {
    *ncol = 0;
    return {};
  }
  std::vector<double> u, s;
  svd_thin(raw.data(), ndof, nraw, u, s);
  int num = nrt;
  if (num < 0) {
    const double smax = s.empty() ? 0.0 : *std::max_element(s.begin(), s.end());
    const double tol = ndof * smax * std::numeric_limits<double>::epsilon();
    num = (int)std::count_if(s.begin(), s.end(), [&](double x) { return x > tol; });
  }
  num = std::min(num, nraw);
  *ncol = num;
  u.resize((size_t)ndof * num);
  return u;
}

void rotation_const(const double* mass, const double* geom_c, int natm, int unit, double* out) {
  double rr[9] = {0};  // sum_a m_a r_a r_a^T
  for (int a = 0; a < natm; ++a)
    for (int r = 0; r < 3; ++r)
      for (int s = 0; s < 3; ++s) rr[r + 3 * s] += mass[a] * geom_c[3 * a + r] * geom_c[3 * a + s];
  const double tr = rr[0] + rr[4] + rr[8];
  double im[9];
  for (int i = 0; i < 9; ++i) im[i] = -rr[i];
  for (int i = 0; i < 3; ++i) im[i + 3 * i] += tr;
  std::vector<double> e, v;
  eigh(im, 3, e, v);
  const double unit_im = AMU2KG * 5.2917721092e-11 * 5.2917721092e-11;
  const double unit_hz = 1.0545718001391127e-34 / (4.0 * PI * unit_im);
  const double conv = unit == 0 ? 1e-9 : 1.0 / CLIGHT * 1e-2;
  for (int i = 0; i < 3; ++i) {
    double x = e[i];
    if (std::fabs(x) < 1e-9) x = 0.0;
    out[i] = std::fabs(x) < 1e-30 ? std::numeric_limits<double>::infinity() : unit_hz / x * conv;
  }
}

int rotor_type(const double* rc) {
  if (rc[0] > 1e8 && rc[1] > 1e8 && rc[2] > 1e8) return ROTOR_ATOM;
  if (rc[0] > 1e8 && std::fabs(rc[1] - rc[2]) < 1e-3) return ROTOR_LINEAR;
  return ROTOR_REGULAR;
}

Result harmonic_analysis(const double* hess, const double* geom, const double* mass, int natm, bool project_trans,
                         bool project_rot) {
  const int n = 3 * natm;
  Result R;
  R.ndof = n;
  // rotor type from the mass-centred geometry fixes the number of projected directions
  double msum = 0.0, mc[3] = {0, 0, 0};
  for (int a = 0; a < natm; ++a) {
    msum += mass[a];
    for (int c = 0; c < 3; ++c) mc[c] += mass[a] * geom[3 * a + c];
  }
  std::vector<double> gc(3 * natm);
  for (int a = 0; a < natm; ++a)
    for (int c = 0; c < 3; ++c) gc[3 * a + c] = geom[3 * a + c] - mc[c] / msum;
  double rc[3];
  rotation_const(mass, gc.data(), natm, 0, rc);
  const int rt = rotor_type(rc);
  int nrt = rt == ROTOR_ATOM ? 0 : rt == ROTOR_LINEAR ? 5 : 6;
  if (!project_trans) nrt = std::max(nrt, 3) - 3;
  if (!project_rot) {
    const int nrot = rt == ROTOR_ATOM ? 0 : rt == ROTOR_LINEAR ? 2 : 3;
    nrt = std::max(nrt, nrot) - nrot;
  }
  int ntr = 0;
  const std::vector<double> tr =
      tr_space(mass, geom, natm, (project_trans ? 1 : 0) | (project_rot ? 2 : 0), nrt, &ntr);
  // P = I - sum |tr><tr|
  std::vector<double> P((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) P[i + (size_t)n * i] = 1.0;
  for (int c = 0; c < ntr; ++c)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) P[i + (size_t)n * j] -= tr[i + (size_t)n * c] * tr[j + (size_t)n * c];
  // mass-weighted, projected Hessian P^T (M^-1/2 H M^-1/2) P
  std::vector<double> sminv(n);
  for (int a = 0; a < natm; ++a)
    for (int c = 0; c < 3; ++c) sminv[3 * a + c] = 1.0 / std::sqrt(mass[a]);
  std::vector<double> mw((size_t)n * n), tmp((size_t)n * n, 0.0), proj((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) mw[i + (size_t)n * j] = hess[i + (size_t)n * j] * sminv[i] * sminv[j];
  for (int j = 0; j < n; ++j)
    for (int l = 0; l < n; ++l) {
      const double plj = P[l + (size_t)n * j];
      if (plj == 0.0) continue;
      for (int i = 0; i < n; ++i) tmp[i + (size_t)n * j] += mw[i + (size_t)n * l] * plj;
    }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double sacc = 0.0;
      for (int l = 0; l < n; ++l) sacc += P[l + (size_t)n * i] * tmp[l + (size_t)n * j];
      proj[i + (size_t)n * j] = sacc;
    }
  std::vector<double> fc, q;
  eigh(proj.data(), n, fc, q);
  // phase: the element of largest magnitude (first within 1e-2 of it) is positive
  for (int v = 0; v < n; ++v) {
    double vext = 0.0;
    for (int r = 0; r < n; ++r) vext = std::max(vext, std::fabs(q[r + (size_t)n * v]));
    int iext = 0;
    for (int r = 0; r < n; ++r)
      if (vext - std::fabs(q[r + (size_t)n * v]) < 1e-2) {
        iext = r;
        break;
      }
    if (q[iext + (size_t)n * v] < 0.0)
      for (int r = 0; r < n; ++r) q[r + (size_t)n * v] = -q[r + (size_t)n * v];
  }
  const double uc = uconv_cm1();
  R.omega.resize(n);
  R.imag.resize(n);
  std::vector<double> wre(n);
  for (int i = 0; i < n; ++i) {
    R.imag[i] = fc[i] < 0.0;
    R.omega[i] = std::sqrt(std::fabs(fc[i])) * uc;
    wre[i] = R.imag[i] ? 0.0 : R.omega[i];
  }
  // degeneracy: modes sharing round(10 * real frequency)
  R.degeneracy.assign(n, 0);
  {
    std::vector<std::pair<long long, int>> keys(n);
    for (int i = 0; i < n; ++i) keys[i] = {std::llround(wre[i] * 10.0), i};
    std::stable_sort(keys.begin(), keys.end(), [](auto& x, auto& y) { return x.first < y.first; });
    for (int s0 = 0; s0 < n;) {
      int e0 = s0 + 1;
      while (e0 < n && keys[e0].first == keys[s0].first) ++e0;
      for (int j = s0; j < e0; ++j) R.degeneracy[keys[j].second] = e0 - s0;
      s0 = e0;
    }
  }
  // TR: the nrt modes of smallest |force constant|; "-" for other near-zero modes
  {
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return std::fabs(fc[x]) < std::fabs(fc[y]); });
    std::vector<bool> is_tr(n, false);
    for (int i = 0; i < std::min(nrt, n); ++i) is_tr[ord[i]] = true;
    R.trv.resize(n);
    for (int i = 0; i < n; ++i) R.trv[i] = is_tr[i] ? TRV_TR : (std::fabs(fc[i]) < 1e-3 ? TRV_NONE : TRV_V);
  }
  const double uconv_mdyne_a = 0.1 * (2.0 * PI * CLIGHT) * (2.0 * PI * CLIGHT) / NA;
  const double uconv_s = std::sqrt((CLIGHT * (2.0 * PI * BOHR2ANG) * (2.0 * PI * BOHR2ANG)) / (HPLANCK * NA * 1.0e21));
  R.q = q;
  R.w.resize((size_t)n * n);
  R.x.resize((size_t)n * n);
  R.mu.resize(n);
  R.k.resize(n);
  R.qtp0.resize(n);
  R.xtp0.resize(n);
  R.dq0.resize(n);
  R.theta_vib.resize(n);
  const double uconv_k = 100.0 * HPLANCK * CLIGHT / KBOLTZ;
  for (int j = 0; j < n; ++j) {
    double nn = 0.0;
    for (int i = 0; i < n; ++i) {
      const double wv = q[i + (size_t)n * j] * sminv[i];
      R.w[i + (size_t)n * j] = wv;
      nn += wv * wv;
    }
    R.mu[j] = 1.0 / nn;
    const double smu = std::sqrt(R.mu[j]);
    for (int i = 0; i < n; ++i) R.x[i + (size_t)n * j] = R.w[i + (size_t)n * j] * smu;
    R.k[j] = R.mu[j] * wre[j] * wre[j] * uconv_mdyne_a;
    const double dq = std::sqrt(wre[j]) * uconv_s, dx = std::sqrt(wre[j] * R.mu[j]) * uconv_s;
    R.qtp0[j] = (dq == 0.0 || !std::isfinite(dq)) ? 0.0 : 1.0 / dq;
    R.xtp0[j] = (dx == 0.0 || !std::isfinite(dx)) ? 0.0 : 1.0 / dx;
    R.dq0[j] = R.qtp0[j] / std::sqrt(2.0);
    R.theta_vib[j] = wre[j] * uconv_k;
  }
  return R;
}

}  // namespace vib
}  // namespace hh
