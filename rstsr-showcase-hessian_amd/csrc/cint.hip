// Gaussian-basis AO values and two-electron Coulomb integrals on the device (see cint.hpp).
//
// AO values (eval_gto): one thread per (grid point, contracted shell); the Cartesian monomial
// x^i y^j z^k e^{-a r^2} and its derivatives factor per direction, with
// d^n/dx^n [x^k e^{-a x^2}] = p_{k,n}(x) e^{-a x^2}, p_{k,0} = x^k,
// p_{k,n+1} = k p_{k-1,n} - 2 a p_{k+1,n}; contracted over primitives, then taken to real solid
// harmonics.
//
// Coulomb integrals (ab|c) over primitive Cartesian Gaussians by McMurchie-Davidson: the product
// a b expanded in Hermite Gaussians at P (coefficients E^{ij}_t per direction), c at its own
// centre, (Lambda_tuv | Lambda_tau nu phi) = (-1)^{tau+nu+phi} 2 pi^{5/2} / (p g sqrt(p + g))
// R_{t+tau, u+nu, v+phi}(rho, P - C), the R table from the Boys function by the Hermite
// recursion.  Derivative integrals (libcint's ip: the gradient with respect to the electron
// coordinate acting on one function) are combinations of integrals over Cartesian functions of
// shifted angular momentum: d/dx [x^i e^{-a x^2}] = i x^{i-1} e^{..} - 2 a x^{i+1} e^{..}.
// 2-centre integrals (a|c) are the same with the unit function (exponent 0) as b.
//
// One workgroup (one wave) per contracted shell triple; every per-triple table lives in LDS.
#include "cint.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>

namespace hh {

namespace {

constexpr int LMAXS = Cint::LMAX;
constexpr int C2S_TOTAL = 476;  // sum over l <= 5 of (2l + 1)(l + 1)(l + 2)/2
__constant__ double c_c2s[C2S_TOTAL];

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int c2s_off(int l) {
  return l == 0 ? 0 : l == 1 ? 1 : l == 2 ? 10 : l == 3 ? 40 : l == 4 ? 110 : 245;
}
// index of (lx, ly, lz) within its shell in libcint's Cartesian order (x^l, x^{l-1} y, x^{l-1} z, ...)
__host__ __device__ constexpr int cart_index(int l, int lx, int lz) { return (l - lx) * (l - lx + 1) / 2 + lz; }
// the s / p factors libcint puts on Cartesian functions
__host__ __device__ constexpr double cart_fac(int l) {
  return l == 0 ? 0.282094791773878143 : l == 1 ? 0.488602511902919921 : 1.0;
}

double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}
double fact(int n) {
  double r = 1.0;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}

// Real solid harmonics r^l Y_lm (unit-sphere normalised) as Cartesian polynomials: the closed
// form of Helgaker, Jorgensen and Olsen, Molecular Electronic-Structure Theory, eqs. 6.4.47-48
// (S_lm with norm 4 pi / (2l + 1) on the sphere), times sqrt((2l + 1) / (4 pi)).  Rows in
// libcint's order: m = -l .. l, and x, y, z for l = 1.
std::vector<double> c2s_host() {
  std::vector<double> t(C2S_TOTAL, 0.0);
  for (int l = 0; l <= LMAXS; ++l) {
    const int nc = ncart(l);
    const double unit = std::sqrt((2.0 * l + 1.0) / (4.0 * M_PI));
    for (int m = -l; m <= l; ++m) {
      const int am = std::abs(m);
      const double nlm = 1.0 / (std::pow(2.0, am) * fact(l)) *
                         std::sqrt(2.0 * fact(l + am) * fact(l - am) / (m == 0 ? 2.0 : 1.0));
      int row = m + l;
      if (l == 1) row = (m == 1) ? 0 : (m == -1) ? 1 : 2;
      double* r = &t[c2s_off(l) + row * nc];
      for (int tt = 0; tt <= (l - am) / 2; ++tt)
        for (int u = 0; u <= tt; ++u)
          for (int v2 = (m < 0 ? 1 : 0); v2 <= am; v2 += 2) {
            // (-1)^(t + v - v_m) with v = v2 / 2, v_m = 1/2 for m < 0
            const int sgn_exp = tt + (v2 - (m < 0 ? 1 : 0)) / 2;
            const double c = ((sgn_exp & 1) ? -1.0 : 1.0) * std::pow(0.25, tt) * binom(l, tt) * binom(l - tt, am + tt) *
                             binom(tt, u) * binom(am, v2);
            const int px = 2 * tt + am - 2 * u - v2, py = 2 * u + v2, pz = l - 2 * tt - am;
            r[cart_index(l, px, pz)] += unit * nlm * c;
            (void)py;
          }
    }
  }
  return t;
}

void upload_c2s() {
  static bool done = false;
  if (done) return;
  std::vector<double> t = c2s_host();
  HH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_c2s), t.data(), sizeof(double) * C2S_TOTAL));
  done = true;
}

// ---------------------------------------------------------------------------------------------
// AO values

// out[k] = p_{k, n}(x) for k <= L (powers xp[k] = x^k, k <= L + n)
template <int L, int NMAX>
__device__ __forceinline__ void ptab(const double* xp, int n, double a, double* out) {
  double b[L + NMAX + 1];
#pragma unroll
  for (int k = 0; k <= L + NMAX; ++k) b[k] = xp[k];
#pragma unroll
  for (int j = 0; j < NMAX; ++j) {
    if (j < n) {
      double prev = 0.0;  // old b[k - 1]
#pragma unroll
      for (int k = 0; k < L + NMAX - j; ++k) {
        const double cur = b[k];
        b[k] = k * prev - 2.0 * a * b[k + 1];
        prev = cur;
      }
    }
  }
#pragma unroll
  for (int k = 0; k <= L; ++k) out[k] = b[k];
}

template <int L, int D>
__global__ __launch_bounds__(256) void eval_gto_kernel(const CShell* __restrict__ cs, const int* __restrict__ list,
                                                       const double* __restrict__ prim,
                                                       const double* __restrict__ coords, int64_t ng, int64_t nao,
                                                       int cart, double* __restrict__ out) {
  constexpr int NC = ncart(L);
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= ng) return;
  const CShell sh = cs[list[blockIdx.y]];
  const double x = coords[3 * g] - sh.x, y = coords[3 * g + 1] - sh.y, z = coords[3 * g + 2] - sh.z;
  const double r2 = x * x + y * y + z * z;
  double xp[L + D + 1], yp[L + D + 1], zp[L + D + 1];
  xp[0] = yp[0] = zp[0] = 1.0;
#pragma unroll
  for (int k = 1; k <= L + D; ++k) {
    xp[k] = xp[k - 1] * x;
    yp[k] = yp[k - 1] * y;
    zp[k] = zp[k - 1] * z;
  }
  const double* ex = prim + sh.prim_off;
  const double* co = ex + sh.nprim;
  int d = 0;
#pragma unroll
  for (int n = 0; n <= D; ++n)
#pragma unroll
    for (int dx = n; dx >= 0; --dx)
#pragma unroll
      for (int dy = n - dx; dy >= 0; --dy) {
        const int dz = n - dx - dy;
        double acc[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) acc[i] = 0.0;
        for (int p = 0; p < sh.nprim; ++p) {
          const double a = ex[p];
          if (a * r2 > 60.0) continue;
          const double e = co[p] * exp(-a * r2);
          double px[L + 1], py[L + 1], pz[L + 1];
          ptab<L, D>(xp, dx, a, px);
          ptab<L, D>(yp, dy, a, py);
          ptab<L, D>(zp, dz, a, pz);
#pragma unroll
          for (int lx = L; lx >= 0; --lx)
#pragma unroll
            for (int ly = L - lx; ly >= 0; --ly) {
              const int lz = L - lx - ly;
              acc[cart_index(L, lx, lz)] += e * px[lx] * py[ly] * pz[lz];
            }
        }
        double* o = out + g + ng * (sh.ao0 + nao * (int64_t)d);
        if (cart) {
#pragma unroll
          for (int i = 0; i < NC; ++i) o[ng * i] = cart_fac(L) * acc[i];
        } else {
#pragma unroll
          for (int m = 0; m < 2 * L + 1; ++m) {
            double v = 0.0;
#pragma unroll
            for (int i = 0; i < NC; ++i) v += c_c2s[c2s_off(L) + m * NC + i] * acc[i];
            o[ng * m] = v;
          }
        }
        ++d;
      }
}

// ---------------------------------------------------------------------------------------------
// Coulomb integrals

// Boys function F_n(T), n = 0 .. nmax: for T <= 35 the series
// F_n(T) = e^{-T} sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1)) at n = nmax and the downward
// recursion F_{n-1} = (2T F_n + e^{-T}) / (2n - 1); above, F_0 from erf and the upward recursion
// F_{n+1} = ((2n+1) F_n - e^{-T}) / (2T) (stable there: T > nmax).
// Boys function table F_m(T_i), T_i = i h (i < BOYS_NT, h = 1/20, so T_i <= 35), m < BOYS_M:
// generated on the host by the series at F_{BOYS_M - 1} and downward recursion in long double.
// The device takes F_nmax(T) from the nearest point by the Taylor series dF_m/dT = -F_{m+1},
// eight terms (|T - T_i| <= h/2 and F_{m+8} <= F_m: remainder below 4e-18 of F_nmax), then
// recurses downward as before.
constexpr int BOYS_NT = 701, BOYS_M = 40;
constexpr double BOYS_H = 0.05;
__device__ double g_boys[BOYS_NT * BOYS_M];

std::vector<double> boys_host() {
  std::vector<double> t((size_t)BOYS_NT * BOYS_M);
  const int top = BOYS_M - 1;
  for (int i = 0; i < BOYS_NT; ++i) {
    const long double T = (long double)(i * BOYS_H);  // the double the device subtracts
    const long double et = expl(-T);
    long double term = 1.0L / (2 * top + 1), sum = term;
    for (int k = 1; k < 4000; ++k) {
      term *= 2.0L * T / (2 * top + 2 * k + 1);
      sum += term;
      if (term < 1e-24L * sum) break;
    }
    long double F = et * sum;
    t[(size_t)i * BOYS_M + top] = (double)F;
    for (int n = top; n > 0; --n) {
      F = (2.0L * T * F + et) / (2 * n - 1);
      t[(size_t)i * BOYS_M + n - 1] = (double)F;
    }
  }
  return t;
}

void upload_boys() {
  static bool done = false;
  if (done) return;
  std::vector<double> b = boys_host();
  HH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_boys), b.data(), sizeof(double) * b.size()));
  done = true;
}

__device__ void boys(int nmax, double T, double* F) {
  if (T <= 35.0 && nmax + 8 <= BOYS_M) {
    const int i = (int)(T * 20.0 + 0.5);
    const double x = i * BOYS_H - T;  // -(T - T_i)
    const double* row = g_boys + i * BOYS_M + nmax;
    double s = row[7];
#pragma unroll
    for (int k = 6; k >= 0; --k) s = row[k] + s * x * (1.0 / (k + 1));
    const double et = exp(-T);
    F[nmax] = s;
    for (int n = nmax; n > 0; --n) F[n - 1] = (2.0 * T * F[n] + et) / (2 * n - 1);
  } else if (T <= 35.0) {
    const double et = exp(-T);
    double term = 1.0 / (2 * nmax + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
      term *= 2.0 * T / (2 * nmax + 2 * k + 1);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    F[nmax] = et * sum;
    for (int n = nmax; n > 0; --n) F[n - 1] = (2.0 * T * F[n] + et) / (2 * n - 1);
  } else {
    const double et = exp(-T);
    F[0] = 0.5 * sqrt(M_PI / T) * erf(sqrt(T));
    for (int n = 0; n < nmax; ++n) F[n + 1] = ((2 * n + 1) * F[n] - et) / (2.0 * T);
  }
}

__device__ __forceinline__ int tet(int n) { return n * (n + 1) * (n + 2) / 6; }
// compact index of (t, u, v), t + u + v <= L
__device__ __forceinline__ int ridx(int t, int u, int v, int L) {
  const int m = L - t;
  return tet(L + 1) - tet(L + 1 - t) + u * (m + 1) - u * (u - 1) / 2 + v;
}

struct TripleArgs {
  const CShell* csa;  // shells of the a / c positions (3c: a, b from mol, c from aux)
  const CShell* csb;
  const CShell* csc;
  const double* prima;
  const double* primb;
  const double* primc;
  const int* la_list;  // contracted shells of this group
  const int* lb_list;
  const int* lc_list;
  int na_list, nb_list, nc_list;
  int la, lb, lc;      // angular momenta of the group (lb = 0 and two_c for 2-centre)
  int two_c;
  int sym_ab;          // (ab|c) = (ba|c) (same a / b shells, no derivative on a or b, or one on
                       // each with the components swapped): only la >= lb groups are launched,
                       // la == lb triples with a >= b, both blocks written
  int nda, ndb, ndc;   // derivatives on a / b / c
  int kind;            // HH_INT_*
  int q0, q1;          // derivative components of this launch
  int cart;
  // output: element (ia, ib, ic, q) of the triple at
  // out[(ao_a + ia) + s1 ((ao_b + ib) + s2 ((ao_c + ic - c0) + s3 q))]
  double* out;
  int64_t s1, s2, s3, c0;
  // LDS layout (doubles)
  int LA, LB, LC, LT, NR;
  int nA, nB, nC;      // sizes of the shifted Cartesian sets
  int loA, loB, loC;   // lowest l of each set
  int offR, offE, offEc, offF, offI, offAcc, offSet, lds_doubles;
};

// Index of (lx, ly, lz) in a shifted Cartesian set holding the l = lo, lo + 2, .. of one parity
// (n derivatives move l by +-1 each, so only l0 + n, l0 + n - 2, .. >= 0 occur)
__device__ __forceinline__ int set_index(int lo, int lx, int ly, int lz) {
  const int l = lx + ly + lz;
  int k = 0;
  for (int m = lo; m < l; m += 2) k += (m + 1) * (m + 2) / 2;
  return k + cart_index(l, lx, lz);
}

// Derivative terms of libcint's ip operators on a Cartesian function, ND (0..2) compile-time:
// up to 2^ND terms (coefficient, shifted powers) in registers; a lowering step on a zero power
// leaves a term of coefficient 0 (skipped by the consumer).  Exponents are taken as 1: the
// contracted shifted-set integrals carry their powers.
template <int ND>
struct DTerms {
  double c[1 << ND];
  int x[1 << ND], y[1 << ND], z[1 << ND];
};
template <int ND>
__device__ __forceinline__ void dterms(int lx, int ly, int lz, int d1, int d2, DTerms<ND>& t) {
  double c[4] = {1.0, 0.0, 0.0, 0.0};
  int x[4] = {lx, 0, 0, 0}, y[4] = {ly, 0, 0, 0}, z[4] = {lz, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < ND; ++s) {
    const int d = s == 0 ? d1 : d2;
    const int n = 1 << s;
#pragma unroll
    for (int k = n - 1; k >= 0; --k) {
      const int l = d == 0 ? x[k] : (d == 1 ? y[k] : z[k]);
      const int ex = d == 0, ey = d == 1, ez = d == 2;
      c[2 * k + 1] = -2.0 * c[k];
      x[2 * k + 1] = x[k] + ex; y[2 * k + 1] = y[k] + ey; z[2 * k + 1] = z[k] + ez;
      c[2 * k] = c[k] * l;
      x[2 * k] = x[k] - ex; y[2 * k] = y[k] - ey; z[2 * k] = z[k] - ez;
    }
  }
#pragma unroll
  for (int k = 0; k < (1 << ND); ++k) {
    t.c[k] = c[k]; t.x[k] = x[k]; t.y[k] = y[k]; t.z[k] = z[k];
  }
}

// Cartesian powers of component k (libcint order: lx descending, then ly descending) of l
__device__ __forceinline__ void cart_comp(int l, int k, int& lx, int& ly, int& lz) {
  int m = 0;
  while ((m + 1) * (m + 2) / 2 <= k) ++m;
  lx = l - m;
  ly = m - (k - m * (m + 1) / 2);
  lz = m - ly;
}

template <int KIND>
__global__ __launch_bounds__(64) void int_triple_kernel(TripleArgs A) {
  constexpr int NDA = KIND == HH_INT_IP1 || KIND == HH_INT_IPVIP1 || KIND == HH_INT_IP1IP2 ? 1
                      : (KIND == HH_INT_IPIP1 ? 2 : 0);
  constexpr int NDB = KIND == HH_INT_IPVIP1 ? 1 : 0;
  constexpr int NDC = KIND == HH_INT_IP2 || KIND == HH_INT_IP1IP2 ? 1 : (KIND == HH_INT_IPIP2 ? 2 : 0);
  extern __shared__ double lds[];
  const int tid = threadIdx.x;
  const int ia_s = A.la_list[blockIdx.x % A.na_list];
  const int ib_s = A.two_c ? -1 : A.lb_list[blockIdx.x / A.na_list];
  const int ic_s = A.lc_list[blockIdx.y];
  if (A.sym_ab && A.la == A.lb && ia_s < ib_s) return;  // its mirror image writes this block
  const CShell sa = A.csa[ia_s];
  CShell sb;
  if (A.two_c) {
    sb.x = sa.x; sb.y = sa.y; sb.z = sa.z; sb.l = 0; sb.nprim = 1; sb.prim_off = 0; sb.ao0 = 0;
  } else {
    sb = A.csb[ib_s];
  }
  const CShell sc = A.csc[ic_s];
  double* R0 = lds + A.offR;
  double* R1 = R0 + A.NR;
  double* Eab = lds + A.offE;
  double* Ec = lds + A.offEc;
  double* F = lds + A.offF;
  double* I = lds + A.offI;
  double* acc = lds + A.offAcc;
  int* setA = reinterpret_cast<int*>(lds + A.offSet);
  int* setB = setA + A.nA;
  int* setC = setB + A.nB;
  const int LA = A.LA, LB = A.LB, LC = A.LC, LT = A.LT;
  const int ncA = ncart(A.la), ncB = ncart(A.lb), ncC = ncart(A.lc);
  const int nq = A.q1 - A.q0;
  const int nacc = ncA * ncB * ncC * nq;
  // shifted Cartesian sets (packed lx | ly << 8 | lz << 16)
  if (tid == 0) {
    int k = 0;
    for (int l = A.loA; l <= LA; l += 2)
      for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) setA[k++] = lx | (ly << 8) | ((l - lx - ly) << 16);
    k = 0;
    for (int l = A.loB; l <= LB; l += 2)
      for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) setB[k++] = lx | (ly << 8) | ((l - lx - ly) << 16);
    k = 0;
    for (int l = A.loC; l <= LC; l += 2)
      for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) setC[k++] = lx | (ly << 8) | ((l - lx - ly) << 16);
  }
  for (int e = tid; e < A.nA * A.nB * A.nC; e += 64) I[e] = 0.0;
  __syncthreads();
  const double* exa = A.prima + sa.prim_off;
  const double* exb = A.two_c ? nullptr : A.primb + sb.prim_off;
  const double* exc = A.primc + sc.prim_off;
  const double ABx = sa.x - sb.x, ABy = sa.y - sb.y, ABz = sa.z - sb.z;
  const double ab2 = ABx * ABx + ABy * ABy + ABz * ABz;
  const int nEab = (LA + 1) * (LB + 1) * (LA + LB + 1);
  const int nEc = (LC + 1) * (LC + 1);
  for (int pa = 0; pa < sa.nprim; ++pa)
    for (int pb = 0; pb < sb.nprim; ++pb) {
      const double al = exa[pa], be = A.two_c ? 0.0 : exb[pb];
      const double ca = exa[sa.nprim + pa], cb = A.two_c ? 1.0 : exb[sb.nprim + pb];
      const double p = al + be;
      const double mu = al * be / p;
      if (mu * ab2 > 40.0) continue;  // K_AB < 4e-18
      const double Kab = exp(-mu * ab2);
      const double Px = (al * sa.x + be * sb.x) / p, Py = (al * sa.y + be * sb.y) / p, Pz = (al * sa.z + be * sb.z) / p;
      // E^{ij}_t per direction (threads 0..2), E^{ab} index ((i (LB+1) + j) (LA+LB+1) + t)
      if (tid < 3) {
        const double PA = (tid == 0 ? Px - sa.x : tid == 1 ? Py - sa.y : Pz - sa.z);
        const double PB = (tid == 0 ? Px - sb.x : tid == 1 ? Py - sb.y : Pz - sb.z);
        double* E = Eab + tid * nEab;
        const int W = LA + LB + 1;
        for (int e = 0; e < nEab; ++e) E[e] = 0.0;
        E[0] = 1.0;
        const double h = 0.5 / p;
        for (int i = 0; i <= LA; ++i) {
          if (i > 0) {  // (i, 0) from (i-1, 0)
            const double* s = E + ((i - 1) * (LB + 1)) * W;
            double* d = E + (i * (LB + 1)) * W;
            for (int t = 0; t <= i; ++t)
              d[t] = (t > 0 ? h * s[t - 1] : 0.0) + PA * s[t] + (t + 1 <= i - 1 ? (t + 1) * s[t + 1] : 0.0);
          }
          for (int j = 1; j <= LB; ++j) {  // (i, j) from (i, j-1)
            const double* s = E + (i * (LB + 1) + j - 1) * W;
            double* d = E + (i * (LB + 1) + j) * W;
            for (int t = 0; t <= i + j; ++t)
              d[t] = (t > 0 ? h * s[t - 1] : 0.0) + PB * s[t] + (t + 1 <= i + j - 1 ? (t + 1) * s[t + 1] : 0.0);
          }
        }
      }
      for (int pc = 0; pc < sc.nprim; ++pc) {
        const double ga = exc[pc], cc = exc[sc.nprim + pc];
        const double rho = p * ga / (p + ga);
        const double PCx = Px - sc.x, PCy = Py - sc.y, PCz = Pz - sc.z;
        const double T = rho * (PCx * PCx + PCy * PCy + PCz * PCz);
        const double pref = 2.0 * pow(M_PI, 2.5) / (p * ga * sqrt(p + ga)) * Kab;
        if (tid == 3) boys(LT, T, F);
        if (tid >= 4 && tid < 7) {  // E^c per direction: E[k (LC+1) + tau]
          const int dd = tid - 4;
          double* E = Ec + dd * nEc;
          for (int e = 0; e < nEc; ++e) E[e] = 0.0;
          E[0] = 1.0;
          const double h = 0.5 / ga;
          for (int k = 1; k <= LC; ++k)
            for (int t = 0; t <= k; ++t)
              E[k * (LC + 1) + t] = (t > 0 ? h * E[(k - 1) * (LC + 1) + t - 1] : 0.0) +
                                    (t + 1 <= k - 1 ? (t + 1) * E[(k - 1) * (LC + 1) + t + 1] : 0.0);
        }
        __syncthreads();
        // R^n_tuv, n = LT .. 0 (level n holds t + u + v <= LT - n)
        double* Rn = (LT % 2 == 0) ? R0 : R1;
        if (tid == 0) Rn[0] = pow(-2.0 * rho, LT) * F[LT];
        __syncthreads();
        for (int n = LT - 1; n >= 0; --n) {
          const double* Rp = Rn;
          Rn = (n % 2 == 0) ? R0 : R1;
          const int Ln = LT - n;
          const int cnt = tet(Ln + 1);
          const double r00 = pow(-2.0 * rho, n) * F[n];
          for (int e = tid; e < cnt; e += 64) {
            // decode e (compact order of sums <= Ln) into (t, u, v)
            int t = 0, rem = e;
            while (true) {
              const int sz = (Ln - t + 1) * (Ln - t + 2) / 2;
              if (rem < sz) break;
              rem -= sz;
              ++t;
            }
            int u = 0;
            while (true) {
              const int sz = Ln - t - u + 1;
              if (rem < sz) break;
              rem -= sz;
              ++u;
            }
            const int v = rem;
            double val;
            if (t > 0)
              val = (t > 1 ? (t - 1) * Rp[ridx(t - 2, u, v, LT)] : 0.0) + PCx * Rp[ridx(t - 1, u, v, LT)];
            else if (u > 0)
              val = (u > 1 ? (u - 1) * Rp[ridx(t, u - 2, v, LT)] : 0.0) + PCy * Rp[ridx(t, u - 1, v, LT)];
            else if (v > 0)
              val = (v > 1 ? (v - 1) * Rp[ridx(t, u, v - 2, LT)] : 0.0) + PCz * Rp[ridx(t, u, v - 1, LT)];
            else
              val = r00;
            Rn[ridx(t, u, v, LT)] = val;
          }
          __syncthreads();
        }
        // primitive integrals over the shifted sets, contracted into I[(a nB + b) nC + c] with the
        // exponent powers of their derivative terms: an element of a set at angular momentum l
        // arises from r = (l - l0 + n) / 2 raising steps, each carrying -2 alpha, so the sum over
        // primitives weighs it by alpha^r and the derivative coefficients are applied once, after
        // the primitive loops, with alpha = 1
        const double wgt = ca * cb * cc * pref;
        const double pwa[3] = {1.0, al, al * al}, pwb[3] = {1.0, be, be * be}, pwc[3] = {1.0, ga, ga * ga};
        const int W = LA + LB + 1;
        const int ntot = A.nA * A.nB * A.nC;
        for (int e = tid; e < ntot; e += 64) {
          const int ic = e % A.nC, ib = (e / A.nC) % A.nB, iA = e / (A.nC * A.nB);
          const int a = setA[iA], b = setB[ib], c = setC[ic];
          const int ax = a & 255, ay = (a >> 8) & 255, az = a >> 16;
          const int bx = b & 255, by = (b >> 8) & 255, bz = b >> 16;
          const int cx = c & 255, cy = (c >> 8) & 255, cz = c >> 16;
          const double* Ex = Eab + (ax * (LB + 1) + bx) * W;
          const double* Ey = Eab + nEab + (ay * (LB + 1) + by) * W;
          const double* Ez = Eab + 2 * nEab + (az * (LB + 1) + bz) * W;
          const double* Gx = Ec + cx * (LC + 1);
          const double* Gy = Ec + nEc + cy * (LC + 1);
          const double* Gz = Ec + 2 * nEc + cz * (LC + 1);
          double s = 0.0;
          for (int t = 0; t <= ax + bx; ++t)
            for (int u = 0; u <= ay + by; ++u)
              for (int v = 0; v <= az + bz; ++v) {
                const double eab = Ex[t] * Ey[u] * Ez[v];
                if (eab == 0.0) continue;
                double hsum = 0.0;
                for (int tau = cx & 1; tau <= cx; tau += 2)
                  for (int nu = cy & 1; nu <= cy; nu += 2)
                    for (int ph = cz & 1; ph <= cz; ph += 2) {
                      const double g = Gx[tau] * Gy[nu] * Gz[ph];
                      hsum += (((tau + nu + ph) & 1) ? -g : g) * R0[ridx(t + tau, u + nu, v + ph, LT)];
                    }
                s += eab * hsum;
              }
          const int ra = (ax + ay + az - A.la + NDA) >> 1, rb = (bx + by + bz - A.lb + NDB) >> 1;
          const int rc = (cx + cy + cz - A.lc + NDC) >> 1;
          I[e] += wgt * pwa[ra] * pwb[rb] * pwc[rc] * s;
        }
        __syncthreads();
      }
    }
  // derivative integrals of the Cartesian components from the contracted shifted-set integrals
  for (int e = tid; e < nacc; e += 64) {
    const int q = A.q0 + e % nq;
    const int rest = e / nq;
    const int jc = rest % ncC, jb = (rest / ncC) % ncB, ja = rest / (ncC * ncB);
    int lxa, lya, lza, lxb = 0, lyb = 0, lzb = 0, lxc, lyc, lzc;
    cart_comp(A.la, ja, lxa, lya, lza);
    if (!A.two_c) cart_comp(A.lb, jb, lxb, lyb, lzb);
    cart_comp(A.lc, jc, lxc, lyc, lzc);
    // derivative directions per function: comp q of 3 (one derivative) or 9 = 3 t + s
    const int d1 = NDA + NDB + NDC == 1 ? q : q / 3, d2 = q % 3;
    DTerms<NDA> ta;
    DTerms<NDB> tb;
    DTerms<NDC> tc;
    dterms<NDA>(lxa, lya, lza, d1, d2, ta);
    dterms<NDB>(lxb, lyb, lzb, d2, 0, tb);
    dterms<NDC>(lxc, lyc, lzc, KIND == HH_INT_IP1IP2 ? d2 : d1, d2, tc);
    double v = 0.0;
#pragma unroll
    for (int x = 0; x < (1 << NDA); ++x) {
      if (ta.c[x] == 0.0) continue;
      const int iA = set_index(A.loA, ta.x[x], ta.y[x], ta.z[x]);
#pragma unroll
      for (int y = 0; y < (1 << NDB); ++y) {
        if (tb.c[y] == 0.0) continue;
        const int iB = set_index(A.loB, tb.x[y], tb.y[y], tb.z[y]);
#pragma unroll
        for (int z = 0; z < (1 << NDC); ++z) {
          if (tc.c[z] == 0.0) continue;
          const int iC = set_index(A.loC, tc.x[z], tc.y[z], tc.z[z]);
          v += ta.c[x] * tb.c[y] * tc.c[z] * I[(iA * A.nB + iB) * A.nC + iC];
        }
      }
    }
    acc[e] = v;
  }
  __syncthreads();
  // to spherical (or Cartesian with libcint's s / p factors) and out
  const int nsa = A.cart ? ncA : 2 * A.la + 1, nsb = A.cart ? ncB : 2 * A.lb + 1, nsc = A.cart ? ncC : 2 * A.lc + 1;
  const int nout = nsa * nsb * nsc * nq;
  for (int e = tid; e < nout; e += 64) {
    const int qq = e % nq;
    const int rest = e / nq;
    const int mc = rest % nsc, mb = (rest / nsc) % nsb, ma = rest / (nsc * nsb);
    double v = 0.0;
    if (A.cart) {
      v = acc[((ma * ncB + mb) * ncC + mc) * nq + qq] * cart_fac(A.la) * (A.two_c ? 1.0 : cart_fac(A.lb)) *
          cart_fac(A.lc);
    } else {
      const double* ra = c_c2s + c2s_off(A.la) + ma * ncA;
      const double* rb = c_c2s + c2s_off(A.lb) + mb * ncB;
      const double* rc = c_c2s + c2s_off(A.lc) + mc * ncC;
      for (int ja = 0; ja < ncA; ++ja) {
        if (ra[ja] == 0.0) continue;
        for (int jb = 0; jb < ncB; ++jb) {
          const double wb = A.two_c ? 1.0 : rb[jb];
          if (wb == 0.0) continue;
          for (int jc = 0; jc < ncC; ++jc)
            v += ra[ja] * wb * rc[jc] * acc[((ja * ncB + jb) * ncC + jc) * nq + qq];
        }
      }
    }
    const int64_t ao_b = A.two_c ? 0 : sb.ao0 + mb;
    const int64_t cq = A.s2 * ((sc.ao0 + mc - A.c0) + A.s3 * (A.q0 + qq));
    A.out[(sa.ao0 + ma) + A.s1 * (ao_b + cq)] = v;
    if (A.sym_ab && ia_s != ib_s) {
      const int q = A.q0 + qq, qm = KIND == HH_INT_IPVIP1 ? 3 * (q % 3) + q / 3 : q;
      A.out[ao_b + A.s1 * ((sa.ao0 + ma) + A.s2 * ((sc.ao0 + mc - A.c0) + A.s3 * qm))] = v;
    }
  }
}

int kind_ncomp(int kind) {
  switch (kind) {
    case HH_INT_PLAIN: return 1;
    case HH_INT_IP1: case HH_INT_IP2: return 3;
    case HH_INT_IPIP1: case HH_INT_IPVIP1: case HH_INT_IP1IP2: case HH_INT_IPIP2: return 9;
    default: throw Error(HH_ERR_INVALID, "invalid argument: unknown integral kind");
  }
}

void kind_derivs(int kind, int* na, int* nb, int* nc) {
  *na = *nb = *nc = 0;
  switch (kind) {
    case HH_INT_IP1: *na = 1; break;
    case HH_INT_IP2: *nc = 1; break;
    case HH_INT_IPIP1: *na = 2; break;
    case HH_INT_IPVIP1: *na = 1; *nb = 1; break;
    case HH_INT_IP1IP2: *na = 1; *nc = 1; break;
    case HH_INT_IPIP2: *nc = 2; break;
    default: break;
  }
}

int tet_h(int n) { return n * (n + 1) * (n + 2) / 6; }

// Launch the triple kernel for every (la, lb, lc) group.  a / c lists: contracted shells per
// angular momentum; b likewise (ignored for 2-centre).
void launch_triples(const Cint* ma, const Cint* mb, const Cint* mc, const std::vector<int>& alist_all,
                    const std::vector<int>& blist_all, const std::vector<int>& clist_all, bool two_c, int kind,
                    bool cart, double* out, int64_t s1, int64_t s2, int64_t s3, int64_t c0, hipStream_t s,
                    const DevBuf& d_csa, const DevBuf& d_prima, const DevBuf& d_csb, const DevBuf& d_primb,
                    const DevBuf& d_csc, const DevBuf& d_primc) {
  upload_c2s();
  upload_boys();
  const int ncomp = kind_ncomp(kind);
  int na, nb, nc;
  kind_derivs(kind, &na, &nb, &nc);
  if (two_c) HH_REQUIRE(nb == 0, "2-centre integrals: no derivative on the unit function");
  // group shells by l
  auto group = [](const Cint* m, const std::vector<int>& list) {
    std::map<int, std::vector<int>> g;
    for (int i : list) g[m->cshells()[i].l].push_back(i);
    return g;
  };
  auto ga = group(ma, alist_all);
  auto gc = group(mc, clist_all);
  std::map<int, std::vector<int>> gb;
  if (two_c) gb[0] = {0};
  else gb = group(mb, blist_all);
  static int max_lds = 0;
  if (max_lds == 0) {
    int dev = 0;
    HH_HIP(hipGetDevice(&dev));
    HH_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const void* fns[] = {(const void*)int_triple_kernel<HH_INT_PLAIN>, (const void*)int_triple_kernel<HH_INT_IP1>,
                         (const void*)int_triple_kernel<HH_INT_IP2>, (const void*)int_triple_kernel<HH_INT_IPIP1>,
                         (const void*)int_triple_kernel<HH_INT_IPVIP1>, (const void*)int_triple_kernel<HH_INT_IP1IP2>,
                         (const void*)int_triple_kernel<HH_INT_IPIP2>};
    for (const void* f : fns) HH_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
  }
  std::vector<DevBuf*> keep;
  const bool sym = !two_c && ma == mb && alist_all == blist_all && ((na == 0 && nb == 0) || kind == HH_INT_IPVIP1);
  for (auto& [la, al] : ga)
    for (auto& [lb, bl] : gb)
      for (auto& [lc, cl] : gc) {
        if (sym && la < lb) continue;
        TripleArgs A{};
        A.csa = d_csa.as<CShell>(); A.prima = d_prima.as<double>();
        A.csb = two_c ? nullptr : d_csb.as<CShell>(); A.primb = two_c ? nullptr : d_primb.as<double>();
        A.csc = d_csc.as<CShell>(); A.primc = d_primc.as<double>();
        A.la = la; A.lb = lb; A.lc = lc; A.two_c = two_c ? 1 : 0; A.sym_ab = sym ? 1 : 0;
        A.nda = na; A.ndb = nb; A.ndc = nc; A.kind = kind; A.cart = cart ? 1 : 0;
        A.out = out; A.s1 = s1; A.s2 = s2; A.s3 = s3; A.c0 = c0;
        A.LA = la + na; A.LB = lb + nb; A.LC = lc + nc; A.LT = A.LA + A.LB + A.LC;
        // shifted sets: l = l0 + n, l0 + n - 2, .. >= 0 (the parity n derivatives reach)
        auto lowest = [](int l0, int n) { return l0 >= n ? l0 - n : (l0 + n) % 2; };
        auto set_size = [](int lo, int hi) {
          int k = 0;
          for (int l = lo; l <= hi; l += 2) k += (l + 1) * (l + 2) / 2;
          return k;
        };
        A.loA = lowest(la, na); A.loB = lowest(lb, nb); A.loC = lowest(lc, nc);
        A.nA = set_size(A.loA, A.LA);
        A.nB = set_size(A.loB, A.LB);
        A.nC = set_size(A.loC, A.LC);
        A.NR = tet_h(A.LT + 1);
        const int nEab = (A.LA + 1) * (A.LB + 1) * (A.LA + A.LB + 1);
        const int nEc = (A.LC + 1) * (A.LC + 1);
        const int ncarts = ncart(la) * ncart(lb) * ncart(lc);
        int off = 0;
        A.offR = off; off += 2 * A.NR;
        A.offE = off; off += 3 * nEab;
        A.offEc = off; off += 3 * nEc;
        A.offF = off; off += A.LT + 1;
        A.offI = off; off += A.nA * A.nB * A.nC;
        A.offSet = off; off += (A.nA + A.nB + A.nC + 1) / 2;
        const int fixed = off;
        // derivative components per launch so that the accumulator fits
        int qchunk = ncomp;
        while (qchunk > 1 && (size_t)(fixed + ncarts * qchunk) * 8 > (size_t)max_lds) qchunk = (qchunk + 1) / 2;
        HH_REQUIRE((size_t)(fixed + ncarts * qchunk) * 8 <= (size_t)max_lds,
                   "integral engine: angular momenta too high for the device's LDS");
        // device lists of this group
        auto* la_d = new DevBuf; auto* lb_d = new DevBuf; auto* lc_d = new DevBuf;
        keep.push_back(la_d); keep.push_back(lb_d); keep.push_back(lc_d);
        la_d->reserve(sizeof(int) * al.size());
        HH_HIP(hipMemcpyAsync(la_d->ptr, al.data(), sizeof(int) * al.size(), hipMemcpyHostToDevice, s));
        lb_d->reserve(sizeof(int) * bl.size());
        HH_HIP(hipMemcpyAsync(lb_d->ptr, bl.data(), sizeof(int) * bl.size(), hipMemcpyHostToDevice, s));
        lc_d->reserve(sizeof(int) * cl.size());
        HH_HIP(hipMemcpyAsync(lc_d->ptr, cl.data(), sizeof(int) * cl.size(), hipMemcpyHostToDevice, s));
        A.la_list = la_d->as<int>(); A.lb_list = lb_d->as<int>(); A.lc_list = lc_d->as<int>();
        A.na_list = (int)al.size(); A.nb_list = (int)bl.size(); A.nc_list = (int)cl.size();
        const int64_t nab = (int64_t)al.size() * (two_c ? 1 : (int64_t)bl.size());
        HH_REQUIRE(nab < (1LL << 31) && cl.size() < 65536, "integral engine: too many shells");
        for (int q0 = 0; q0 < ncomp; q0 += qchunk) {
          A.q0 = q0;
          A.q1 = std::min(ncomp, q0 + qchunk);
          A.offAcc = fixed;
          A.lds_doubles = fixed + ncarts * (A.q1 - A.q0);
          const dim3 grid((unsigned)nab, (unsigned)cl.size());
          const size_t lds = sizeof(double) * A.lds_doubles;
          switch (kind) {
            case HH_INT_IP1: hipLaunchKernelGGL(int_triple_kernel<HH_INT_IP1>, grid, dim3(64), lds, s, A); break;
            case HH_INT_IP2: hipLaunchKernelGGL(int_triple_kernel<HH_INT_IP2>, grid, dim3(64), lds, s, A); break;
            case HH_INT_IPIP1: hipLaunchKernelGGL(int_triple_kernel<HH_INT_IPIP1>, grid, dim3(64), lds, s, A); break;
            case HH_INT_IPVIP1: hipLaunchKernelGGL(int_triple_kernel<HH_INT_IPVIP1>, grid, dim3(64), lds, s, A); break;
            case HH_INT_IP1IP2: hipLaunchKernelGGL(int_triple_kernel<HH_INT_IP1IP2>, grid, dim3(64), lds, s, A); break;
            case HH_INT_IPIP2: hipLaunchKernelGGL(int_triple_kernel<HH_INT_IPIP2>, grid, dim3(64), lds, s, A); break;
            default: hipLaunchKernelGGL(int_triple_kernel<HH_INT_PLAIN>, grid, dim3(64), lds, s, A);
          }
          HH_HIP(hipGetLastError());
        }
      }
  HH_HIP(hipStreamSynchronize(s));  // the group lists are freed on return
  for (DevBuf* b : keep) delete b;
}

// out[nu (nu + 1) / 2 + mu + ntp p] = full[mu + nao (nu + nao p)], mu <= nu (packed upper triangle)
__global__ void pack_upper_kernel(const double* __restrict__ full, int64_t nao, int64_t np, double* __restrict__ out) {
  const int64_t ntp = nao * (nao + 1) / 2, total = ntp * np;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = e % ntp, p = e / ntp;
    int64_t nu = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (nu * (nu + 1) / 2 > t) --nu;
    while ((nu + 1) * (nu + 2) / 2 <= t) ++nu;
    const int64_t mu = t - nu * (nu + 1) / 2;
    out[e] = full[mu + nao * (nu + nao * p)];
  }
}

// ---------------------------------------------------------------------------------------------
// One-electron overlap and kinetic integrals <a| op |b> over contracted shells: per primitive
// pair the 1D overlaps S_d[i][j] = int x_A^i x_B^j e^{-alpha x_A^2 - beta x_B^2} dx by the
// Obara-Saika recursion from S[0][0] = sqrt(pi / p) e^{-mu X_AB^2}
//   S[i+1][j] = X_PA S[i][j] + (i S[i-1][j] + j S[i][j-1]) / 2p
//   S[i][j+1] = X_PB S[i][j] + (i S[i-1][j] + j S[i][j-1]) / 2p,
// then derivatives of either function by d/dx [x^i e^{-a x^2}] = i x^{i-1} e^{..} - 2 a x^{i+1} e^{..}
// and the kinetic operator as -1/2 sum_e d^2/de^2 on b.  One wave per contracted shell pair.

constexpr int P_LA = LMAXS + 2;  // highest a power: l + 2 (ipip on a)
constexpr int P_LB = LMAXS + 3;  // highest b power: l + 3 (ip on b and the kinetic d^2)

struct PairArgs {
  const CShell* cs;
  const double* prim;
  const int* alist;
  const int* blist;
  int na_list, nb_list;
  int la, lb, op, kind, cart;
  double* out;  // element (ia, ib, q) at out[(ao_a + ia) + nao ((ao_b + ib) + nao q)]
  int64_t nao;
};

// 1D factor: (d/dx)^ka of x^i e^{-al x^2} against (d/dx)^kb of x^j e^{-be x^2}
__device__ __forceinline__ double pair_1d(const double* S, int i, int j, int ka, int kb, double al, double be) {
  // ka <= 2: at most 3 distinct powers (4 before merging); kb <= 3: at most 4 (6 before merging)
  double ca[4], cb[5];
  int pa[4], pb[5];
  int na = 1, nb = 1;
  ca[0] = 1.0; pa[0] = i;
  cb[0] = 1.0; pb[0] = j;
  for (int s = 0; s < ka; ++s) {
    double c2[4];
    int p2[4];
    int m = 0;
    for (int k = 0; k < na; ++k) {
      if (pa[k] > 0) { c2[m] = ca[k] * pa[k]; p2[m] = pa[k] - 1; ++m; }
      c2[m] = -2.0 * al * ca[k]; p2[m] = pa[k] + 1; ++m;
    }
    // merge equal powers (at most 3 distinct after two steps)
    na = 0;
    for (int k = 0; k < m; ++k) {
      int f = -1;
      for (int t = 0; t < na; ++t) if (pa[t] == p2[k]) f = t;
      if (f >= 0) ca[f] += c2[k]; else { ca[na] = c2[k]; pa[na] = p2[k]; ++na; }
    }
  }
  for (int s = 0; s < kb; ++s) {
    double c2[8];
    int p2[8];
    int m = 0;
    for (int k = 0; k < nb; ++k) {
      if (pb[k] > 0) { c2[m] = cb[k] * pb[k]; p2[m] = pb[k] - 1; ++m; }
      c2[m] = -2.0 * be * cb[k]; p2[m] = pb[k] + 1; ++m;
    }
    nb = 0;
    for (int k = 0; k < m; ++k) {
      int f = -1;
      for (int t = 0; t < nb; ++t) if (pb[t] == p2[k]) f = t;
      if (f >= 0) cb[f] += c2[k]; else { cb[nb] = c2[k]; pb[nb] = p2[k]; ++nb; }
    }
  }
  double v = 0.0;
  for (int x = 0; x < na; ++x)
    for (int y = 0; y < nb; ++y) v += ca[x] * cb[y] * S[pa[x] * (P_LB + 1) + pb[y]];
  return v;
}

__global__ __launch_bounds__(64) void int_pair_kernel(PairArgs A) {
  __shared__ double S1[3][(P_LA + 1) * (P_LB + 1)];
  __shared__ double acc[21 * 21 * 9];
  const int tid = threadIdx.x;
  const CShell sa = A.cs[A.alist[blockIdx.x % A.na_list]];
  const CShell sb = A.cs[A.blist[blockIdx.x / A.na_list]];
  const int ncA = ncart(A.la), ncB = ncart(A.lb);
  const int nq = A.kind == HH_INT_PLAIN ? 1 : (A.kind == HH_INT_IP1 ? 3 : 9);
  const int nacc = ncA * ncB * nq;
  int nda = 0, ndb = 0;
  if (A.kind == HH_INT_IP1) nda = 1;
  if (A.kind == HH_INT_IPIP1) nda = 2;
  if (A.kind == HH_INT_IPVIP1) { nda = 1; ndb = 1; }
  const int LA = A.la + nda, LB = A.lb + ndb + (A.op == 1 ? 2 : 0);
  for (int e = tid; e < nacc; e += 64) acc[e] = 0.0;
  const double* exa = A.prim + sa.prim_off;
  const double* exb = A.prim + sb.prim_off;
  const double AB[3] = {sa.x - sb.x, sa.y - sb.y, sa.z - sb.z};
  const double ab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
  const double Ac[3] = {sa.x, sa.y, sa.z}, Bc[3] = {sb.x, sb.y, sb.z};
  for (int pa = 0; pa < sa.nprim; ++pa)
    for (int pb = 0; pb < sb.nprim; ++pb) {
      const double al = exa[pa], be = exb[pb];
      const double p = al + be, mu = al * be / p;
      if (mu * ab2 > 40.0) continue;  // e^{-mu AB^2} < 4e-18
      __syncthreads();
      if (tid < 3) {
        double* S = S1[tid];
        const double P = (al * Ac[tid] + be * Bc[tid]) / p;
        const double xpa = P - Ac[tid], xpb = P - Bc[tid], h = 0.5 / p;
        S[0] = sqrt(M_PI / p) * exp(-mu * AB[tid] * AB[tid]);
        for (int i = 0; i <= LA; ++i) {
          if (i > 0) {
            const double sm = i >= 2 ? S[(i - 2) * (P_LB + 1)] : 0.0;
            S[i * (P_LB + 1)] = xpa * S[(i - 1) * (P_LB + 1)] + h * (i - 1) * sm;
          }
          for (int j = 1; j <= LB; ++j) {
            const double a1 = i > 0 ? S[(i - 1) * (P_LB + 1) + j - 1] : 0.0;
            const double b1 = j >= 2 ? S[i * (P_LB + 1) + j - 2] : 0.0;
            S[i * (P_LB + 1) + j] = xpb * S[i * (P_LB + 1) + j - 1] + h * (i * a1 + (j - 1) * b1);
          }
        }
      }
      __syncthreads();
      const double wgt = exa[sa.nprim + pa] * exb[sb.nprim + pb];
      for (int e = tid; e < nacc; e += 64) {
        const int q = e % nq;
        const int rest = e / nq;
        const int jb = rest % ncB, ja = rest / ncB;
        int la3[3] = {0, 0, 0}, lb3[3] = {0, 0, 0};
        {
          int k = 0;
          for (int lx = A.la; lx >= 0; --lx)
            for (int ly = A.la - lx; ly >= 0; --ly, ++k)
              if (k == ja) { la3[0] = lx; la3[1] = ly; la3[2] = A.la - lx - ly; }
          k = 0;
          for (int lx = A.lb; lx >= 0; --lx)
            for (int ly = A.lb - lx; ly >= 0; --ly, ++k)
              if (k == jb) { lb3[0] = lx; lb3[1] = ly; lb3[2] = A.lb - lx - ly; }
        }
        int ka[3] = {0, 0, 0}, kb[3] = {0, 0, 0};
        switch (A.kind) {
          case HH_INT_IP1: ka[q] += 1; break;
          case HH_INT_IPIP1: ka[q / 3] += 1; ka[q % 3] += 1; break;
          case HH_INT_IPVIP1: ka[q / 3] += 1; kb[q % 3] += 1; break;
          default: break;
        }
        double v;
        if (A.op == 0) {
          v = 1.0;
          for (int d = 0; d < 3; ++d) v *= pair_1d(S1[d], la3[d], lb3[d], ka[d], kb[d], al, be);
        } else {
          v = 0.0;
          for (int ed = 0; ed < 3; ++ed) {
            double t = 1.0;
            for (int d = 0; d < 3; ++d) t *= pair_1d(S1[d], la3[d], lb3[d], ka[d], kb[d] + (d == ed ? 2 : 0), al, be);
            v += t;
          }
          v *= -0.5;
        }
        acc[(ja * ncB + jb) * nq + q] += wgt * v;
      }
    }
  __syncthreads();
  const int nsa = A.cart ? ncA : 2 * A.la + 1, nsb = A.cart ? ncB : 2 * A.lb + 1;
  for (int e = tid; e < nsa * nsb * nq; e += 64) {
    const int q = e % nq;
    const int rest = e / nq;
    const int mb = rest % nsb, ma = rest / nsb;
    double v = 0.0;
    if (A.cart) {
      v = acc[(ma * ncB + mb) * nq + q] * cart_fac(A.la) * cart_fac(A.lb);
    } else {
      const double* ra = c_c2s + c2s_off(A.la) + ma * ncA;
      const double* rb = c_c2s + c2s_off(A.lb) + mb * ncB;
      for (int ja = 0; ja < ncA; ++ja) {
        if (ra[ja] == 0.0) continue;
        for (int jb = 0; jb < ncB; ++jb)
          if (rb[jb] != 0.0) v += ra[ja] * rb[jb] * acc[(ja * ncB + jb) * nq + q];
      }
    }
    A.out[(sa.ao0 + ma) + A.nao * ((sb.ao0 + mb) + A.nao * (int64_t)q)] = v;
  }
}

// out[i + n2 q] = sum_c w[c] in[i + n2 (c + nc q)], i < n2
__global__ void weighted_sum_kernel(const double* __restrict__ in, int64_t n2, int nc, int nq,
                                    const double* __restrict__ w, double* __restrict__ out) {
  const int64_t total = n2 * nq;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % n2, q = e / n2;
    double v = 0.0;
    for (int c = 0; c < nc; ++c) v += w[c] * in[i + n2 * (c + (int64_t)nc * q)];
    out[e] = v;
  }
}

}  // namespace

void Cint::int3c2e_s2ij(const Cint& mol, const Cint& aux, int sh0, int sh1, double* out, hipStream_t s) {
  HH_REQUIRE(out, "int3c2e: null output");
  HH_REQUIRE(0 <= sh0 && sh0 <= sh1 && sh1 <= aux.nbas(), "int3c2e: auxiliary shell range");
  const int64_t n01 = aux.bas_ao_[sh1] - aux.bas_ao_[sh0];
  if (n01 == 0) return;
  DevBuf full;
  full.reserve(sizeof(double) * mol.nao_ * mol.nao_ * n01);
  int3c2e(mol, aux, HH_INT_PLAIN, sh0, sh1, full.as<double>(), s);
  const int64_t total = mol.nao_ * (mol.nao_ + 1) / 2 * n01;
  hipLaunchKernelGGL(pack_upper_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(total, 256), 65536)), dim3(256), 0,
                     s, full.as<double>(), mol.nao_, n01, out);
  HH_HIP(hipGetLastError());
  HH_HIP(hipStreamSynchronize(s));  // the full block is freed on return
}

Cint::Cint(const int32_t* atm, int natm, const int32_t* bas, int nbas, const double* env, int64_t nenv, bool cart)
    : cart_(cart) {
  HH_REQUIRE(atm && bas && env && natm > 0 && nbas > 0 && nenv > 0, "cint: empty molecule");
  for (int a = 0; a < natm; ++a) {
    const int pc = atm[6 * a + 1];
    HH_REQUIRE(pc >= 0 && pc + 3 <= nenv, "cint: atom coordinates outside env");
    atom_z_.push_back((double)atm[6 * a]);
    for (int d = 0; d < 3; ++d) atom_xyz_.push_back(env[pc + d]);
  }
  bas_cs_.push_back(0);
  bas_ao_.push_back(0);
  for (int i = 0; i < nbas; ++i) {
    const int32_t* b = bas + 8 * i;
    const int atom = b[0], l = b[1], nprim = b[2], nctr = b[3], pexp = b[5], pcoef = b[6];
    HH_REQUIRE(atom >= 0 && atom < natm, "cint: shell atom out of range");
    HH_REQUIRE(l >= 0 && l <= LMAX, "cint: angular momentum above 5");
    HH_REQUIRE(nprim >= 1 && nctr >= 1, "cint: empty shell");
    HH_REQUIRE(pexp >= 0 && pexp + nprim <= nenv && pcoef >= 0 && pcoef + (int64_t)nprim * nctr <= nenv,
               "cint: shell data outside env");
    const int pc = atm[6 * atom + 1];
    HH_REQUIRE(pc >= 0 && pc + 3 <= nenv, "cint: atom coordinates outside env");
    const int ncomp = cart ? ncart(l) : 2 * l + 1;
    for (int c = 0; c < nctr; ++c) {
      CShell s{};
      s.x = env[pc]; s.y = env[pc + 1]; s.z = env[pc + 2];
      s.l = l;
      s.nprim = nprim;
      s.prim_off = (int)prim_h_.size();
      s.ao0 = (int)nao_;
      for (int p = 0; p < nprim; ++p) prim_h_.push_back(env[pexp + p]);
      for (int p = 0; p < nprim; ++p) prim_h_.push_back(env[pcoef + p + (int64_t)nprim * c]);
      cs_.push_back(s);
      nao_ += ncomp;
      lmax_ = std::max(lmax_, l);
    }
    bas_cs_.push_back((int)cs_.size());
    bas_ao_.push_back(nao_);
  }
  // AO range per atom; an atom without shells gets an empty range at its position
  atom_ao_.assign(2 * (size_t)natm, -1);
  for (int i = 0; i < nbas; ++i) {
    const int atom = bas[8 * i];
    int64_t& lo = atom_ao_[2 * atom];
    int64_t& hi = atom_ao_[2 * atom + 1];
    if (lo < 0) {
      lo = bas_ao_[i];
      hi = bas_ao_[i + 1];
    } else {
      hi = std::max(hi, bas_ao_[i + 1]);
      lo = std::min(lo, bas_ao_[i]);
    }
  }
  {
    int64_t prev = 0, count = 0;
    for (int a = 0; a < natm; ++a) {
      if (atom_ao_[2 * a] < 0) {
        atom_ao_[2 * a] = atom_ao_[2 * a + 1] = prev;
      } else {
        prev = atom_ao_[2 * a + 1];
      }
      count += atom_ao_[2 * a + 1] - atom_ao_[2 * a];
    }
    // grouped by atom: the ranges tile [0, nao) without overlap
    atom_contiguous_ = count == nao_;
  }
  d_cs_.reserve(sizeof(CShell) * cs_.size());
  HH_HIP(hipMemcpy(d_cs_.ptr, cs_.data(), sizeof(CShell) * cs_.size(), hipMemcpyHostToDevice));
  d_prim_.reserve(sizeof(double) * prim_h_.size());
  HH_HIP(hipMemcpy(d_prim_.ptr, prim_h_.data(), sizeof(double) * prim_h_.size(), hipMemcpyHostToDevice));
}

void Cint::eval_gto(int deriv, const double* coords, int64_t ng, double* out, hipStream_t s) const {
  HH_REQUIRE(deriv >= 0 && deriv <= 3, "eval_gto: deriv 0..3");
  HH_REQUIRE(ng >= 0 && (ng == 0 || (coords && out)), "eval_gto: null pointer");
  if (ng == 0) return;
  upload_c2s();
  std::vector<std::vector<int>> byl(LMAX + 1);
  for (int i = 0; i < (int)cs_.size(); ++i) byl[cs_[i].l].push_back(i);
  std::vector<int> flat;
  std::vector<int> start(LMAX + 2, 0);
  for (int l = 0; l <= LMAX; ++l) {
    start[l] = (int)flat.size();
    flat.insert(flat.end(), byl[l].begin(), byl[l].end());
  }
  start[LMAX + 1] = (int)flat.size();
  DevBuf lst;
  lst.reserve(sizeof(int) * flat.size());
  HH_HIP(hipMemcpyAsync(lst.ptr, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice, s));
  const unsigned gx = (unsigned)ceil_div(ng, 256);
  HH_REQUIRE(ceil_div(ng, 256) < (1LL << 31), "eval_gto: too many points");
#define HH_EVAL_L(L)                                                                                               \
  if (!byl[L].empty()) {                                                                                           \
    const dim3 grid(gx, (unsigned)byl[L].size());                                                                  \
    const int* lp = lst.as<int>() + start[L];                                                                      \
    switch (deriv) {                                                                                               \
      case 0: hipLaunchKernelGGL((eval_gto_kernel<L, 0>), grid, dim3(256), 0, s, d_cs_.as<CShell>(), lp,           \
                                 d_prim_.as<double>(), coords, ng, nao_, (int)cart_, out); break;                  \
      case 1: hipLaunchKernelGGL((eval_gto_kernel<L, 1>), grid, dim3(256), 0, s, d_cs_.as<CShell>(), lp,           \
                                 d_prim_.as<double>(), coords, ng, nao_, (int)cart_, out); break;                  \
      case 2: hipLaunchKernelGGL((eval_gto_kernel<L, 2>), grid, dim3(256), 0, s, d_cs_.as<CShell>(), lp,           \
                                 d_prim_.as<double>(), coords, ng, nao_, (int)cart_, out); break;                  \
      default: hipLaunchKernelGGL((eval_gto_kernel<L, 3>), grid, dim3(256), 0, s, d_cs_.as<CShell>(), lp,          \
                                  d_prim_.as<double>(), coords, ng, nao_, (int)cart_, out); break;                 \
    }                                                                                                              \
    HH_HIP(hipGetLastError());                                                                                     \
  }
  HH_EVAL_L(0) HH_EVAL_L(1) HH_EVAL_L(2) HH_EVAL_L(3) HH_EVAL_L(4) HH_EVAL_L(5)
#undef HH_EVAL_L
  HH_HIP(hipStreamSynchronize(s));  // the shell list is freed on return
}

void Cint::int2c2e(int kind, int sh0, int sh1, double* out, hipStream_t s) const {
  HH_REQUIRE(kind == HH_INT_PLAIN || kind == HH_INT_IP1 || kind == HH_INT_IPIP1 || kind == HH_INT_IP1IP2,
             "int2c2e: kinds PLAIN, IP1, IPIP1, IP1IP2");
  HH_REQUIRE(out, "int2c2e: null output");
  HH_REQUIRE(0 <= sh0 && sh0 <= sh1 && sh1 <= nbas(), "int2c2e: shell range");
  if (sh0 == sh1) return;
  std::vector<int> rows, cols;
  for (int i = bas_cs_[sh0]; i < bas_cs_[sh1]; ++i) rows.push_back(i);
  for (int i = 0; i < (int)cs_.size(); ++i) cols.push_back(i);
  const int64_t n01 = bas_ao_[sh1] - bas_ao_[sh0];
  // element (a, c, q) at out[(a - ao(sh0)) + n01 (c + nao q)]: the a offset is folded into the
  // output pointer (rows start at AO ao(sh0))
  double* o = out - bas_ao_[sh0];
  launch_triples(this, nullptr, this, rows, {}, cols, true, kind, cart_, o, n01, 1, nao_, 0, s, d_cs_, d_prim_, d_cs_,
                 d_prim_, d_cs_, d_prim_);
}

void Cint::int3c2e(const Cint& mol, const Cint& aux, int kind, int sh0, int sh1, double* out, hipStream_t s) {
  kind_ncomp(kind);
  HH_REQUIRE(out, "int3c2e: null output");
  HH_REQUIRE(mol.cart_ == aux.cart_, "int3c2e: both bases spherical or both Cartesian");
  HH_REQUIRE(0 <= sh0 && sh0 <= sh1 && sh1 <= aux.nbas(), "int3c2e: auxiliary shell range");
  if (sh0 == sh1) return;
  std::vector<int> a, c;
  for (int i = 0; i < (int)mol.cs_.size(); ++i) a.push_back(i);
  for (int i = aux.bas_cs_[sh0]; i < aux.bas_cs_[sh1]; ++i) c.push_back(i);
  const int64_t n01 = aux.bas_ao_[sh1] - aux.bas_ao_[sh0];
  launch_triples(&mol, &mol, &aux, a, a, c, false, kind, mol.cart_, out, mol.nao_, mol.nao_, n01, aux.bas_ao_[sh0], s,
                 mol.d_cs_, mol.d_prim_, mol.d_cs_, mol.d_prim_, aux.d_cs_, aux.d_prim_);
}


void Cint::int1e(int op, int kind, double* out, hipStream_t s) const {
  HH_REQUIRE(op == HH_INT1E_OVLP || op == HH_INT1E_KIN, "int1e: operator OVLP or KIN here");
  HH_REQUIRE(kind == HH_INT_PLAIN || kind == HH_INT_IP1 || kind == HH_INT_IPIP1 || kind == HH_INT_IPVIP1,
             "int1e: kinds PLAIN, IP1, IPIP1, IPVIP1");
  HH_REQUIRE(out, "int1e: null output");
  upload_c2s();
  std::map<int, std::vector<int>> g;
  for (int i = 0; i < (int)cs_.size(); ++i) g[cs_[i].l].push_back(i);
  std::vector<DevBuf*> keep;
  for (auto& [la, al] : g) {
    auto* d = new DevBuf;
    keep.push_back(d);
    d->reserve(sizeof(int) * al.size());
    HH_HIP(hipMemcpyAsync(d->ptr, al.data(), sizeof(int) * al.size(), hipMemcpyHostToDevice, s));
  }
  int ia = 0;
  for (auto& [la, al] : g) {
    int ib = 0;
    for (auto& [lb, bl] : g) {
      PairArgs A{};
      A.cs = d_cs_.as<CShell>();
      A.prim = d_prim_.as<double>();
      A.alist = keep[ia]->as<int>();
      A.blist = keep[ib]->as<int>();
      A.na_list = (int)al.size();
      A.nb_list = (int)bl.size();
      A.la = la;
      A.lb = lb;
      A.op = op == HH_INT1E_KIN ? 1 : 0;
      A.kind = kind;
      A.cart = cart_ ? 1 : 0;
      A.out = out;
      A.nao = nao_;
      const int64_t nab = (int64_t)al.size() * (int64_t)bl.size();
      HH_REQUIRE(nab < (1LL << 31), "int1e: too many shell pairs");
      hipLaunchKernelGGL(int_pair_kernel, dim3((unsigned)nab), dim3(64), 0, s, A);
      HH_HIP(hipGetLastError());
      ++ib;
    }
    ++ia;
  }
  HH_HIP(hipStreamSynchronize(s));  // the shell lists are freed on return
  for (DevBuf* b : keep) delete b;
}

const Cint& Cint::point_charges() const {
  if (!points_) {
    // one s function per nucleus, c e^{-g r^2} times the s factor 1 / (2 sqrt(pi)), normalised to
    // unit charge: c = (g / pi)^(3/2) 2 sqrt(pi); with g = 1e20 the Coulomb integral (ab|c) equals
    // <a| 1 / |r - C| |b> to a relative 1e-16 p (p the exponent sum of the pair)
    const double g = 1e20;
    const int natm = (int)atom_z_.size();
    std::vector<int32_t> atm(6 * natm, 0), bas(8 * natm, 0);
    std::vector<double> env(20, 0.0);
    for (int a = 0; a < natm; ++a) {
      atm[6 * a + 1] = (int32_t)env.size();
      env.push_back(atom_xyz_[3 * a]);
      env.push_back(atom_xyz_[3 * a + 1]);
      env.push_back(atom_xyz_[3 * a + 2]);
    }
    const int pexp = (int)env.size();
    env.push_back(g);
    env.push_back(pow(g / M_PI, 1.5) * 2.0 * sqrt(M_PI));
    for (int a = 0; a < natm; ++a) {
      int32_t* b = &bas[8 * a];
      b[0] = a; b[1] = 0; b[2] = 1; b[3] = 1; b[5] = pexp; b[6] = pexp + 1;
    }
    points_ = std::make_unique<Cint>(atm.data(), natm, bas.data(), natm, env.data(), (int64_t)env.size(), cart_);
  }
  return *points_;
}

void Cint::int1e_rinv_all(int kind, double* out, hipStream_t s) const {
  HH_REQUIRE(kind == HH_INT_PLAIN || kind == HH_INT_IP1 || kind == HH_INT_IPIP1 || kind == HH_INT_IPVIP1,
             "int1e: kinds PLAIN, IP1, IPIP1, IPVIP1");
  HH_REQUIRE(out, "int1e: null output");
  int3c2e(*this, point_charges(), kind, 0, natm(), out, s);
}

void Cint::int1e_nuc(int kind, int atom, double* out, hipStream_t s) const {
  HH_REQUIRE(kind == HH_INT_PLAIN || kind == HH_INT_IP1 || kind == HH_INT_IPIP1 || kind == HH_INT_IPVIP1,
             "int1e: kinds PLAIN, IP1, IPIP1, IPVIP1");
  HH_REQUIRE(out, "int1e: null output");
  const int natm = (int)atom_z_.size();
  HH_REQUIRE(atom >= -1 && atom < natm, "int1e: nucleus index out of range");
  const Cint& pc = point_charges();
  if (atom >= 0) {
    // <a| 1 / |r - R_atom| |b>: element (a, b, 0, q) of the 3-centre block of one point
    int3c2e(*this, pc, kind, atom, atom + 1, out, s);
    return;
  }
  const int nq = kind == HH_INT_PLAIN ? 1 : (kind == HH_INT_IP1 ? 3 : 9);
  const int64_t n2 = nao_ * nao_;
  DevBuf tmp, w;
  tmp.reserve(sizeof(double) * n2 * natm * nq);
  int3c2e(*this, pc, kind, 0, natm, tmp.as<double>(), s);
  std::vector<double> mz(natm);
  for (int a = 0; a < natm; ++a) mz[a] = -atom_z_[a];
  w.reserve(sizeof(double) * natm);
  HH_HIP(hipMemcpyAsync(w.ptr, mz.data(), sizeof(double) * natm, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(weighted_sum_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(n2 * nq, 256), 65536)), dim3(256),
                     0, s, tmp.as<double>(), n2, natm, nq, w.as<double>(), out);
  HH_HIP(hipGetLastError());
  HH_HIP(hipStreamSynchronize(s));  // tmp and w are freed on return
}
}  // namespace hh
