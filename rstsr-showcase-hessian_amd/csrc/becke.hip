// Becke partition on the device (reference: src/numint_matmul/becke_partition.rs).
//
// The grid is processed in generating-atom order (ByAtom boundaries, or the ByGrid attribution
// sorted by atom on the host), so every grid point knows its atom G by a search over the
// boundaries.
//
// w, dw: one thread per grid point; the switch products P_A live in LDS columns.  dw[A] for
// A != G is the fixed-grid quotient-rule derivative, assembled per row A from the pairs (A, N);
// the row of G is minus the sum of the others (the translation-invariance fix,
// becke_partition.rs:1119-1152).
//
// ddw / ddc: the fixed-grid second derivative of q = P_G / Z, weighted by c~ = wquad * c, is
//   S = sum_M gam_M (l_M l_M^T + d2 log P_M) - c~/Z (dq dZ^T + dZ dq^T),
//   gam_M = c~ P_M ([M == G] - q) / Z,  l_M = d log P_M,  dZ = sum_M P_M l_M,
//   dq = (P_G l_G - q dZ) / Z,
// which is the reference's quotient rule with the rank-1 cross terms C = sum_M P_M l_M l_M^T
// and P_G l_G l_G^T folded into one weighted sum (becke_partition.rs:1155-1300).  d2 log P_M only
// has the pair blocks of log s_MN.  The sum over the grid points of one generating atom is taken
// before the translation fix T S T^T (T: the row of G becomes minus the sum of the other rows),
// since the fix is linear: one workgroup accumulates one chunk of an atom's points, keeping the
// per-point tables l_M (3 natm^2 values) in LDS, each thread owning 3x3 atom-pair blocks.
#include "becke.hpp"

#include <algorithm>
#include <cmath>

#include "elementwise.hpp"

namespace hh {
namespace {

constexpr double INVTOL = 1e-14;  // becke_partition.rs:24

struct Sw {
  double f, df, ddf;
};

// f = p^h(nu), nu = mu + a (1 - mu^2), with derivatives with respect to nu (becke_partition.rs:1575-1590)
template <int ORDER>
__device__ inline Sw switch_eval(double mu, double a, int h) {
  const double nu = mu + (1.0 - mu * mu) * a;
  double f = nu, df = 1.0, ddf = 0.0;
  for (int k = 0; k < h; ++k) {
    const double g = 1.5 * (1.0 - f * f);
    if (ORDER >= 2) ddf = -3.0 * f * df * df + g * ddf;
    if (ORDER >= 1) df = g * df;
    f = (1.5 - 0.5 * f * f) * f;
  }
  return {f, df, ddf};
}

// table layout: R [3 natm], adj [natm^2], irab [natm^2] (1 / |R_AB|, 0 on the diagonal),
// U [natm^2, 3] ((R_A - R_B) / |R_AB|), rab [natm^2] (|R_AB|)
struct Tab {
  const double *R, *adj, *irab, *U, *rab;
  __device__ Tab(const double* t, int64_t natm)
      : R(t), adj(t + 3 * natm), irab(t + 3 * natm + natm * natm), U(t + 3 * natm + 2 * natm * natm),
        rab(t + 3 * natm + 5 * natm * natm) {}
};

// generating atom of sorted index i: the last A with bounds[A] <= i
__device__ inline int64_t atom_of(const int64_t* bounds, int64_t natm, int64_t i) {
  int64_t lo = 0, hi = natm;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (bounds[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// ---- w and dw, one thread per grid point (becke_partition.rs:796-863, 865-960, 1102-1153) ----
// LDS: dist [natm][bs], P [natm][bs]
template <bool DW>
__global__ void becke_w_dw_kernel(const double* __restrict__ coords, const double* __restrict__ wquad,
                                  const int64_t* __restrict__ perm, const int64_t* __restrict__ bounds, int64_t n,
                                  const double* __restrict__ tab, int64_t natm, int h, double* __restrict__ w_out,
                                  double* __restrict__ dw_out, int64_t ng) {
  extern __shared__ double sm[];
  const int bs = blockDim.x, tid = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * bs + tid;
  if (i >= n) return;
  const Tab T(tab, natm);
  double* dist = sm;
  double* P = sm + natm * bs;
  const int64_t g = perm ? perm[i] : i;
  const int64_t G = atom_of(bounds, natm, i);
  const double r0 = coords[3 * g], r1 = coords[3 * g + 1], r2 = coords[3 * g + 2];
  for (int64_t A = 0; A < natm; ++A) {
    const double d0 = T.R[3 * A] - r0, d1 = T.R[3 * A + 1] - r1, d2 = T.R[3 * A + 2] - r2;
    dist[A * bs + tid] = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    P[A * bs + tid] = 1.0;
  }
  for (int64_t A = 0; A < natm; ++A)
    for (int64_t B = 0; B < A; ++B) {
      const double mu = (dist[A * bs + tid] - dist[B * bs + tid]) * T.irab[A * natm + B];
      const double f = switch_eval<0>(mu, T.adj[A * natm + B], h).f;
      P[A * bs + tid] *= 0.5 * (1.0 - f);
      P[B * bs + tid] *= 0.5 * (1.0 + f);
    }
  double Z = 0.0;
  for (int64_t A = 0; A < natm; ++A) Z += P[A * bs + tid];
  const double Pg = P[G * bs + tid];
  const double wq = wquad[g];
  if (w_out) w_out[g] = wq * Pg / Z;
  if (!DW) return;
  const double q = Pg / Z, fac = wq / Z;
  double sum0 = 0.0, sum1 = 0.0, sum2 = 0.0;
  for (int64_t A = 0; A < natm; ++A) {
    const double dA = dist[A * bs + tid], PA = P[A * bs + tid];
    const double e0 = (T.R[3 * A] - r0) / dA, e1 = (T.R[3 * A + 1] - r1) / dA, e2 = (T.R[3 * A + 2] - r2) / dA;
    double z0 = 0.0, z1 = 0.0, z2 = 0.0, p0 = 0.0, p1 = 0.0, p2 = 0.0;
    for (int64_t N = 0; N < natm; ++N) {
      if (N == A) continue;
      const double ir = T.irab[A * natm + N], a = T.adj[A * natm + N];
      const double mu = (dA - dist[N * bs + tid]) * ir;
      const Sw s = switch_eval<1>(mu, a, h);
      const double dsA = -0.5 * s.df * (1.0 - 2.0 * mu * a);
      const double lA = dsA / fmax(0.5 * (1.0 - s.f), INVTOL);
      const double lN = -dsA / fmax(0.5 * (1.0 + s.f), INVTOL);
      const double PN = P[N * bs + tid];
      const double cZ = PA * lA + PN * lN;
      const double cP = (A == G ? PA * lA : 0.0) + (N == G ? PN * lN : 0.0);
      const double* U = T.U + 3 * (A * natm + N);
      const double u0 = (e0 - mu * U[0]) * ir, u1 = (e1 - mu * U[1]) * ir, u2 = (e2 - mu * U[2]) * ir;
      z0 += cZ * u0; z1 += cZ * u1; z2 += cZ * u2;
      p0 += cP * u0; p1 += cP * u1; p2 += cP * u2;
    }
    const double v0 = fac * (p0 - q * z0), v1 = fac * (p1 - q * z1), v2 = fac * (p2 - q * z2);
    if (A != G) {
      dw_out[(3 * A) * ng + g] = v0;
      dw_out[(3 * A + 1) * ng + g] = v1;
      dw_out[(3 * A + 2) * ng + g] = v2;
      sum0 += v0; sum1 += v1; sum2 += v2;
    }
  }
  dw_out[(3 * G) * ng + g] = -sum0;
  dw_out[(3 * G + 1) * ng + g] = -sum1;
  dw_out[(3 * G + 2) * ng + g] = -sum2;
}

// ---- contracted second derivative, one workgroup per chunk of one atom's points ----
// chunks [nchunk][3] = (first sorted index, end, G); part [nchunk][n3 * n3] receives S.
// Per-point LDS record (doubles): dist [natm], 1 / dist [natm], P [natm], e [natm][3],
// L [natm][n3], gam [natm], dZ [n3], dq [n3], sc [4] = (Z, q, c~).
struct Rec {
  int natm, n3, stride;
  int o_id, o_P, o_e, o_L, o_gam, o_dZ, o_dq, o_sc;
  __host__ __device__ Rec(int natm_) : natm(natm_), n3(3 * natm_) {
    o_id = natm;
    o_P = 2 * natm;
    o_e = 3 * natm;
    o_L = 6 * natm;
    o_gam = o_L + natm * n3;
    o_dZ = o_gam + natm;
    o_dq = o_dZ + n3;
    o_sc = o_dq + n3;
    stride = o_sc + 4;
  }
};

// Work units of a workgroup: units [0, nblk) own the atom-pair blocks (A <= B): the rank-natm
// and rank-2 parts and, for A != B, the pair (A, B); units [nblk, nblk + natm * ngrp) carry the
// pairs (A, N), N in a group of DIAG_GRP atoms, of the diagonal block (A, A) (its natm - 1 pairs
// would otherwise serialise on one thread).
constexpr int DIAG_GRP = 3;

// Pair part of one block from the pair (A, N) at one point, A in the first role (the reference
// pair loop, becke_partition.rs:960-1090): coefficient (gam_A ddl_A + gam_N ddl_N) on
// d mu / dR_A (x) d mu / dR_X and (gam_A l_A + gam_N l_N) on d2 mu / dR_A dR_X, X = N (off-diagonal
// block) or X = A (diagonal block).
template <bool DIAG>
__device__ inline void pair_block(const Tab& T, int natm, int h, const double* rec, const Rec& L, int A, int N,
                                  double* a9) {
  const double* e = rec + L.o_e;
  const double* gam = rec + L.o_gam;
  const int an = A * natm + N;
  const double ir = T.irab[an], a = T.adj[an], gab = T.rab[an];
  const double dA = rec[A], dN = rec[N];
  const double fab = dA - dN, mu = fab * ir;
  const Sw s = switch_eval<2>(mu, a, h);
  const double sAs = fmax(0.5 * (1.0 - s.f), INVTOL), sNs = fmax(0.5 * (1.0 + s.f), INVTOL);
  const double inv = 1.0 / (sAs * sNs);
  const double isA = sNs * inv, isN = sAs * inv;
  const double dmu_nu = 1.0 - 2.0 * mu * a;
  const double dsA = -0.5 * s.df * dmu_nu;
  const double lA = dsA * isA, lN = -dsA * isN;
  const double ddsA = -0.5 * s.ddf * dmu_nu * dmu_nu + s.df * a;
  const double ddlA = ddsA * isA - lA * lA, ddlN = -ddsA * isN - lN * lN;
  const double cdd = gam[A] * ddlA + gam[N] * ddlN;
  const double c1 = gam[A] * lA + gam[N] * lN;
  const double* U = T.U + 3 * an;
  const double* eA = e + 3 * A;
  const double* eN = e + 3 * N;
  double rA[3], rX[3];
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    rA[t] = (eA[t] - mu * U[t]) * ir;
    rX[t] = DIAG ? rA[t] : (-eN[t] + mu * U[t]) * ir;
  }
  const double ir2 = ir * ir, ir3 = ir2 * ir, idA = rec[L.o_id + A];
#pragma unroll
  for (int t = 0; t < 3; ++t)
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const double dl = t == u ? 1.0 : 0.0;
      const double PU = (dl - U[t] * U[u]) * ir;
      double hm;
      if (DIAG) {
        const double PrA = (dl - eA[t] * eA[u]) * idA;
        hm = (PrA * gab - (eA[t] * U[u] + U[t] * eA[u]) - fab * PU) * ir2 + 2.0 * fab * U[t] * U[u] * ir3;
      } else {
        hm = (eA[t] * U[u] + U[t] * eN[u] + fab * PU) * ir2 - 2.0 * fab * U[t] * U[u] * ir3;
      }
      a9[3 * t + u] += cdd * rA[t] * rX[u] + c1 * hm;
    }
}

template <int NBPT>
__global__ void __launch_bounds__(256) becke_dd_kernel(const double* __restrict__ coords,
                                                       const double* __restrict__ wquad,
                                                       const double* __restrict__ cw, const int64_t* __restrict__ perm,
                                                       const int64_t* __restrict__ chunks,
                                                       const double* __restrict__ tab, int natm, int h, int GB,
                                                       double* __restrict__ part) {
  extern __shared__ double sm[];
  const int tid = threadIdx.x;
  const Tab T(tab, natm);
  const Rec L(natm);
  const int n3 = L.n3;
  const int64_t c = blockIdx.x;
  const int64_t i0 = chunks[3 * c], i1 = chunks[3 * c + 1];
  const int G = (int)chunks[3 * c + 2];
  const int nblk = natm * (natm + 1) / 2;
  const int ngrp = (natm + DIAG_GRP - 1) / DIAG_GRP;
  const int nunit = nblk + natm * ngrp;
  int uA[NBPT], uB[NBPT];
  double acc[NBPT][9];
#pragma unroll
  for (int k = 0; k < NBPT; ++k) {
    int p = tid + 256 * k, A = 0, B = 0;
    if (p < nblk) {
      while (p >= natm - A) {
        p -= natm - A;
        ++A;
      }
      B = A + p;
    } else if (p < nunit) {
      A = (p - nblk) / ngrp;
      B = ((p - nblk) % ngrp) * DIAG_GRP;  // first N of the group
    }
    uA[k] = A;
    uB[k] = B;
#pragma unroll
    for (int j = 0; j < 9; ++j) acc[k][j] = 0.0;
  }
  for (int64_t base = i0; base < i1; base += GB) {
    const int nb = (int)(i1 - base < GB ? i1 - base : GB);
    // distances and unit vectors towards the atoms
    for (int t = tid; t < nb * natm; t += 256) {
      const int gi = t / natm, A = t % natm;
      double* rec = sm + gi * L.stride;
      const int64_t g = perm ? perm[base + gi] : base + gi;
      const double d0 = T.R[3 * A] - coords[3 * g], d1 = T.R[3 * A + 1] - coords[3 * g + 1],
                   d2 = T.R[3 * A + 2] - coords[3 * g + 2];
      const double d = sqrt(d0 * d0 + d1 * d1 + d2 * d2), id = 1.0 / d;
      rec[A] = d;
      rec[L.o_id + A] = id;
      rec[L.o_e + 3 * A] = d0 * id;
      rec[L.o_e + 3 * A + 1] = d1 * id;
      rec[L.o_e + 3 * A + 2] = d2 * id;
    }
    __syncthreads();
    // cell functions P_M
    for (int t = tid; t < nb * natm; t += 256) {
      const int gi = t / natm, M = t % natm;
      double* rec = sm + gi * L.stride;
      double p = 1.0;
      for (int N = 0; N < natm; ++N) {
        if (N == M) continue;
        const double mu = (rec[M] - rec[N]) * T.irab[M * natm + N];
        p *= 0.5 * (1.0 - switch_eval<0>(mu, T.adj[M * natm + N], h).f);
      }
      rec[L.o_P + M] = p;
    }
    __syncthreads();
    for (int gi = tid; gi < nb; gi += 256) {
      double* rec = sm + gi * L.stride;
      const int64_t g = perm ? perm[base + gi] : base + gi;
      double Z = 0.0;
      for (int M = 0; M < natm; ++M) Z += rec[L.o_P + M];
      rec[L.o_sc] = Z;
      rec[L.o_sc + 1] = rec[L.o_P + G] / Z;
      rec[L.o_sc + 2] = wquad[g] * (cw ? cw[g] : 1.0);
    }
    __syncthreads();
    // l_M = d log P_M: block N != M is l_MN d mu_MN / dR_N, block M sums l_MN d mu_MN / dR_M
    for (int t = tid; t < nb * natm; t += 256) {
      const int gi = t / natm, M = t % natm;
      double* rec = sm + gi * L.stride;
      const double* e = rec + L.o_e;
      double* Lrow = rec + L.o_L + M * n3;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
      for (int N = 0; N < natm; ++N) {
        if (N == M) continue;
        const double ir = T.irab[M * natm + N], a = T.adj[M * natm + N];
        const double mu = (rec[M] - rec[N]) * ir;
        const Sw s = switch_eval<1>(mu, a, h);
        const double lM = -0.5 * s.df * (1.0 - 2.0 * mu * a) / fmax(0.5 * (1.0 - s.f), INVTOL);
        const double* U = T.U + 3 * (M * natm + N);
        Lrow[3 * N] = lM * (-e[3 * N] + mu * U[0]) * ir;
        Lrow[3 * N + 1] = lM * (-e[3 * N + 1] + mu * U[1]) * ir;
        Lrow[3 * N + 2] = lM * (-e[3 * N + 2] + mu * U[2]) * ir;
        s0 += lM * (e[3 * M] - mu * U[0]) * ir;
        s1 += lM * (e[3 * M + 1] - mu * U[1]) * ir;
        s2 += lM * (e[3 * M + 2] - mu * U[2]) * ir;
      }
      Lrow[3 * M] = s0;
      Lrow[3 * M + 1] = s1;
      Lrow[3 * M + 2] = s2;
      const double Z = rec[L.o_sc], q = rec[L.o_sc + 1], ct = rec[L.o_sc + 2];
      rec[L.o_gam + M] = ct * rec[L.o_P + M] * ((M == G ? 1.0 : 0.0) - q) / Z;
    }
    __syncthreads();
    for (int t = tid; t < nb * n3; t += 256) {
      const int gi = t / n3, a = t % n3;
      double* rec = sm + gi * L.stride;
      const double* Lm = rec + L.o_L;
      double z = 0.0;
      for (int M = 0; M < natm; ++M) z += rec[L.o_P + M] * Lm[M * n3 + a];
      const double Z = rec[L.o_sc], q = rec[L.o_sc + 1];
      rec[L.o_dZ + a] = z;
      rec[L.o_dq + a] = (rec[L.o_P + G] * Lm[G * n3 + a] - q * z) / Z;
    }
    __syncthreads();
    for (int gi = 0; gi < nb; ++gi) {
      const double* rec = sm + gi * L.stride;
      const double* Lm = rec + L.o_L;
      const double* gam = rec + L.o_gam;
      const double* dZ = rec + L.o_dZ;
      const double* dq = rec + L.o_dq;
      const double fac = -rec[L.o_sc + 2] / rec[L.o_sc];
#pragma unroll
      for (int k = 0; k < NBPT; ++k) {
        const int p = tid + 256 * k;
        if (p >= nunit) continue;
        const int A = uA[k], B = uB[k];
        double* a9 = acc[k];
        if (p >= nblk) {
          const int n1 = B + DIAG_GRP < natm ? B + DIAG_GRP : natm;
          for (int N = B; N < n1; ++N)
            if (N != A) pair_block<true>(T, natm, h, rec, L, A, N, a9);
          continue;
        }
        for (int M = 0; M < natm; ++M) {
          const double gm = gam[M];
          const double* la = Lm + M * n3 + 3 * A;
          const double* lb = Lm + M * n3 + 3 * B;
          const double x0 = gm * la[0], x1 = gm * la[1], x2 = gm * la[2];
          const double y0 = lb[0], y1 = lb[1], y2 = lb[2];
          a9[0] += x0 * y0; a9[1] += x0 * y1; a9[2] += x0 * y2;
          a9[3] += x1 * y0; a9[4] += x1 * y1; a9[5] += x1 * y2;
          a9[6] += x2 * y0; a9[7] += x2 * y1; a9[8] += x2 * y2;
        }
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
          for (int s = 0; s < 3; ++s)
            a9[3 * t + s] += fac * (dq[3 * A + t] * dZ[3 * B + s] + dZ[3 * A + t] * dq[3 * B + s]);
        if (A != B) pair_block<false>(T, natm, h, rec, L, A, B, a9);
      }
    }
    __syncthreads();
  }
  // diagonal-block pair units add into their block through LDS
  double* D = sm;
  for (int i = tid; i < 9 * natm; i += 256) D[i] = 0.0;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NBPT; ++k) {
    const int p = tid + 256 * k;
    if (p >= nblk && p < nunit)
      for (int j = 0; j < 9; ++j) atomicAdd(&D[9 * uA[k] + j], acc[k][j]);
  }
  __syncthreads();
  double* S = part + c * (int64_t)n3 * n3;
#pragma unroll
  for (int k = 0; k < NBPT; ++k) {
    const int p = tid + 256 * k;
    if (p >= nblk) continue;
    const int A = uA[k], B = uB[k];
    for (int t = 0; t < 3; ++t)
      for (int s = 0; s < 3; ++s) {
        double v = acc[k][3 * t + s];
        if (A == B) v += D[9 * A + 3 * t + s];
        S[(3 * A + t) * n3 + 3 * B + s] = v;
        if (A != B) S[(3 * B + s) * n3 + 3 * A + t] = v;
      }
  }
}

// Translation fix T S T^T of one chunk (becke_partition.rs:1300-1380 row / column / corner):
// rows and columns of G become minus the sums over the other atoms.  dst == nullptr: in place;
// else (one point per chunk) written to ddw[(a n3 + b) ng + g].
__global__ void __launch_bounds__(256) becke_fix_kernel(double* __restrict__ part, const int64_t* __restrict__ chunks,
                                                        const int64_t* __restrict__ perm, int64_t natm,
                                                        double* __restrict__ ddw, int64_t ng) {
  extern __shared__ double cs[];  // colsum [3][n3]: sum over A != G of S[(A, t), b]
  const int64_t n3 = 3 * natm;
  const int64_t c = blockIdx.x;
  const int64_t G = chunks[3 * c + 2];
  double* S = part + c * n3 * n3;
  for (int64_t idx = threadIdx.x; idx < 3 * n3; idx += blockDim.x) {
    const int64_t t = idx / n3, b = idx % n3;
    double v = 0.0;
    for (int64_t A = 0; A < natm; ++A)
      if (A != G) v += S[(3 * A + t) * n3 + b];
    cs[idx] = v;
  }
  __syncthreads();
  const int64_t g = ddw ? (perm ? perm[chunks[3 * c]] : chunks[3 * c]) : 0;
  for (int64_t idx = threadIdx.x; idx < n3 * n3; idx += blockDim.x) {
    const int64_t a = idx / n3, b = idx % n3;
    const int64_t A = a / 3, t = a % 3, B = b / 3, s = b % 3;
    double v;
    if (A != G && B != G) {
      v = S[idx];
    } else if (A == G && B != G) {
      v = -cs[t * n3 + b];
    } else if (A != G) {
      v = -cs[s * n3 + a];
    } else {
      v = 0.0;
      for (int64_t Bp = 0; Bp < natm; ++Bp)
        if (Bp != G) v += cs[t * n3 + 3 * Bp + s];
    }
    if (ddw) ddw[idx * ng + g] = v;
    else S[idx] = v;
  }
}

// ddc[e * nset + iset] = sum over chunks of the fixed S
__global__ void becke_reduce_kernel(const double* __restrict__ part, int64_t nchunk, int64_t nn, int64_t nset,
                                    int64_t iset, double* __restrict__ ddc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn) return;
  double v = 0.0;
  for (int64_t c = 0; c < nchunk; ++c) v += part[c * nn + e];
  ddc[e * nset + iset] = v;
}

unsigned nblocks(int64_t n, int tpb) { return (unsigned)((n + tpb - 1) / tpb); }

// C[m, n] = sum_k A[m * lda + k] B[n * ldb + k]  (both operands k-fastest), FP64
void gemm_kk(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
             int64_t ldc, hipStream_t s, DevBuf& ws) {
  GemmArgs g;
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.a_kmajor = true;
  g.B = B; g.ldb = ldb; g.b_kmajor = true;
  g.C = C; g.ldc = ldc;
  g.alpha = 1.0; g.beta = 0.0;
  gemm<double, double>(g, PREC_F64, s, ws);
}

}  // namespace

Becke::Becke(int64_t natm, const double* atm_coords, const double* adjustment, hipStream_t s)
    : natm_(natm), stream_(s) {
  HH_REQUIRE(natm >= 1 && atm_coords && adjustment, "becke: null tables or no atoms");
  std::vector<double> t(3 * natm + 6 * natm * natm, 0.0);
  double* R = t.data();
  double* adj = R + 3 * natm;
  double* irab = adj + natm * natm;
  double* U = irab + natm * natm;
  double* rab = U + 3 * natm * natm;
  for (int64_t i = 0; i < 3 * natm; ++i) R[i] = atm_coords[i];
  for (int64_t A = 0; A < natm; ++A)
    for (int64_t B = 0; B < natm; ++B) {
      adj[A * natm + B] = adjustment[A * natm + B];
      if (A == B) continue;
      const double d0 = R[3 * A] - R[3 * B], d1 = R[3 * A + 1] - R[3 * B + 1], d2 = R[3 * A + 2] - R[3 * B + 2];
      const double d = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      HH_REQUIRE(d > 0.0, "becke: two atoms share a position");
      irab[A * natm + B] = 1.0 / d;
      rab[A * natm + B] = d;
      U[3 * (A * natm + B)] = d0 / d;
      U[3 * (A * natm + B) + 1] = d1 / d;
      U[3 * (A * natm + B) + 2] = d2 / d;
    }
  double* dt = static_cast<double*>(tab_.reserve(sizeof(double) * t.size()));
  HH_HIP(hipMemcpy(dt, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
}

void Becke::partition(const double* coords, int64_t ng, bool by_atom, const int64_t* idx, const double* wquad,
                      int hardness, int deriv, const Out& o) {
  const int64_t natm = natm_, n3 = 3 * natm;
  HH_REQUIRE(deriv >= 0 && deriv <= 2, "becke: deriv must be 0, 1 or 2");
  HH_REQUIRE(hardness >= 1, "becke: hardness must be positive");
  HH_REQUIRE(ng >= 0 && idx && (ng == 0 || (coords && wquad)), "becke: null argument");
  HH_REQUIRE(o.nset_w >= 0 && o.nset_dw >= 0 && o.nset_ddw >= 0, "becke: negative contraction set count");
  const bool want_c = o.contract_w && o.nset_w > 0 && o.c;
  const bool want_dc = deriv >= 1 && o.contract_dw && o.nset_dw > 0 && o.dc;
  const bool want_ddc = deriv >= 2 && o.contract_ddw && o.nset_ddw > 0 && o.ddc;
  const bool want_dw = deriv >= 1 && (o.dw || want_dc);
  const bool want_ddw = deriv >= 2 && o.ddw;
  HH_REQUIRE(natm <= 256, "becke: at most 256 atoms");
  HH_REQUIRE(!(want_ddw || want_ddc) || natm <= 48,
             "becke: second derivatives keep per-point atom-pair tables in LDS and support at most 48 atoms");

  // sorted order: boundaries per atom and (ByGrid) the permutation
  std::vector<int64_t> bounds(natm + 1, 0), perm;
  int64_t n = ng;
  bool padding = false;
  if (by_atom) {
    HH_REQUIRE(idx[0] == 0, "ByAtom atm_indices must start with 0");
    HH_REQUIRE(idx[natm] == ng, "ByAtom atm_indices must end with ngrids");
    for (int64_t A = 0; A <= natm; ++A) {
      HH_REQUIRE(idx[A] >= 0 && idx[A] <= ng, "ByAtom atm_indices must not exceed ngrids");
      if (A > 0) HH_REQUIRE(idx[A - 1] <= idx[A], "ByAtom atm_indices must be non-decreasing");
      bounds[A] = idx[A];
    }
  } else {
    std::vector<int64_t> cnt(natm + 1, 0);
    for (int64_t g = 0; g < ng; ++g) {
      HH_REQUIRE(idx[g] >= 0, "ByGrid atm_indices must be non-negative");
      if (idx[g] < natm) ++cnt[idx[g] + 1];
      else padding = true;
    }
    for (int64_t A = 0; A < natm; ++A) cnt[A + 1] += cnt[A];
    bounds = cnt;
    n = cnt[natm];
    perm.resize(n);
    for (int64_t g = 0; g < ng; ++g)
      if (idx[g] < natm) perm[cnt[idx[g]]++] = g;
  }
  // chunks of one atom's points for the second derivatives: one point each for ddw, <= 256 for ddc
  auto make_chunks = [&](int64_t len) {
    std::vector<int64_t> ch;
    for (int64_t A = 0; A < natm; ++A)
      for (int64_t s = bounds[A]; s < bounds[A + 1]; s += len) {
        ch.push_back(s);
        ch.push_back(std::min(bounds[A + 1], s + len));
        ch.push_back(A);
      }
    return ch;
  };
  std::vector<int64_t> ch_ddw = want_ddw ? make_chunks(1) : std::vector<int64_t>();
  std::vector<int64_t> ch_ddc = want_ddc ? make_chunks(256) : std::vector<int64_t>();
  const int64_t n_meta = (natm + 1) + (int64_t)perm.size() + (int64_t)ch_ddw.size() + (int64_t)ch_ddc.size();
  int64_t* meta = static_cast<int64_t*>(meta_.reserve(sizeof(int64_t) * n_meta));
  int64_t* d_bounds = meta;
  int64_t* d_perm = perm.empty() ? nullptr : d_bounds + natm + 1;
  int64_t* d_chw = d_bounds + natm + 1 + perm.size();
  int64_t* d_chc = d_chw + ch_ddw.size();
  HH_HIP(hipMemcpyAsync(d_bounds, bounds.data(), sizeof(int64_t) * (natm + 1), hipMemcpyHostToDevice, stream_));
  if (!perm.empty())
    HH_HIP(hipMemcpyAsync(d_perm, perm.data(), sizeof(int64_t) * perm.size(), hipMemcpyHostToDevice, stream_));
  if (!ch_ddw.empty())
    HH_HIP(hipMemcpyAsync(d_chw, ch_ddw.data(), sizeof(int64_t) * ch_ddw.size(), hipMemcpyHostToDevice, stream_));
  if (!ch_ddc.empty())
    HH_HIP(hipMemcpyAsync(d_chc, ch_ddc.data(), sizeof(int64_t) * ch_ddc.size(), hipMemcpyHostToDevice, stream_));

  // scratch for w / dw when only their contractions are requested
  const bool w_scr = want_c && !o.w, dw_scr = want_dw && !o.dw;
  double* scr = static_cast<double*>(scr_.reserve(sizeof(double) * ((w_scr ? ng : 0) + (dw_scr ? n3 * ng : 0) + 1)));
  double* w = o.w ? o.w : (w_scr ? scr : nullptr);
  double* dw = o.dw && deriv >= 1 ? o.dw : (dw_scr ? scr + (w_scr ? ng : 0) : nullptr);
  if (padding) {
    if (w) fill<double>(w, ng, 0.0, stream_);
    if (dw) fill<double>(dw, n3 * ng, 0.0, stream_);
    if (want_ddw) fill<double>(o.ddw, n3 * n3 * ng, 0.0, stream_);
  }
  const double* tab = tab_.as<double>();
  if ((w || dw) && n > 0) {
    const int bs = natm <= 64 ? 64 : (natm <= 128 ? 32 : 16);
    const size_t lds = sizeof(double) * 2 * natm * bs;
    if (dw)
      hipLaunchKernelGGL(becke_w_dw_kernel<true>, dim3(nblocks(n, bs)), dim3(bs), lds, stream_, coords, wquad, d_perm,
                         d_bounds, n, tab, natm, hardness, w, dw, ng);
    else
      hipLaunchKernelGGL(becke_w_dw_kernel<false>, dim3(nblocks(n, bs)), dim3(bs), lds, stream_, coords, wquad,
                         d_perm, d_bounds, n, tab, natm, hardness, w, nullptr, ng);
    HH_HIP(hipGetLastError());
  }
  if (want_c) gemm_kk(o.nset_w, 1, ng, o.contract_w, ng, w, ng, o.c, o.nset_w, stream_, ws_);
  if (want_dc) gemm_kk(o.nset_dw, n3, ng, o.contract_dw, ng, dw, ng, o.dc, o.nset_dw, stream_, ws_);

  if (want_ddw || want_ddc) {
    const Rec L((int)natm);
    const int64_t gb = std::max<int64_t>(1, std::min<int64_t>(8, 8192 / L.stride));
    const size_t lds = sizeof(double) * gb * L.stride;
    const int64_t nunit = natm * (natm + 1) / 2 + natm * ((natm + DIAG_GRP - 1) / DIAG_GRP);
    const int na = (int)natm;
    auto run = [&](const int64_t* d_ch, int64_t nchunk, const double* cwv, double* part) {
      if (nchunk == 0) return;
      if (nunit <= 256)
        hipLaunchKernelGGL(becke_dd_kernel<1>, dim3((unsigned)nchunk), dim3(256), lds, stream_, coords, wquad, cwv,
                           d_perm, d_ch, tab, na, hardness, (int)gb, part);
      else if (nunit <= 512)
        hipLaunchKernelGGL(becke_dd_kernel<2>, dim3((unsigned)nchunk), dim3(256), lds, stream_, coords, wquad, cwv,
                           d_perm, d_ch, tab, na, hardness, (int)gb, part);
      else if (nunit <= 1024)
        hipLaunchKernelGGL(becke_dd_kernel<4>, dim3((unsigned)nchunk), dim3(256), lds, stream_, coords, wquad, cwv,
                           d_perm, d_ch, tab, na, hardness, (int)gb, part);
      else
        hipLaunchKernelGGL(becke_dd_kernel<8>, dim3((unsigned)nchunk), dim3(256), lds, stream_, coords, wquad, cwv,
                           d_perm, d_ch, tab, na, hardness, (int)gb, part);
      HH_HIP(hipGetLastError());
    };
    const size_t lds_fix = sizeof(double) * 3 * n3;
    if (want_ddw) {
      const int64_t nch = (int64_t)ch_ddw.size() / 3;
      double* part = static_cast<double*>(part_.reserve(sizeof(double) * std::max<int64_t>(1, nch) * n3 * n3));
      run(d_chw, nch, nullptr, part);
      if (nch > 0) {
        hipLaunchKernelGGL(becke_fix_kernel, dim3((unsigned)nch), dim3(256), lds_fix, stream_, part, d_chw, d_perm,
                           natm, o.ddw, ng);
        HH_HIP(hipGetLastError());
      }
    }
    if (want_ddc) {
      const int64_t nch = (int64_t)ch_ddc.size() / 3;
      double* part = static_cast<double*>(part_.reserve(sizeof(double) * std::max<int64_t>(1, nch) * n3 * n3));
      for (int64_t k = 0; k < o.nset_ddw; ++k) {
        if (nch == 0) {
          hipLaunchKernelGGL(becke_reduce_kernel, dim3(nblocks(n3 * n3, 256)), dim3(256), 0, stream_, part, 0,
                             n3 * n3, o.nset_ddw, k, o.ddc);
          HH_HIP(hipGetLastError());
          continue;
        }
        run(d_chc, nch, o.contract_ddw + k * ng, part);
        hipLaunchKernelGGL(becke_fix_kernel, dim3((unsigned)nch), dim3(256), lds_fix, stream_, part, d_chc, d_perm,
                           natm, nullptr, ng);
        HH_HIP(hipGetLastError());
        hipLaunchKernelGGL(becke_reduce_kernel, dim3(nblocks(n3 * n3, 256)), dim3(256), 0, stream_, part, nch,
                           n3 * n3, o.nset_ddw, k, o.ddc);
        HH_HIP(hipGetLastError());
      }
    }
  }
}

}  // namespace hh
