// RI-JK skeleton Hessian and first-derivative Fock terms (see rijk_skeleton.hpp for the algebra).
//
// Indices: u, v, k AOs (n = nao); P, Q auxiliary (m = naux); i, j occupied; x = t + 3 A a nuclear
// coordinate of listed atom A (N atoms).  V = j2c = L L^T, Li = L^-1, B = j3c L^-T (the handle's
// cderi, unpacked), so j3c = B L^T and V^-1 = Li^T Li.
// Per spin the occupied columns enter as C = mocc sqrt(occ); D = sum_s C_s C_s^T.
//   J:  d = V^-1 (P|D) = Li^T B^T vec(D)
//   K:  dK[ij, P] = V^-1 (ij|P) = (C^T B_R C) Li;   T~[u, j, P] = V^-1 (uj|P) = (B_R C) Li
// Derivative-integral passes per auxiliary range; everything else is global.
#include <algorithm>
#include <vector>

#include "elementwise.hpp"
#include "gemm_nt.hpp"
#include "linalg.hpp"
#include "rijk_skeleton.hpp"

namespace hh {
namespace {

unsigned nbk(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 1 << 20)); }

// C[m, n] = alpha sum_k A[m, k] B[k, n] + beta C; A[m,k] at akm ? A[m*lda+k] : A[m+k*lda];
// B[k,n] at bkm ? B[k+n*ldb] : B[k*ldb+n]; optional batch (nb0, strides)
void gm(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, bool akm, const double* B, int64_t ldb, bool bkm,
        double* C, int64_t ldc, double alpha, double beta, hipStream_t s, DevBuf& ws, int64_t nb0 = 1, int64_t sA = 0,
        int64_t sB = 0, int64_t sC = 0) {
  if (M <= 0 || N <= 0 || nb0 <= 0) return;
  if (K <= 0) {
    HH_REQUIRE(beta == 1.0, "skeleton: empty contraction with beta != 1");
    return;
  }
  GemmArgs g;
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.a_kmajor = akm;
  g.B = B; g.ldb = ldb; g.b_kmajor = bkm;
  g.C = C; g.ldc = ldc;
  g.alpha = alpha; g.beta = beta;
  g.nb0 = nb0; g.sA0 = sA; g.sB0 = sB; g.sC0 = sC;
  gemm<double, double>(g, PREC_F64, s, ws);
}

// The same product for ng groups in one launch: group g sums over the K slice
// [krange[2 g], krange[2 g + 1]) (device table) of the shared operands and writes C + g sC1.
void gmg(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, bool akm, const double* B, int64_t ldb,
         bool bkm, double* C, int64_t ldc, double alpha, double beta, hipStream_t s, DevBuf& ws, int64_t nb0,
         int64_t sA, int64_t sB, int64_t sC, int64_t ng, int64_t sC1, const int64_t* krange, int64_t kmax) {
  if (M <= 0 || N <= 0 || nb0 <= 0 || ng <= 0 || K <= 0) return;
  GemmArgs g;
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.a_kmajor = akm;
  g.B = B; g.ldb = ldb; g.b_kmajor = bkm;
  g.C = C; g.ldc = ldc;
  g.alpha = alpha; g.beta = beta;
  g.nb0 = nb0; g.sA0 = sA; g.sB0 = sB; g.sC0 = sC;
  g.nb1 = ng; g.sC1 = sC1;
  g.krange = krange;
  g.krange_max = kmax;
  gemm<double, double>(g, PREC_F64, s, ws);
}

#define GRID(n) dim3(nbk(n)), dim3(256), 0, st
#define ELEM (blockIdx.x * (int64_t)blockDim.x + threadIdx.x)

__global__ void eye_kernel(double* X, int64_t m) {
  const int64_t e = ELEM;
  if (e < m * m) X[e] = (e % m == e / m) ? 1.0 : 0.0;
}

// mocc[:, c] = src[:, idx[c]], Cs[:, c] = sqrt(occ) * mocc[:, c]
__global__ void gather_cols_kernel(const double* src, int64_t n, const int64_t* idx, const double* sc, int64_t no,
                                   double* mocc, double* C) {
  const int64_t e = ELEM;
  if (e >= n * no) return;
  const int64_t u = e % n, c = e / n;
  const double v = src[u + n * idx[c]];
  mocc[e] = v;
  C[e] = sc[c] * v;
}

// out[u, P, c] = sum_v I[u, v, P, c] M[u, v (, P)]
__global__ void rowsum_kernel(const double* __restrict__ I, int64_t n, int64_t nb, int64_t ncomp,
                              const double* __restrict__ M, int perP, double* __restrict__ out) {
  const int64_t e = ELEM;
  if (e >= n * nb * ncomp) return;
  const int64_t u = e % n, r = e / n, P = r % nb, c = r / nb;
  const double* Ip = I + u + n * n * (P + nb * c);
  const double* Mp = M + u + (perP ? n * n * P : 0);
  double acc = 0.0;
  for (int64_t v = 0; v < n; ++v) acc += Ip[n * v] * Mp[n * v];
  out[e] = acc;
}

// z[u, v, c] += sum_P I[u, v, P, c] M[u, v, P]
__global__ void psum_kernel(const double* __restrict__ I, int64_t n, int64_t nb, const double* __restrict__ M,
                            double* __restrict__ z) {
  const int64_t e = ELEM;
  if (e >= n * n * 9) return;
  const int64_t uv = e % (n * n), c = e / (n * n);
  double acc = 0.0;
  for (int64_t P = 0; P < nb; ++P) acc += I[uv + n * n * (P + nb * c)] * M[uv + n * n * P];
  z[e] += acc;
}

// XJ[p0 + P, t + 3A] -= 2 sum_{u in A} Y1[u, P, t]
__global__ void xj_ip1_kernel(const double* Y1, int64_t n, int64_t nb, const int64_t* aosl, int64_t N, double* XJ,
                              int64_t m, int64_t p0) {
  const int64_t e = ELEM;
  if (e >= nb * 3 * N) return;
  const int64_t P = e % nb, t = (e / nb) % 3, A = e / (3 * nb);
  double acc = 0.0;
  for (int64_t u = aosl[2 * A]; u < aosl[2 * A + 1]; ++u) acc += Y1[u + n * (P + nb * t)];
  XJ[(p0 + P) + m * (t + 3 * A)] -= 2.0 * acc;
}

// XJ[p0 + P, t + 3 A(P)] -= sum_u Y2[u, P, t]
__global__ void xj_ip2_kernel(const double* Y2, int64_t n, int64_t nb, const int64_t* auxatom, double* XJ, int64_t m,
                              int64_t p0) {
  const int64_t e = ELEM;
  if (e >= nb * 3) return;
  const int64_t P = e % nb, t = e / nb;
  const int64_t A = auxatom[p0 + P];
  if (A < 0) return;
  double acc = 0.0;
  for (int64_t u = 0; u < n; ++u) acc += Y2[u + n * (P + nb * t)];
  XJ[(p0 + P) + m * (t + 3 * A)] -= acc;
}

// Cm[u, i + no A] = C[u, i] for AOs u of atom A, else 0 (the per-atom restriction as one GEMM)
__global__ void mask_cols_kernel(const double* C, const int64_t* aoatom, int64_t n, int64_t no, int64_t N, double* Cm) {
  const int64_t e = ELEM;
  if (e >= n * no * N) return;
  const int64_t u = e % n, r = e / n, i = r % no, A = r / no;
  Cm[e] = aoatom[u] == A ? C[u + n * i] : 0.0;
}

// XK[p0 + P, i, j, t + 3A] -= HA[(i, A), j, P, t] + HA[(j, A), i, P, t]   (HA rows i + no A, ld no N)
__global__ void xk_ip1_kernel(const double* HA, int64_t no, int64_t nb, int64_t N, double* XK, int64_t m, int64_t p0) {
  const int64_t e = ELEM;
  if (e >= nb * no * no * 3 * N) return;
  const int64_t P = e % nb, r = e / nb, i = r % no, j = (r / no) % no, r2 = r / (no * no), t = r2 % 3, A = r2 / 3;
  const int64_t ld = no * N;
  const double v = HA[(i + no * A) + ld * (j + no * (P + nb * t))] + HA[(j + no * A) + ld * (i + no * (P + nb * t))];
  XK[(p0 + P) + m * (i + no * (j + no * (t + 3 * A)))] -= v;
}

// XK[p0 + P, i, j, t + 3 A(P)] -= H2[i, j, P, t]
__global__ void xk_ip2_kernel(const double* H2, int64_t no, int64_t nb, const int64_t* auxatom, double* XK, int64_t m,
                              int64_t p0) {
  const int64_t e = ELEM;
  if (e >= nb * no * no * 3) return;
  const int64_t P = e % nb, r = e / nb, i = r % no, j = (r / no) % no, t = r / (no * no);
  const int64_t A = auxatom[p0 + P];
  if (A < 0) return;
  XK[(p0 + P) + m * (i + no * (j + no * (t + 3 * A)))] -= H2[i + no * (j + no * (P + nb * t))];
}

// M[u, v, P] = sj d[p0 + P] D[u, v]
__global__ void mtot_j_kernel(const double* D, const double* d, int64_t n, int64_t nb, int64_t p0, double sj,
                              double* M) {
  const int64_t e = ELEM;
  if (e >= n * n * nb) return;
  M[e] = sj * d[p0 + e / (n * n)] * D[e % (n * n)];
}

// out[P, k, u, t] = I[k, u, P, t]
__global__ void transpose_ip_kernel(const double* __restrict__ I, int64_t n, int64_t nb, double* __restrict__ out) {
  const int64_t e = ELEM;
  if (e >= n * n * nb * 3) return;
  const int64_t k = e % n, r = e / n, u = r % n, r2 = r / n, P = r2 % nb, t = r2 / nb;
  out[P + nb * (k + n * (u + n * t))] = I[e];
}

// out[P, k, v] = Mp[k, P, v]
__global__ void transpose_m_kernel(const double* __restrict__ Mp, int64_t n, int64_t nb, double* __restrict__ out) {
  const int64_t e = ELEM;
  if (e >= n * n * nb) return;
  const int64_t k = e % n, r = e / n, P = r % nb, v = r / nb;
  out[P + nb * (k + n * v)] = Mp[e];
}

// r[u, c] += sum_P T[u, P, c]
__global__ void reduce_p_kernel(const double* T, int64_t n, int64_t nb, double* r) {
  const int64_t e = ELEM;
  if (e >= n * 9) return;
  const int64_t u = e % n, c = e / n;
  double acc = 0.0;
  for (int64_t P = 0; P < nb; ++P) acc += T[u + n * (P + nb * c)];
  r[e] += acc;
}

// y[u, p0 + P, c] = T[u, P, c]
__global__ void store_up_kernel(const double* T, int64_t n, int64_t nb, int64_t m, int64_t p0, double* y) {
  const int64_t e = ELEM;
  if (e >= n * nb * 9) return;
  const int64_t u = e % n, P = (e / n) % nb, c = e / (n * nb);
  y[u + n * ((p0 + P) + m * c)] = T[e];
}

// q[p0 + P, c] = sum_u T[u, P, c]
__global__ void reduce_u_kernel(const double* T, int64_t n, int64_t nb, int64_t m, int64_t p0, double* q) {
  const int64_t e = ELEM;
  if (e >= nb * 9) return;
  const int64_t P = e % nb, c = e / nb;
  double acc = 0.0;
  for (int64_t u = 0; u < n; ++u) acc += T[u + n * (P + nb * c)];
  q[(p0 + P) + m * c] = acc;
}

// j1ao[u, v, t + 3A] = acc - [A(u) = A] G1[u, v, t] - [A(v) = A] G1[v, u, t]
__global__ void j1_final_kernel(const double* acc, const double* G1, const int64_t* aoatom, int64_t n, int64_t N,
                                double* j1) {
  const int64_t e = ELEM;
  if (e >= n * n * 3 * N) return;
  const int64_t u = e % n, v = (e / n) % n, x = e / (n * n), t = x % 3, A = x / 3;
  double val = acc[e];
  if (aoatom[u] == A) val -= G1[u + n * (v + n * t)];
  if (aoatom[v] == A) val -= G1[v + n * (u + n * t)];
  j1[e] = val;
}

// k1ao[u, v, x] = acc[u, v, x] + acc[v, u, x] - [A(u) = A] r1[u, v, t] - [A(v) = A] r1[v, u, t]
__global__ void k1_final_kernel(const double* acc, const double* r1, const int64_t* aoatom, int64_t n, int64_t N,
                                double* k1) {
  const int64_t e = ELEM;
  if (e >= n * n * 3 * N) return;
  const int64_t u = e % n, v = (e / n) % n, x = e / (n * n), t = x % 3, A = x / 3;
  double val = acc[e] + acc[v + n * (u + n * x)];
  if (aoatom[u] == A) val -= r1[u + n * (v + n * t)];
  if (aoatom[v] == A) val -= r1[v + n * (u + n * t)];
  k1[e] = val;
}

// One workgroup per (t, s, A, B): the second-derivative-integral terms (contracted with M and F)
// added into H[t + 3A, s + 3B].  Component c = a + 3 b of a [.., 3, 3] derivative integral.
__global__ void __launch_bounds__(256) t1t2_kernel(const double* __restrict__ rpp, const double* __restrict__ z,
                                                   const double* __restrict__ y, const double* __restrict__ q,
                                                   const double* __restrict__ F, const double* __restrict__ v2pp,
                                                   const double* __restrict__ v2pq, const int64_t* __restrict__ aosl,
                                                   const int64_t* __restrict__ auxsl, int64_t n, int64_t m, int64_t N,
                                                   double* __restrict__ H) {
  const int64_t e = blockIdx.x;
  const int64_t t = e % 3, s = (e / 3) % 3, A = (e / 9) % N, B = e / (9 * N);
  const int64_t a0 = aosl[2 * A], na = aosl[2 * A + 1] - a0, b0 = aosl[2 * B], nbb = aosl[2 * B + 1] - b0;
  const int64_t pa0 = auxsl[2 * A], ma = auxsl[2 * A + 1] - pa0, pb0 = auxsl[2 * B], mb = auxsl[2 * B + 1] - pb0;
  const int64_t cts = t + 3 * s, cst = s + 3 * t;
  double v = 0.0;
  const int tid = threadIdx.x, nt = blockDim.x;
  if (A == B) {
    for (int64_t i = tid; i < na; i += nt) v += 2.0 * rpp[(a0 + i) + n * cts];
    for (int64_t i = tid; i < ma; i += nt) v += q[(pa0 + i) + m * cts];
    for (int64_t i = tid; i < ma * m; i += nt) {  // -sum_{P in A, Q} V_ipip1[P, Q, t, s] F[P, Q]
      const int64_t P = pa0 + i % ma, Q = i / ma;
      v -= v2pp[P + m * (Q + m * cts)] * F[P + m * Q];
    }
  }
  for (int64_t i = tid; i < na * nbb; i += nt) {  // ipvip1 over (u in A, v in B) and (u in B, v in A)
    const int64_t u = a0 + i % na, w = b0 + i / na;
    v += z[u + n * (w + n * cst)] + z[w + n * (u + n * cts)];
  }
  for (int64_t i = tid; i < na * mb; i += nt) {  // ip1ip2 (u in A, P in B)
    const int64_t u = a0 + i % na, P = pb0 + i / na;
    v += 2.0 * y[u + n * (P + m * cst)];
  }
  for (int64_t i = tid; i < nbb * ma; i += nt) {  // ip1ip2 (u in B, P in A)
    const int64_t u = b0 + i % nbb, P = pa0 + i / nbb;
    v += 2.0 * y[u + n * (P + m * cts)];
  }
  for (int64_t i = tid; i < ma * mb; i += nt) {  // -1/2 V_ip1ip2 (P in A, Q in B) and (P in B, Q in A)
    const int64_t P = pa0 + i % ma, Q = pb0 + i / ma;
    v -= 0.5 * (v2pq[P + m * (Q + m * cst)] * F[P + m * Q] + v2pq[Q + m * (P + m * cts)] * F[Q + m * P]);
  }
  __shared__ double red[256];
  red[tid] = v;
  __syncthreads();
  for (int h = nt / 2; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  if (tid == 0) H[(t + 3 * A) + 3 * N * (s + 3 * B)] += red[0];
}

// de[t, s, A, B] = H[t + 3A, s + 3B]
__global__ void de_out_kernel(const double* H, int64_t N, double* de) {
  const int64_t e = ELEM;
  if (e >= 9 * N * N) return;
  const int64_t t = e % 3, s = (e / 3) % 3, A = (e / 9) % N, B = e / (9 * N);
  de[e] = H[(t + 3 * A) + 3 * N * (s + 3 * B)];
}

// X[P, c, t, A(P)] += Y[P, c, t] for the auxiliary functions P of listed atoms (A(P) = auxatom[P])
__global__ void scatter_rows_kernel(const double* Y, int64_t m, int64_t ncol, const int64_t* auxatom, double* X) {
  const int64_t e = ELEM;
  if (e >= m * ncol * 3) return;
  const int64_t P = e % m, c = (e / m) % ncol, t = e / (m * ncol);
  const int64_t A = auxatom[P];
  if (A >= 0) X[P + m * (c + ncol * (t + 3 * A))] += Y[e];
}

// T[i + m k] = X[k + m i]
__global__ void transpose_sq_kernel(const double* X, int64_t m, double* T) {
  const int64_t e = ELEM;
  if (e < m * m) T[e] = X[(e / m) + m * (e % m)];
}

void launch_check() { HH_HIP(hipGetLastError()); }

struct SkeletonWs {
  DevBuf ws, bL, bLi, bmeta, bmo[2], bC[2], bD, bbD, bd, bBo[2], bbK[2], bdKij[2], bdKP[2], bTt[2], bg[2], bF, bXJ,
      bXK[2], bG1, bj1, bk1[2], br1[2], brpp, bz, by, bq, bH, bI, bIT, bY, bGs, bHA, bM, bMp, bMpp, bE1, bW, bout,
      bYu, bktab, bLiT, bCm[2];
};

}  // namespace

void rijk_skeleton(Rijk& r, const SkeletonInput& in, double* de, double* const* deriv1_bra) {
  const int64_t n = r.nao(), m = r.naux(), N = in.natm;
  const int ns = r.nspin();
  const hipStream_t st = r.stream();
  const double sj = r.scale_j(), sk = r.scale_k(), fk = ns == 1 ? 0.5 : 1.0;
  const bool doJ = sj != 0.0, doK = sk != 0.0;
  HH_REQUIRE(in.int3c && in.j2c_ip1 && in.j2c_ipip1 && in.j2c_ip1ip2 && de, "skeleton: null argument");
  HH_REQUIRE(N > 0 && in.aoslice && in.auxslice, "skeleton: no atoms");
  HH_REQUIRE(in.j2c || r.j2c_l() || r.j2c_l_inv(), "skeleton: j2c is required for a handle made from cderi");
  // ranges and slices
  HH_REQUIRE(in.nrange > 0 && in.aux_ranges, "skeleton: no auxiliary ranges");
  int64_t nbmax = 0;
  for (int64_t i = 0; i < in.nrange; ++i) {
    const int64_t p0 = in.aux_ranges[2 * i], p1 = in.aux_ranges[2 * i + 1];
    HH_REQUIRE(p0 == (i == 0 ? 0 : in.aux_ranges[2 * i - 1]) && p1 > p0, "skeleton: aux ranges must tile [0, naux)");
    nbmax = std::max(nbmax, p1 - p0);
  }
  HH_REQUIRE(in.aux_ranges[2 * in.nrange - 1] == m, "skeleton: aux ranges must end at naux");
  std::vector<int64_t> aoatom(n, -1), auxatom(m, -1);
  for (int64_t A = 0; A < N; ++A) {
    const int64_t a0 = in.aoslice[2 * A], a1 = in.aoslice[2 * A + 1];
    const int64_t q0 = in.auxslice[2 * A], q1 = in.auxslice[2 * A + 1];
    HH_REQUIRE(0 <= a0 && a0 <= a1 && a1 <= n && 0 <= q0 && q0 <= q1 && q1 <= m, "skeleton: slice out of range");
    for (int64_t u = a0; u < a1; ++u) {
      HH_REQUIRE(aoatom[u] < 0, "skeleton: AO slices overlap");
      aoatom[u] = A;
    }
    for (int64_t P = q0; P < q1; ++P) {
      HH_REQUIRE(auxatom[P] < 0, "skeleton: auxiliary slices overlap");
      auxatom[P] = A;
    }
  }
  // occupied orbitals per spin
  std::vector<std::vector<int64_t>> occidx(ns);
  std::vector<std::vector<double>> occsq(ns);
  int64_t no[2] = {0, 0};
  for (int s = 0; s < ns; ++s) {
    HH_REQUIRE(in.mo_coeff[s] && in.mo_occ[s] && in.nmo[s] > 0, "skeleton: null orbitals");
    for (int64_t i = 0; i < in.nmo[s]; ++i)
      if (in.mo_occ[s][i] > 0.0) {
        occidx[s].push_back(i);
        occsq[s].push_back(std::sqrt(in.mo_occ[s][i]));
      }
    no[s] = (int64_t)occidx[s].size();
    HH_REQUIRE(no[s] > 0, "skeleton: no occupied orbitals");
    if (deriv1_bra) HH_REQUIRE(deriv1_bra[s], "skeleton: null deriv1_bra");
  }

  // buffers: the handle's skeleton workspace (allocated on the first call, grown as needed)
  std::shared_ptr<void>& slot = r.skeleton_workspace();
  if (!slot) slot = std::make_shared<SkeletonWs>();
  SkeletonWs& W_ = *static_cast<SkeletonWs*>(slot.get());
  DevBuf &ws = W_.ws, &bL = W_.bL, &bLi = W_.bLi, &bmeta = W_.bmeta, &bD = W_.bD, &bbD = W_.bbD, &bd = W_.bd,
         &bF = W_.bF, &bXJ = W_.bXJ, &bG1 = W_.bG1, &bj1 = W_.bj1, &brpp = W_.brpp, &bz = W_.bz, &by = W_.by,
         &bq = W_.bq, &bH = W_.bH, &bI = W_.bI, &bIT = W_.bIT, &bY = W_.bY, &bGs = W_.bGs, &bHA = W_.bHA,
         &bM = W_.bM, &bMp = W_.bMp, &bMpp = W_.bMpp, &bE1 = W_.bE1, &bW = W_.bW, &bout = W_.bout, &bYu = W_.bYu,
         &bktab = W_.bktab, &bLiT = W_.bLiT;
  DevBuf(&bmo)[2] = W_.bmo;
  DevBuf(&bC)[2] = W_.bC;
  DevBuf(&bBo)[2] = W_.bBo;
  DevBuf(&bbK)[2] = W_.bbK;
  DevBuf(&bdKij)[2] = W_.bdKij;
  DevBuf(&bdKP)[2] = W_.bdKP;
  DevBuf(&bTt)[2] = W_.bTt;
  DevBuf(&bg)[2] = W_.bg;
  DevBuf(&bXK)[2] = W_.bXK;
  DevBuf(&bk1)[2] = W_.bk1;
  DevBuf(&br1)[2] = W_.br1;
  DevBuf(&bCm)[2] = W_.bCm;
  auto D = [](DevBuf& b, int64_t count) { return static_cast<double*>(b.reserve(sizeof(double) * std::max<int64_t>(1, count))); };
  auto zero = [&](double* p, int64_t count) { fill<double>(p, count, 0.0, st); };

  // metadata on the device: slices, atom maps, orbital index lists
  const int64_t nmeta = 4 * N + n + m + 2 * (n + 1);
  int64_t* meta = static_cast<int64_t*>(bmeta.reserve(sizeof(int64_t) * nmeta + sizeof(double) * 2 * (n + 1)));
  int64_t* d_aosl = meta;
  int64_t* d_auxsl = meta + 2 * N;
  int64_t* d_aoatom = meta + 4 * N;
  int64_t* d_auxatom = d_aoatom + n;
  int64_t* d_idx[2] = {d_auxatom + m, d_auxatom + m + (n + 1)};
  double* d_sc[2] = {reinterpret_cast<double*>(meta + nmeta), reinterpret_cast<double*>(meta + nmeta) + (n + 1)};
  HH_HIP(hipMemcpyAsync(d_aosl, in.aoslice, sizeof(int64_t) * 2 * N, hipMemcpyHostToDevice, st));
  HH_HIP(hipMemcpyAsync(d_auxsl, in.auxslice, sizeof(int64_t) * 2 * N, hipMemcpyHostToDevice, st));
  HH_HIP(hipMemcpyAsync(d_aoatom, aoatom.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, st));
  HH_HIP(hipMemcpyAsync(d_auxatom, auxatom.data(), sizeof(int64_t) * m, hipMemcpyHostToDevice, st));
  for (int s = 0; s < ns; ++s) {
    HH_REQUIRE(no[s] <= n + 1, "skeleton: more occupied orbitals than AOs");
    HH_HIP(hipMemcpyAsync(d_idx[s], occidx[s].data(), sizeof(int64_t) * no[s], hipMemcpyHostToDevice, st));
    HH_HIP(hipMemcpyAsync(d_sc[s], occsq[s].data(), sizeof(double) * no[s], hipMemcpyHostToDevice, st));
  }

  // metric factor inverse Li with V^-1 = Li^T Li: L^-1 of the Cholesky factor, or the handle's
  // symmetric J^-1/2 (eigen decomposition, the pseudo-inverse over the kept eigenvalues)
  const double* L = r.j2c_l();
  const double* Li = nullptr;
  if (in.j2c) {
    double* Lw = D(bL, m * m);
    copy_bytes(Lw, in.j2c, (int64_t)sizeof(double) * m * m, st);
    HH_REQUIRE(cholesky(Lw, m, st, ws), "skeleton: j2c is not positive definite");
    L = Lw;
  } else if (!L) {
    Li = r.j2c_l_inv();
  }
  HH_REQUIRE(L || Li, "skeleton: no metric given and the handle keeps no factor of it");
  if (!Li) {
    double* Lw = D(bLi, m * m);
    hipLaunchKernelGGL(eye_kernel, GRID(m * m), Lw, m);
    launch_check();
    trsm_lower(L, m, Lw, m, m, false, st, ws);
    Li = Lw;
  }

  const double* Bf = r.bfull();
  double* mo[2];
  double* C[2];
  for (int s = 0; s < ns; ++s) {
    mo[s] = D(bmo[s], n * no[s]);
    C[s] = D(bC[s], n * no[s]);
    hipLaunchKernelGGL(gather_cols_kernel, GRID(n * no[s]), in.mo_coeff[s], n, d_idx[s], d_sc[s], no[s], mo[s], C[s]);
    launch_check();
  }
  // D = sum_s C C^T;  d = Li^T B^T vec(D)
  double* Dm = D(bD, n * n);
  zero(Dm, n * n);
  for (int s = 0; s < ns; ++s) gm(n, n, no[s], C[s], n, false, C[s], n, false, Dm, n, 1.0, 1.0, st, ws);
  double* bDv = D(bbD, m);
  double* d = D(bd, m);
  gm(m, 1, n * n, Bf, n * n, true, Dm, n * n, true, bDv, m, 1.0, 0.0, st, ws);
  gm(m, 1, m, Li, m, true, bDv, m, true, d, m, 1.0, 0.0, st, ws);

  // K intermediates per spin
  double *dKij[2] = {nullptr, nullptr}, *dKP[2] = {nullptr, nullptr}, *Tt[2] = {nullptr, nullptr},
         *g[2] = {nullptr, nullptr};
  if (doK)
    for (int s = 0; s < ns; ++s) {
      const int64_t o = no[s], o2 = o * o;
      double* Bo = D(bBo[s], n * o * m);
      gm(n, o, n, Bf, n, false, C[s], n, true, Bo, n, 1.0, 0.0, st, ws, m, n * n, 0, n * o);
      double* bK = D(bbK[s], o2 * m);
      gm(o, o, n, C[s], n, true, Bo, n, true, bK, o, 1.0, 0.0, st, ws, m, 0, n * o, o2);
      dKij[s] = D(bdKij[s], o2 * m);
      gm(o2, m, m, bK, o2, false, Li, m, true, dKij[s], o2, 1.0, 0.0, st, ws);
      dKP[s] = D(bdKP[s], m * o2);
      gm(m, o2, m, Li, m, true, bK, o2, false, dKP[s], m, 1.0, 0.0, st, ws);
      Tt[s] = D(bTt[s], n * o * m);
      gm(n * o, m, m, Bo, n * o, false, Li, m, true, Tt[s], n * o, 1.0, 0.0, st, ws);
      if (deriv1_bra) {
        g[s] = D(bg[s], n * o * m * 3);  // g[v, j, P, t] = sum_Q j2c_ip1[P, Q, t] T~[v, j, Q]
        for (int t = 0; t < 3; ++t)
          gm(n * o, m, m, Tt[s], n * o, false, in.j2c_ip1 + m * m * t, m, false, g[s] + n * o * m * t, n * o, 1.0, 0.0,
             st, ws);
      }
    }
  // F = sj d d^T - fk sk sum_s dK dK^T
  double* F = D(bF, m * m);
  zero(F, m * m);
  if (doJ) gm(m, m, 1, d, m, false, d, 1, false, F, m, sj, 1.0, st, ws);
  if (doK)
    for (int s = 0; s < ns; ++s)
      gm(m, m, no[s] * no[s], dKP[s], m, false, dKP[s], m, false, F, m, -fk * sk, 1.0, st, ws);

  // Z = X - U: start with -U = +(first-derivative metric terms applied to d)
  double* XJ = D(bXJ, m * 3 * N);
  zero(XJ, m * 3 * N);
  double* XK[2] = {nullptr, nullptr};
  // X[:, c, s, B] (one block per listed atom B and component s):
  //   rows P in B: + sum_Q ip1[P, Q, s] d[Q, c]   (one product for all rows, scattered to P's atom)
  //   all rows:    + sum_{Q in B} ip1[Q, P, s] d[Q, c]   (the atoms as K slices of one launch)
  const int64_t ncol_max = std::max<int64_t>(1, std::max(no[0] * no[0], ns > 1 ? no[1] * no[1] : 0));
  int64_t kmax_aux = 0;
  for (int64_t A = 0; A < N; ++A) kmax_aux = std::max(kmax_aux, in.auxslice[2 * A + 1] - in.auxslice[2 * A]);
  double* Yu = D(bYu, m * ncol_max * 3);
  auto add_minus_u = [&](double* X, const double* dcols, int64_t ncol) {
    gm(m, ncol, m, in.j2c_ip1, m, false, dcols, m, true, Yu, m, 1.0, 0.0, st, ws, 3, m * m, 0, m * ncol);
    hipLaunchKernelGGL(scatter_rows_kernel, GRID(m * ncol * 3), Yu, m, ncol, d_auxatom, X);
    launch_check();
    gmg(m, ncol, m, in.j2c_ip1, m, true, dcols, m, true, X, m, 1.0, 1.0, st, ws, 3, m * m, 0, m * ncol, N,
        3 * m * ncol, d_auxsl, kmax_aux);
  };
  if (doJ) add_minus_u(XJ, d, 1);
  if (doK)
    for (int s = 0; s < ns; ++s) {
      XK[s] = D(bXK[s], m * no[s] * no[s] * 3 * N);
      zero(XK[s], m * no[s] * no[s] * 3 * N);
      add_minus_u(XK[s], dKP[s], no[s] * no[s]);
    }

  // accumulators
  const int64_t nx = 3 * N;
  double* G1 = D(bG1, n * n * 3);
  double* j1acc = D(bj1, n * n * nx);
  zero(G1, n * n * 3);
  zero(j1acc, n * n * nx);
  double *k1acc[2] = {nullptr, nullptr}, *r1[2] = {nullptr, nullptr};
  if (doK && deriv1_bra)
    for (int s = 0; s < ns; ++s) {
      k1acc[s] = D(bk1[s], n * n * nx);
      r1[s] = D(br1[s], n * n * 3);
      zero(k1acc[s], n * n * nx);
      zero(r1[s], n * n * 3);
    }
  double* rpp = D(brpp, n * 9);
  double* z = D(bz, n * n * 9);
  double* y = D(by, n * m * 9);
  double* q = D(bq, m * 9);
  double* H = D(bH, nx * nx);
  zero(rpp, n * 9);
  zero(z, n * n * 9);
  zero(y, n * m * 9);
  zero(q, m * 9);
  zero(H, nx * nx);

  // batch scratch
  int64_t omax = std::max(no[0], ns > 1 ? no[1] : 0);
  double* I = D(bI, n * n * nbmax * 9);
  double* IT = D(bIT, n * n * nbmax * 3);
  double* Yb = D(bY, n * nbmax * 9);
  double* Gs = D(bGs, n * omax * nbmax * 3);
  double* HA = D(bHA, omax * omax * nbmax * 3 * N);
  double* Cm[2] = {nullptr, nullptr};
  if (doK)
    for (int s = 0; s < ns; ++s) {
      Cm[s] = D(bCm[s], n * no[s] * N);
      hipLaunchKernelGGL(mask_cols_kernel, GRID(n * no[s] * N), C[s], d_aoatom, n, no[s], N, Cm[s]);
      launch_check();
    }
  double* M = D(bM, n * n * nbmax);
  double* Mp = D(bMp, n * n * nbmax);
  double* Mpp = D(bMpp, n * n * nbmax);
  double* E1 = D(bE1, n * omax * nbmax);
  auto request = [&](int32_t kind, int64_t p0, int64_t p1) {
    const int32_t rc = in.int3c(in.ctx, kind, p0, p1, I, (void*)st);
    if (rc != 0) throw Error(HH_ERR_INVALID, "int3c generator returned " + std::to_string(rc));
  };

  // per auxiliary range, the K slices of the per-atom products: ip1 k1acc (K = (k, P): the atom's
  // AOs), ip2 j1acc (K = P: the atom's auxiliary functions in the range), ip2 k1acc (K = (k, P))
  std::vector<int64_t> ktab((size_t)in.nrange * 6 * N, 0), kmax_tab((size_t)in.nrange * 3, 0);
  // listed atoms with auxiliary functions in each range: [alo, ahi) (contiguous: the slices are
  // ordered by atom), so the ip2 products launch over those atoms only
  std::vector<int64_t> alo((size_t)in.nrange, 0), ahi((size_t)in.nrange, 0);
  for (int64_t ir = 0; ir < in.nrange; ++ir) {
    const int64_t p0 = in.aux_ranges[2 * ir], p1 = in.aux_ranges[2 * ir + 1], nb = p1 - p0;
    int64_t* t = ktab.data() + ir * 6 * N;
    for (int64_t A = 0; A < N; ++A) {
      t[2 * A] = nb * in.aoslice[2 * A];
      t[2 * A + 1] = nb * in.aoslice[2 * A + 1];
      const int64_t q0 = std::max(p0, in.auxslice[2 * A]), q1 = std::min(p1, in.auxslice[2 * A + 1]);
      if (q1 > q0) {
        if (ahi[ir] == 0) alo[ir] = A;
        ahi[ir] = A + 1;
        t[2 * N + 2 * A] = q0 - p0;
        t[2 * N + 2 * A + 1] = q1 - p0;
        t[4 * N + 2 * A] = n * (q0 - p0);
        t[4 * N + 2 * A + 1] = n * (q1 - p0);
      }
      for (int k = 0; k < 3; ++k)
        kmax_tab[3 * ir + k] = std::max(kmax_tab[3 * ir + k], t[2 * N * k + 2 * A + 1] - t[2 * N * k + 2 * A]);
    }
  }
  // the final K products per spin: K = (j, P) over the atom's auxiliary functions
  for (int s = 0; s < ns; ++s)
    for (int64_t A = 0; A < N; ++A) {
      ktab.push_back(no[s] * in.auxslice[2 * A]);
      ktab.push_back(no[s] * in.auxslice[2 * A + 1]);
    }
  int64_t* d_ktab = static_cast<int64_t*>(bktab.reserve(sizeof(int64_t) * ktab.size()));
  const int64_t* d_ktab_fin = d_ktab + in.nrange * 6 * N;
  HH_HIP(hipMemcpyAsync(d_ktab, ktab.data(), sizeof(int64_t) * ktab.size(), hipMemcpyHostToDevice, st));

  for (int64_t ir = 0; ir < in.nrange; ++ir) {
    const int64_t p0 = in.aux_ranges[2 * ir], p1 = in.aux_ranges[2 * ir + 1], nb = p1 - p0;
    const int64_t* tab_ip1 = d_ktab + ir * 6 * N;
    const int64_t* tab_ip2j = tab_ip1 + 2 * N;
    const int64_t* tab_ip2k = tab_ip1 + 4 * N;
    // M_P = sj d_P D - fk sk sum_s C dK_P C^T for this range
    hipLaunchKernelGGL(mtot_j_kernel, GRID(n * n * nb), Dm, d, n, nb, p0, doJ ? sj : 0.0, M);
    launch_check();
    if (doK)
      for (int s = 0; s < ns; ++s) {
        const int64_t o = no[s];
        gm(n, o, o, C[s], n, false, dKij[s] + o * o * p0, o, true, E1, n, 1.0, 0.0, st, ws, nb, 0, o * o, n * o);
        gm(n, n, o, E1, n, false, C[s], n, false, M, n, -fk * sk, 1.0, st, ws, nb, n * o, 0, n * n);
      }

    // ---- first derivatives on the bra AO: int3c2e_ip1 ----
    request(HH_INT3C_IP1, p0, p1);
    if (doJ) {
      hipLaunchKernelGGL(rowsum_kernel, GRID(n * nb * 3), I, n, nb, 3, Dm, 0, Yb);
      launch_check();
      hipLaunchKernelGGL(xj_ip1_kernel, GRID(nb * 3 * N), Yb, n, nb, d_aosl, N, XJ, m, p0);
      launch_check();
      if (deriv1_bra)  // G1[u, v, t] += sum_P ip1[u, v, P, t] d_P
        gm(n * n, 1, nb, I, n * n, false, d + p0, nb, true, G1, n * n, 1.0, 1.0, st, ws, 3, n * n * nb, 0, n * n);
    }
    if (doK) {
      if (deriv1_bra) {
        hipLaunchKernelGGL(transpose_ip_kernel, GRID(n * n * nb * 3), I, n, nb, IT);
        launch_check();
      }
      for (int s = 0; s < ns; ++s) {
        const int64_t o = no[s];
        // G[u, j, P, t] = sum_v ip1[u, v, P, t] C[v, j]
        gm(n, o, n, I, n, false, C[s], n, true, Gs, n, 1.0, 0.0, st, ws, nb * 3, n * n, 0, n * o);
        // HA[(i, A), (j, P, t)] = sum_{u in A} C[u, i] G[u, j, P, t], all atoms in one product
        gm(o * N, o * nb * 3, n, Cm[s], n, true, Gs, n, true, HA, o * N, 1.0, 0.0, st, ws);
        hipLaunchKernelGGL(xk_ip1_kernel, GRID(nb * o * o * 3 * N), HA, o, nb, N, XK[s], m, p0);
        launch_check();
        if (deriv1_bra) {
          // M'[k, P, v] = sum_j C[k, j] T~[v, j, p0 + P]  and its [P, k, v] copy
          gm(n, n, o, C[s], n, false, Tt[s] + n * o * p0, n, false, Mp, n * nb, 1.0, 0.0, st, ws, nb, 0, n * o, n);
          hipLaunchKernelGGL(transpose_m_kernel, GRID(n * n * nb), Mp, n, nb, Mpp);
          launch_check();
          // r1[u, v, t] += sum_{k, P} ip1[u, k, P, t] M'[k, P, v]
          gm(n, n, n * nb, I, n, false, Mp, n * nb, true, r1[s], n, 1.0, 1.0, st, ws, 3, n * n * nb, 0, n * n);
          // k1acc[u, v, t, A] -= sum_{k in A, P} ip1[k, u, P, t] M'[k, P, v]: the atoms' AO
          // blocks (K = (k, P), P fastest) as K slices of one launch
          gmg(n, n, nb * n, IT, nb * n, true, Mpp, nb * n, true, k1acc[s], n, -1.0, 1.0, st, ws, 3, n * n * nb, 0,
              n * n, N, n * n * 3, tab_ip1, kmax_tab[3 * ir]);
        }
      }
    }

    // ---- first derivatives on the auxiliary function: int3c2e_ip2 ----
    request(HH_INT3C_IP2, p0, p1);
    if (doJ) {
      hipLaunchKernelGGL(rowsum_kernel, GRID(n * nb * 3), I, n, nb, 3, Dm, 0, Yb);
      launch_check();
      hipLaunchKernelGGL(xj_ip2_kernel, GRID(nb * 3), Yb, n, nb, d_auxatom, XJ, m, p0);
      launch_check();
    }
    if (doJ && deriv1_bra)  // j1acc[u, v, t, A] -= sum_{P in A} ip2[u, v, P, t] d_P (atoms: K slices)
      gmg(n * n, 1, nb, I, n * n, false, d + p0, nb, true, j1acc + n * n * 3 * alo[ir], n * n, -1.0, 1.0, st, ws, 3,
          n * n * nb, 0, n * n, ahi[ir] - alo[ir], n * n * 3, tab_ip2j + 2 * alo[ir], kmax_tab[3 * ir + 1]);
    if (doK)
      for (int s = 0; s < ns; ++s) {
        const int64_t o = no[s];
        gm(n, o, n, I, n, false, C[s], n, true, Gs, n, 1.0, 0.0, st, ws, nb * 3, n * n, 0, n * o);
        gm(o, o * nb * 3, n, C[s], n, true, Gs, n, true, HA, o, 1.0, 0.0, st, ws);
        hipLaunchKernelGGL(xk_ip2_kernel, GRID(nb * o * o * 3), HA, o, nb, d_auxatom, XK[s], m, p0);
        launch_check();
        if (deriv1_bra) {
          gm(n, n, o, C[s], n, false, Tt[s] + n * o * p0, n, false, Mp, n * nb, 1.0, 0.0, st, ws, nb, 0, n * o, n);
          // k1acc[u, v, t, A] -= sum_{P in A, k} ip2[u, k, P, t] M'[k, P, v] (atoms: K slices)
          gmg(n, n, n * nb, I, n, false, Mp, n * nb, true, k1acc[s] + n * n * 3 * alo[ir], n, -1.0, 1.0, st, ws, 3,
              n * n * nb, 0, n * n, ahi[ir] - alo[ir], n * n * 3, tab_ip2k + 2 * alo[ir], kmax_tab[3 * ir + 2]);
        }
      }

    // ---- second derivatives, contracted with M ----
    request(HH_INT3C_IPIP1, p0, p1);
    hipLaunchKernelGGL(rowsum_kernel, GRID(n * nb * 9), I, n, nb, 9, M, 1, Yb);
    launch_check();
    hipLaunchKernelGGL(reduce_p_kernel, GRID(n * 9), Yb, n, nb, rpp);
    launch_check();
    request(HH_INT3C_IPVIP1, p0, p1);
    hipLaunchKernelGGL(psum_kernel, GRID(n * n * 9), I, n, nb, M, z);
    launch_check();
    request(HH_INT3C_IP1IP2, p0, p1);
    hipLaunchKernelGGL(rowsum_kernel, GRID(n * nb * 9), I, n, nb, 9, M, 1, Yb);
    launch_check();
    hipLaunchKernelGGL(store_up_kernel, GRID(n * nb * 9), Yb, n, nb, m, p0, y);
    launch_check();
    request(HH_INT3C_IPIP2, p0, p1);
    hipLaunchKernelGGL(rowsum_kernel, GRID(n * nb * 9), I, n, nb, 9, M, 1, Yb);
    launch_check();
    hipLaunchKernelGGL(reduce_u_kernel, GRID(nb * 9), Yb, n, nb, m, p0, q);
    launch_check();
  }

  // ---- products of the first-derivative vectors: Z^T V^-1 Z = (Li Z)^T (Li Z) ----
  if (doJ) {
    double* W = D(bW, m * nx);
    gm(m, nx, m, Li, m, false, XJ, m, true, W, m, 1.0, 0.0, st, ws);
    gm(nx, nx, m, W, m, true, W, m, true, H, nx, sj, 1.0, st, ws);
    if (deriv1_bra) gm(n * n, nx, m, Bf, n * n, false, W, m, true, j1acc, n * n, 1.0, 1.0, st, ws);
  }
  if (doK) {
    // with naux a multiple of 32 both products run on the NT FP64 kernel (K-contiguous operands:
    // Li^T rows, the columns of XK and of W)
    const bool nt = m % 32 == 0;
    double* LiT = nullptr;
    if (nt) {
      LiT = D(bLiT, m * m);
      hipLaunchKernelGGL(transpose_sq_kernel, GRID(m * m), Li, m, LiT);
      launch_check();
    }
    for (int s = 0; s < ns; ++s) {
      const int64_t o2 = no[s] * no[s];
      double* W = D(bW, m * o2 * nx);
      if (nt) {
        NtArgs a;
        a.M = m; a.N = o2 * nx; a.K = m;
        a.A = LiT; a.lda = m; a.B = XK[s]; a.ldb = m;
        a.C = W; a.ldc = m;
        gemm_nt(a, NT_F64, st, ws);
        NtArgs h;
        h.M = nx; h.N = nx; h.K = m * o2;
        h.A = W; h.lda = m * o2; h.B = W; h.ldb = m * o2;
        h.C = H; h.ldc = nx;
        h.alpha = -fk * sk; h.beta = 1.0;
        gemm_nt(h, NT_F64, st, ws);
      } else {
        gm(m, o2 * nx, m, Li, m, false, XK[s], m, true, W, m, 1.0, 0.0, st, ws);
        gm(nx, nx, m * o2, W, m * o2, true, W, m * o2, true, H, nx, -fk * sk, 1.0, st, ws);
      }
    }
  }
  // ---- second-derivative-integral and metric terms ----
  hipLaunchKernelGGL(t1t2_kernel, dim3((unsigned)(9 * N * N)), dim3(256), 0, st, rpp, z, y, q, F, in.j2c_ipip1,
                     in.j2c_ip1ip2, d_aosl, d_auxsl, n, m, N, H);
  launch_check();
  hipLaunchKernelGGL(de_out_kernel, GRID(9 * N * N), H, N, de);
  launch_check();

  // ---- first-derivative Fock terms in bra form ----
  if (deriv1_bra) {
    double* F1 = D(bout, n * n * nx);
    for (int s = 0; s < ns; ++s) {
      double* out = deriv1_bra[s];
      const int64_t o = no[s];
      zero(out, n * o * nx);
      if (doJ) {
        hipLaunchKernelGGL(j1_final_kernel, GRID(n * n * nx), j1acc, G1, d_aoatom, n, N, F1);
        launch_check();
        gm(n, o, n, F1, n, false, mo[s], n, true, out, n, sj, 1.0, st, ws, nx, n * n, 0, n * o);
      }
      if (doK) {
        // S[u, v, t, A] = sum_{j, P in A} T~[u, j, P] g[v, j, P, t], added to the accumulator (the
        // atoms as K slices (j, P), j fastest, of one launch)
        gmg(n, n, o * m, Tt[s], n, false, g[s], n, false, k1acc[s], n, 1.0, 1.0, st, ws, 3, 0, n * o * m, n * n, N,
            n * n * 3, d_ktab_fin + 2 * N * s, o * kmax_aux);
        hipLaunchKernelGGL(k1_final_kernel, GRID(n * n * nx), k1acc[s], r1[s], d_aoatom, n, N, F1);
        launch_check();
        gm(n, o, n, F1, n, false, mo[s], n, true, out, n, -fk * sk, 1.0, st, ws, nx, n * n, 0, n * o);
      }
    }
  }
  HH_HIP(hipStreamSynchronize(st));
}

}  // namespace hh
