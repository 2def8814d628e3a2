// Density-fitted restricted Hartree-Fock on the device (see scf.hpp).
//
// Per cycle: F' = X^T F X diagonalised by the Jacobi eigensolver (X = U s^-1/2 from the overlap),
// warm-started in the previous cycle's eigenvectors V (V^T F' V is nearly diagonal, so the
// quadratically convergent sweeps stop after a few), C = X V, D = 2 C_o C_o^T; Coulomb v_P = sum_{mu nu} B[mu, nu, P] D[mu, nu], J = B v (two FP64
// matrix-vector products over the unpacked cderi B [nao, nao, naux]); exchange T_P = B_P C_o
// (batched over P) and K = 2 T T^T with T as [nao, nocc naux] (one FP64 matrix-core product);
// F = h + J - K / 2; E = (D . (h + F)) / 2 + E_nuc.  DIIS (Pulay) on the orthogonal-basis
// commutator e = X^T (F D S - S D F) X: the Gram matrix, the bordered system and the
// extrapolated Fock matrix stay on the device (the system solved by the library's LU); the host
// reads two numbers per cycle.
#include "scf.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "gemm.hpp"
#include "linalg.hpp"
#include "numint.hpp"
#include "rijk.hpp"
#include "xc_functionals.hpp"

namespace hh {

namespace {

constexpr int NDIIS = 8;

unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 65536)); }

// z = a x + b y
__global__ void lincomb_kernel(double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                               int64_t n, double* __restrict__ z) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    z[e] = a * x[e] + b * y[e];
}

// X[:, k] = U[:, k] / sqrt(s[k])
__global__ void inv_sqrt_cols_kernel(const double* __restrict__ U, const double* __restrict__ sv, int64_t n,
                                     double* __restrict__ X) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x)
    X[e] = U[e] / sqrt(sv[e / n]);
}

// per-block partial results: part[2 b] = sum x y, part[2 b + 1] = max |z|
__global__ __launch_bounds__(256) void dot_max_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                      const double* __restrict__ z, int64_t n,
                                                      double* __restrict__ part) {
  double d = 0.0, m = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    d += x[e] * y[e];
    m = fmax(m, fabs(z[e]));
  }
  __shared__ double sd[256], sm[256];
  sd[threadIdx.x] = d;
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      sd[threadIdx.x] += sd[threadIdx.x + h];
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sd[0];
    part[2 * blockIdx.x + 1] = sm[0];
  }
}

// G[i + NDIIS j] = <e_i, e_j> for i, j < m (slot order), one block per pair
__global__ __launch_bounds__(256) void diis_gram_kernel(const double* __restrict__ E, int64_t n, int m,
                                                        double* __restrict__ G) {
  const int i = blockIdx.x % m, j = blockIdx.x / m;
  const double* a = E + n * i;
  const double* b = E + n * j;
  double d = 0.0;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) d += a[e] * b[e];
  __shared__ double sd[256];
  sd[threadIdx.x] = d;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) sd[threadIdx.x] += sd[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) G[i + NDIIS * j] = sd[0];
}

// bordered Pulay system [[G, -1], [-1, 0]] c = [0, -1] of size m + 1 (column-major, ld m + 1),
// G scaled by its largest diagonal element
__global__ void diis_system_kernel(const double* __restrict__ G, int m, double* __restrict__ H,
                                   double* __restrict__ rhs) {
  const int n = m + 1;
  double sc = 0.0;
  for (int k = 0; k < m; ++k) sc = fmax(sc, G[k + NDIIS * k]);
  sc = sc > 0.0 ? 1.0 / sc : 1.0;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e % n, j = e / n;
    double v;
    if (i < m && j < m)
      v = G[i + NDIIS * j] * sc;
    else if (i == m && j == m)
      v = 0.0;
    else
      v = -1.0;
    H[e] = v;
  }
  for (int e = threadIdx.x; e < n; e += blockDim.x) rhs[e] = e == m ? -1.0 : 0.0;
}

// F = sum_k c[k] Fs[k]
__global__ void diis_combine_kernel(const double* __restrict__ Fs, const double* __restrict__ c, int m, int64_t n,
                                    double* __restrict__ F) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    double v = 0.0;
    for (int k = 0; k < m; ++k) v += c[k] * Fs[e + n * k];
    F[e] = v;
  }
}

struct Mat {
  DevBuf b;
  double* p = nullptr;
  void alloc(int64_t n) { p = static_cast<double*>(b.reserve(sizeof(double) * std::max<int64_t>(n, 1))); }
};

// C = alpha op(A) op(B) + beta C, square or rectangular, column-major
void mm(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, bool ta, const double* B, int64_t ldb, bool tb,
        double* C, int64_t ldc, double alpha, double beta, hipStream_t s, DevBuf& ws) {
  GemmArgs g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.lda = lda;
  g.a_kmajor = ta;  // A^T: element (m, k) at A[m lda + k]
  g.B = B;
  g.ldb = ldb;
  g.b_kmajor = !tb;  // B: element (k, n) at B[k + n ldb]; B^T: at B[k ldb + n]
  g.C = C;
  g.ldc = ldc;
  g.alpha = alpha;
  g.beta = beta;
  gemm<double, double>(g, PREC_F64, s, ws);
}

}  // namespace

namespace {

// The SCF of one (restricted) or two (unrestricted) spin channels.  Channel s has nocc[s]
// occupied orbitals with occupation f = 2 / nspin: D_s = f C_s,o C_s,o^T, F_s = h + J[D_a + D_b]
// - exx T_s T_s^T (T_s,P = B_P C_s,o, so T_s T_s^T = K[D_s] / f) + V_xc,s, and
// E = sum_s (D_s . (h + F_s^quad)) / 2 + E_xc + E_nuc with F^quad the Fock matrix without V_xc.
ScfResult scf_run(const Cint& mol, const Cint& aux, int nspin, const int* nocc, double conv_tol, int max_cycle,
                  double* const* mo_coeff, double* const* mo_energy, hipStream_t s, const ScfXc* xc) {
  const int64_t nao = mol.nao(), naux = aux.nao(), n2 = nao * nao;
  HH_REQUIRE(nspin == 1 || nspin == 2, "scf: nspin must be 1 or 2");
  for (int sp = 0; sp < nspin; ++sp)
    HH_REQUIRE(nocc[sp] >= 0 && nocc[sp] <= nao && (nspin == 2 || nocc[sp] > 0), "scf: nocc out of range");
  HH_REQUIRE(conv_tol > 0.0 && max_cycle > 0, "scf: conv_tol > 0, max_cycle > 0");
  for (int sp = 0; sp < nspin; ++sp) HH_REQUIRE(mo_coeff[sp] && mo_energy[sp], "scf: null output");
  const double focc = nspin == 1 ? 2.0 : 1.0;
  ScfResult res;
  for (int a = 0; a < mol.natm(); ++a)
    for (int b = 0; b < a; ++b) {
      const double* ra = mol.atom_xyz(a);
      const double* rb = mol.atom_xyz(b);
      const double d = sqrt((ra[0] - rb[0]) * (ra[0] - rb[0]) + (ra[1] - rb[1]) * (ra[1] - rb[1]) +
                            (ra[2] - rb[2]) * (ra[2] - rb[2]));
      res.e_nuc += mol.atom_charge(a) * mol.atom_charge(b) / d;
    }
  const int nomax = std::max(nocc[0], nspin == 2 ? nocc[1] : 0);
  DevBuf ws, lws;
  // per-channel matrices are stored channel after channel: F[s * n2], ...
  Mat S, h, X, sv, tmp, F, Fo, C, D, Dt, J, T, v, W1, W2, Fs, Es, G, H, rhs, part, Vev;
  const int64_t nsn2 = nspin * n2;
  S.alloc(n2); h.alloc(n2); X.alloc(n2); sv.alloc(nao); tmp.alloc(nsn2); F.alloc(nsn2); Fo.alloc(n2); C.alloc(nsn2);
  D.alloc(nsn2); Dt.alloc(n2); J.alloc(n2); T.alloc(nao * std::max(nomax, 1) * naux); v.alloc(naux); W1.alloc(n2);
  W2.alloc(n2); Fs.alloc(NDIIS * nsn2); Es.alloc(NDIIS * nsn2); G.alloc(NDIIS * NDIIS);
  H.alloc((NDIIS + 1) * (NDIIS + 1)); rhs.alloc(NDIIS + 1);
  constexpr int NPART = 256;
  part.alloc(2 * NPART);
  std::vector<double> hpart(2 * NPART);
  const double exx = xc ? xc->exx : 1.0;
  Mat rho, exc, vxc, V;
  int64_t ng = 0;
  int nvl = 1;
  if (xc) {
    HH_REQUIRE(xc->ni && xc->nfunc >= 0 && (xc->nfunc == 0 || (xc->ids && xc->scales)), "scf: incomplete XC spec");
    HH_REQUIRE(xc->ni->nao() == nao, "scf: the grid's AO count differs from the basis");
    HH_REQUIRE(xc->ni->stream() == s, "scf: the grid handle runs on another stream");
    ng = xc->ni->ng();
    nvl = xc->den == HH_DEN_RHO ? 1 : (xc->den == HH_DEN_SIGMA ? 4 : 5);
    rho.alloc(ng * nvl * nspin); exc.alloc(ng); vxc.alloc(ng * nvl * nspin); V.alloc(n2);
  }
  // one-electron matrices
  mol.int1e(HH_INT1E_OVLP, HH_INT_PLAIN, S.p, s);
  mol.int1e(HH_INT1E_KIN, HH_INT_PLAIN, h.p, s);
  mol.int1e_nuc(HH_INT_PLAIN, -1, tmp.p, s);
  hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(n2)), dim3(256), 0, s, 1.0, h.p, 1.0, tmp.p, n2, h.p);
  HH_HIP(hipGetLastError());
  // X = U s^-1/2
  HH_HIP(hipMemcpyAsync(X.p, S.p, sizeof(double) * n2, hipMemcpyDeviceToDevice, s));
  eigh(X.p, nao, sv.p, s, lws);
  {
    std::vector<double> hs(nao);
    HH_HIP(hipMemcpyAsync(hs.data(), sv.p, sizeof(double) * nao, hipMemcpyDeviceToHost, s));
    HH_HIP(hipStreamSynchronize(s));
    HH_REQUIRE(hs[0] > 1e-10 * hs[nao - 1], "scf: linearly dependent basis (overlap eigenvalue below 1e-10)");
  }
  hipLaunchKernelGGL(inv_sqrt_cols_kernel, dim3(grid_for(n2)), dim3(256), 0, s, X.p, sv.p, nao, X.p);
  HH_HIP(hipGetLastError());
  // cderi of the RI metric (Cholesky with the eigenvalue-floor refactorisation at 1e-14)
  std::unique_ptr<Rijk> rijk;
  {
    Mat j3c, j2c;
    j3c.alloc(nao * (nao + 1) / 2 * naux);
    j2c.alloc(naux * naux);
    Cint::int3c2e_s2ij(mol, aux, 0, aux.nbas(), j3c.p, s);
    aux.int2c2e(HH_INT_PLAIN, 0, aux.nbas(), j2c.p, s);
    rijk = Rijk::from_j3c(j3c.p, j2c.p, nao, naux, HH_J2C_CD, 1e-14, 1.0, 1.0, 1, s);
  }
  const double* B = rijk->bfull();
  // diagonalise F_s (orthogonal basis) into C_s, mo_energy[s]: F' = X^T F X in the basis of the
  // channel's last eigenvectors V_s, V_s <- V_s W with W the eigenvectors of V_s^T F' V_s
  Vev.alloc(nsn2);
  bool have_v[2] = {false, false};
  auto diagonalise = [&](int sp) {
    double* Fp = F.p + n2 * sp;
    double* Vs = Vev.p + n2 * sp;
    mm(nao, nao, nao, X.p, nao, true, Fp, nao, false, tmp.p, nao, 1.0, 0.0, s, ws);   // X^T F
    mm(nao, nao, nao, tmp.p, nao, false, X.p, nao, false, Fo.p, nao, 1.0, 0.0, s, ws);  // X^T F X
    if (have_v[sp]) {
      mm(nao, nao, nao, Vs, nao, true, Fo.p, nao, false, tmp.p, nao, 1.0, 0.0, s, ws);   // V^T F'
      mm(nao, nao, nao, tmp.p, nao, false, Vs, nao, false, Fo.p, nao, 1.0, 0.0, s, ws);  // V^T F' V
    }
    eigh(Fo.p, nao, mo_energy[sp], s, lws);
    if (have_v[sp]) {
      mm(nao, nao, nao, Vs, nao, false, Fo.p, nao, false, tmp.p, nao, 1.0, 0.0, s, ws);  // V W
      HH_HIP(hipMemcpyAsync(Vs, tmp.p, sizeof(double) * n2, hipMemcpyDeviceToDevice, s));
    } else {
      HH_HIP(hipMemcpyAsync(Vs, Fo.p, sizeof(double) * n2, hipMemcpyDeviceToDevice, s));
      have_v[sp] = true;
    }
    mm(nao, nao, nao, X.p, nao, false, Vs, nao, false, C.p + n2 * sp, nao, 1.0, 0.0, s, ws);  // C = X V
  };
  // core-Hamiltonian guess for every channel
  for (int sp = 0; sp < nspin; ++sp)
    HH_HIP(hipMemcpyAsync(F.p + n2 * sp, h.p, sizeof(double) * n2, hipMemcpyDeviceToDevice, s));
  double e_prev = 0.0;
  int ndiis = 0, slot = 0;
  const unsigned npb = (unsigned)std::min<int64_t>(NPART, ceil_div(n2, 256));
  for (int cycle = 1; cycle <= max_cycle; ++cycle) {
    for (int sp = 0; sp < nspin; ++sp) {
      diagonalise(sp);
      double* Dp = D.p + n2 * sp;
      if (nocc[sp] > 0)
        mm(nao, nao, nocc[sp], C.p + n2 * sp, nao, false, C.p + n2 * sp, nao, true, Dp, nao, focc, 0.0, s, ws);
      else
        HH_HIP(hipMemsetAsync(Dp, 0, sizeof(double) * n2, s));
    }
    // total density and J = B (B^T d)
    if (nspin == 2)
      hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(n2)), dim3(256), 0, s, 1.0, D.p, 1.0, D.p + n2, n2, Dt.p);
    else
      HH_HIP(hipMemcpyAsync(Dt.p, D.p, sizeof(double) * n2, hipMemcpyDeviceToDevice, s));
    HH_HIP(hipGetLastError());
    mm(naux, 1, n2, B, n2, true, Dt.p, n2, false, v.p, naux, 1.0, 0.0, s, ws);
    mm(n2, 1, naux, B, n2, false, v.p, naux, false, J.p, n2, 1.0, 0.0, s, ws);
    double e_quad = 0.0, err = 0.0;
    for (int sp = 0; sp < nspin; ++sp) {
      double* Fp = F.p + n2 * sp;
      hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(n2)), dim3(256), 0, s, 1.0, h.p, 1.0, J.p, n2, Fp);
      HH_HIP(hipGetLastError());
      if (nocc[sp] > 0) {
        // T_P = B_P C_s,o (batched over P); F_s -= exx T T^T
        GemmArgs g;
        g.M = nao;
        g.N = nocc[sp];
        g.K = nao;
        g.A = B;
        g.lda = nao;
        g.a_kmajor = false;
        g.B = C.p + n2 * sp;
        g.ldb = nao;
        g.b_kmajor = true;
        g.C = T.p;
        g.ldc = nao;
        g.nb0 = naux;
        g.sA0 = n2;
        g.sB0 = 0;
        g.sC0 = nao * nocc[sp];
        gemm<double, double>(g, PREC_F64, s, ws);
        mm(nao, nao, (int64_t)nocc[sp] * naux, T.p, nao, false, T.p, nao, true, Fp, nao, -exx, 1.0, s, ws);
      }
      // (D_s . (h + F_s)) / 2 of the quadratic part
      hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(n2)), dim3(256), 0, s, 0.5, h.p, 0.5, Fp, n2, tmp.p);
      HH_HIP(hipGetLastError());
      hipLaunchKernelGGL(dot_max_kernel, dim3(npb), dim3(256), 0, s, D.p + n2 * sp, tmp.p, tmp.p, n2, part.p);
      HH_HIP(hipGetLastError());
      HH_HIP(hipMemcpyAsync(hpart.data(), part.p, sizeof(double) * 2 * npb, hipMemcpyDeviceToHost, s));
      HH_HIP(hipStreamSynchronize(s));
      for (unsigned b = 0; b < npb; ++b) e_quad += hpart[2 * b];
    }
    double e_xc = 0.0;
    if (xc) {
      // rho of the channel densities on the grid, the functional list's exc / vxc_eff, F_s += V_xc,s
      const double* dptr[2] = {D.p, D.p + n2};
      xc->ni->setup_rho(xc->den, nspin, dptr, rho.p);
      HH_HIP(hipMemsetAsync(exc.p, 0, sizeof(double) * ng, s));
      HH_HIP(hipMemsetAsync(vxc.p, 0, sizeof(double) * ng * nvl * nspin, s));
      for (int f = 0; f < xc->nfunc; ++f)
        xc_eval_eff(xc->ids[f], nspin, nvl, rho.p, ng, ng * nvl, ng, xc->scales[f], true, exc.p, vxc.p, nullptr, s);
      for (int sp = 0; sp < nspin; ++sp) {
        e_xc += xc->ni->xc_energy(exc.p, rho.p + ng * nvl * sp);
        xc->ni->vxc_matrix(xc->den, vxc.p + ng * nvl * sp, V.p);
        hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(n2)), dim3(256), 0, s, 1.0, F.p + n2 * sp, 1.0, V.p, n2,
                           F.p + n2 * sp);
        HH_HIP(hipGetLastError());
      }
    }
    // commutators X^T (F_s D_s S - S D_s F_s) X into the DIIS slot (channels stacked)
    double* e_slot = Es.p + nsn2 * slot;
    for (int sp = 0; sp < nspin; ++sp) {
      mm(nao, nao, nao, F.p + n2 * sp, nao, false, D.p + n2 * sp, nao, false, W1.p, nao, 1.0, 0.0, s, ws);  // F D
      mm(nao, nao, nao, W1.p, nao, false, S.p, nao, false, W2.p, nao, 1.0, 0.0, s, ws);  // F D S
      mm(nao, nao, nao, W2.p, nao, true, X.p, nao, false, W1.p, nao, -1.0, 0.0, s, ws);  // -(FDS)^T X
      mm(nao, nao, nao, W2.p, nao, false, X.p, nao, false, W1.p, nao, 1.0, 1.0, s, ws);  // (FDS - SDF) X
      mm(nao, nao, nao, X.p, nao, true, W1.p, nao, false, e_slot + n2 * sp, nao, 1.0, 0.0, s, ws);
    }
    HH_HIP(hipMemcpyAsync(Fs.p + nsn2 * slot, F.p, sizeof(double) * nsn2, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(dot_max_kernel, dim3((unsigned)std::min<int64_t>(NPART, ceil_div(nsn2, 256))), dim3(256), 0, s,
                       e_slot, e_slot, e_slot, nsn2, part.p);
    HH_HIP(hipGetLastError());
    const unsigned npe = (unsigned)std::min<int64_t>(NPART, ceil_div(nsn2, 256));
    HH_HIP(hipMemcpyAsync(hpart.data(), part.p, sizeof(double) * 2 * npe, hipMemcpyDeviceToHost, s));
    HH_HIP(hipStreamSynchronize(s));
    for (unsigned b = 0; b < npe; ++b) err = std::max(err, hpart[2 * b + 1]);
    const double energy = e_quad + e_xc + res.e_nuc;
    HH_REQUIRE(std::isfinite(energy), "scf: non-finite energy");
    res.cycles = cycle;
    res.energy = energy;
    res.last_delta = std::fabs(energy - e_prev);
    res.last_error = err;
    if (cycle > 1 && res.last_delta < conv_tol && err < std::sqrt(conv_tol)) {
      res.converged = true;
      break;
    }
    e_prev = energy;
    // DIIS extrapolation of the stacked Fock matrices over the stored slots
    ndiis = std::min(ndiis + 1, NDIIS);
    slot = (slot + 1) % NDIIS;
    if (ndiis >= 2) {
      hipLaunchKernelGGL(diis_gram_kernel, dim3((unsigned)(ndiis * ndiis)), dim3(256), 0, s, Es.p, nsn2, ndiis, G.p);
      hipLaunchKernelGGL(diis_system_kernel, dim3(1), dim3(128), 0, s, G.p, ndiis, H.p, rhs.p);
      HH_HIP(hipGetLastError());
      lu_solve(H.p, ndiis + 1, rhs.p, 1, s, lws);
      hipLaunchKernelGGL(diis_combine_kernel, dim3(grid_for(nsn2)), dim3(256), 0, s, Fs.p, rhs.p, ndiis, nsn2, F.p);
      HH_HIP(hipGetLastError());
    }
  }
  // canonical orbitals of the last Fock matrices (built from the converged densities, not
  // extrapolated); on failure the last cycle's orbitals
  if (res.converged)
    for (int sp = 0; sp < nspin; ++sp) diagonalise(sp);
  for (int sp = 0; sp < nspin; ++sp)
    HH_HIP(hipMemcpyAsync(mo_coeff[sp], C.p + n2 * sp, sizeof(double) * n2, hipMemcpyDeviceToDevice, s));
  HH_HIP(hipStreamSynchronize(s));
  if (!res.converged)
    throw Error(HH_ERR_NOT_CONVERGED, "scf: not converged in " + std::to_string(max_cycle) + " cycles (dE " +
                                          std::to_string(res.last_delta) + ", error " +
                                          std::to_string(res.last_error) + ")");
  return res;
}

}  // namespace

ScfResult rhf_scf(const Cint& mol, const Cint& aux, int nocc, double conv_tol, int max_cycle, double* mo_coeff,
                  double* mo_energy, hipStream_t s, const ScfXc* xc) {
  int no[1] = {nocc};
  double* c[1] = {mo_coeff};
  double* e[1] = {mo_energy};
  return scf_run(mol, aux, 1, no, conv_tol, max_cycle, c, e, s, xc);
}

ScfResult uhf_scf(const Cint& mol, const Cint& aux, int nalpha, int nbeta, double conv_tol, int max_cycle,
                  double* const* mo_coeff, double* const* mo_energy, hipStream_t s, const ScfXc* xc) {
  int no[2] = {nalpha, nbeta};
  return scf_run(mol, aux, 2, no, conv_tol, max_cycle, mo_coeff, mo_energy, s, xc);
}

}  // namespace hh
