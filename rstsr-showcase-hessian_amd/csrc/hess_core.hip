// One-electron Hessian terms on the device (see hess_core.hpp).
//
// The core and overlap terms contract second-derivative integrals [nao, nao, 9] with a density
// over atom blocks.  Every term is a sum over (mu in A, nu in B) of X[mu, nu, q] M[mu, nu] for
// some atoms A, B, so one reduction kernel forms R[q, A, B] = sum_{mu in A, nu in B} X M for all
// atom pairs (and for every nucleus C of the rinv integrals), and a small assembly kernel combines
// row and column sums of R into the [3, 3, natm, natm] blocks:
//   overlap, B <= A:  de[., ., B, A] = -2 R(s_ipovlpip, W)[A, B] - 2 [A == B] sum_B' R(s_ipipovlp, W)[A, B']
//   core,   A == B:   de = 2 [Tot(raa_A) + Tot(rab_A) - 2 Row(raa_A)[A] - Row(rab_A)[A] - Col(rab_A)[A]
//                            + Row(h_aa)[A] + R(h_ab)[A, A]]
//           A != B:   de = 2 [R(h_ab)[A, B] - Row(raa_A)[B] - Row(rab_A)^T[B] - Row(raa_B)[A] - Row(rab_B)[A]]
// with h_aa = ipipkin + ipipnuc, h_ab = ipkinip + ipnucip, raa_C = -Z_C ipiprinv(C),
// rab_C = -Z_C iprinvip(C), all contracted with the density D; the factor 2 is the AO transpose
// of generator_hcore_deriv2 against a symmetric D.  Blocks with B > A are the component transposes
// of (A, B).  The [3, 3] index (i, j) is component q = i + 3 j of the 9-component integrals.
#include "hess_core.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

#include "gemm.hpp"

namespace hh {

namespace {

// R[q + nq (A + natm (B + natm c))] = sum_{mu in A, nu in B} X[mu + nao (nu + nao (c + nc q))] M[mu + nao nu]
__global__ __launch_bounds__(256) void block_reduce_kernel(const double* __restrict__ X, int64_t nao, int nc, int nq,
                                                           const double* __restrict__ M,
                                                           const int64_t* __restrict__ aoloc, int natm,
                                                           double* __restrict__ R) {
  const int A = blockIdx.x % natm, B = blockIdx.x / natm, c = blockIdx.y;
  const int64_t a0 = aoloc[2 * A], a1 = aoloc[2 * A + 1], b0 = aoloc[2 * B], b1 = aoloc[2 * B + 1];
  const int64_t na = a1 - a0, nblk = na * (b1 - b0);
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = 0.0;
  const int64_t n2 = nao * nao;
  for (int64_t e = threadIdx.x; e < nblk; e += blockDim.x) {
    const int64_t mu = a0 + e % na, nu = b0 + e / na;
    const int64_t off = mu + nao * nu;
    const double m = M[off];
    for (int q = 0; q < nq; ++q) acc[q] += X[off + n2 * (c + (int64_t)nc * q)] * m;
  }
  __shared__ double red[256];
  for (int q = 0; q < nq; ++q) {
    red[threadIdx.x] = acc[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) R[q + nq * (A + (int64_t)natm * (B + (int64_t)natm * c))] = red[0];
    __syncthreads();
  }
}

__device__ __forceinline__ int qT(int q) { return 3 * (q % 3) + q / 3; }

// overlap assembly: Raa, Rab [9, natm, natm] from ipipovlp / ipovlpip against W
__global__ void ovlp_assemble_kernel(const double* __restrict__ Raa, const double* __restrict__ Rab, int natm,
                                     double* __restrict__ out) {
  const int64_t total = 9 * (int64_t)natm * natm;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int q0 = (int)(e % 9);
  const int X = (int)((e / 9) % natm), Y = (int)(e / (9 * (int64_t)natm));
  // de[i, j, X, Y]: computed at (B, A) = (X, Y) when X <= Y, else the transpose of (Y, X)
  int A, B, q;
  if (X <= Y) { B = X; A = Y; q = q0; } else { B = Y; A = X; q = qT(q0); }
  auto Rv = [&](const double* R, int a, int b) { return R[q + 9 * (a + (int64_t)natm * b)]; };
  double v = -2.0 * Rv(Rab, A, B);
  if (A == B) {
    double row = 0.0;
    for (int b = 0; b < natm; ++b) row += Rv(Raa, A, b);
    v -= 2.0 * row;
  }
  out[e] = v;
}

// core assembly: Rh_aa, Rh_ab [9, natm, natm] (h integrals against D); Rr_aa, Rr_ab
// [9, natm, natm, natm] (rinv integrals of every nucleus C, already weighted by -Z_C)
__global__ void hcore_assemble_kernel(const double* __restrict__ Rhaa, const double* __restrict__ Rhab,
                                      const double* __restrict__ Rraa, const double* __restrict__ Rrab, int natm,
                                      double* __restrict__ out) {
  const int64_t total = 9 * (int64_t)natm * natm;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int q0 = (int)(e % 9);
  const int X = (int)((e / 9) % natm), Y = (int)(e / (9 * (int64_t)natm));
  int A, B, q;
  if (X <= Y) { B = X; A = Y; q = q0; } else { B = Y; A = X; q = qT(q0); }
  const int64_t N = natm;
  auto R2 = [&](const double* R, int qq, int a, int b) { return R[qq + 9 * (a + N * b)]; };
  auto R3 = [&](const double* R, int qq, int a, int b, int c) { return R[qq + 9 * (a + N * (b + N * c))]; };
  auto row3 = [&](const double* R, int qq, int a, int c) {
    double t = 0.0;
    for (int b = 0; b < natm; ++b) t += R3(R, qq, a, b, c);
    return t;
  };
  double v;
  if (A == B) {
    double tot_aa = 0.0, tot_ab = 0.0, col_ab = 0.0, row_h = 0.0;
    for (int a = 0; a < natm; ++a) {
      tot_aa += row3(Rraa, q, a, A);
      tot_ab += row3(Rrab, q, a, A);
      col_ab += R3(Rrab, q, a, A, A);
      row_h += R2(Rhaa, q, A, a);
    }
    v = tot_aa + tot_ab - 2.0 * row3(Rraa, q, A, A) - row3(Rrab, q, A, A) - col_ab + row_h + R2(Rhab, q, A, A);
  } else {
    v = R2(Rhab, q, A, B) - row3(Rraa, q, B, A) - row3(Rrab, qT(q), B, A) - row3(Rraa, q, A, B) -
        row3(Rrab, q, A, B);
  }
  out[e] = 2.0 * v;
}

__global__ void nuc_repl_kernel(const double* __restrict__ Z, const double* __restrict__ R, int natm,
                                double* __restrict__ out) {
  const int64_t total = 9 * (int64_t)natm * natm;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int i = (int)(e % 3), j = (int)((e / 3) % 3);
  const int A = (int)((e / 9) % natm), B = (int)(e / (9 * (int64_t)natm));
  // pair term of (A, C): Z_A Z_C (delta_ij / r^3 - 3 r_i r_j / r^5), r = R_A - R_C
  auto pair = [&](int a, int c) {
    const double r[3] = {R[3 * a] - R[3 * c], R[3 * a + 1] - R[3 * c + 1], R[3 * a + 2] - R[3 * c + 2]};
    const double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    const double d = sqrt(d2);
    const double zz = Z[a] * Z[c];
    return zz * ((i == j ? 1.0 : 0.0) / (d2 * d) - 3.0 * r[i] * r[j] / (d2 * d2 * d));
  };
  double v = 0.0;
  if (A != B) {
    v = pair(A, B);
  } else {
    for (int c = 0; c < natm; ++c)
      if (c != A) v -= pair(A, c);
  }
  out[e] = v;
}

// h1[mu, nu, x, A] = H[mu, nu, x] + H[nu, mu, x], H = -(ip rows of atom A) - Z_A iprinv_A
// (rinv [nao, nao, natm, 3]); with rinv null (overlap): H = -(ip rows of atom A)
__global__ void deriv1_kernel(const double* __restrict__ ip, const double* __restrict__ rinv,
                              const double* __restrict__ Z, const int64_t* __restrict__ aoloc, int64_t nao, int natm,
                              double* __restrict__ out) {
  const int64_t n2 = nao * nao, total = n2 * 3 * natm;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t mu = e % nao, nu = (e / nao) % nao;
    const int x = (int)((e / n2) % 3), A = (int)(e / (3 * n2));
    const int64_t a0 = aoloc[2 * A], a1 = aoloc[2 * A + 1];
    double v = 0.0;
    if (mu >= a0 && mu < a1) v -= ip[mu + nao * nu + n2 * x];
    if (nu >= a0 && nu < a1) v -= ip[nu + nao * mu + n2 * x];
    if (rinv) {
      const double* r = rinv + n2 * (A + (int64_t)natm * x);
      v -= Z[A] * (r[mu + nao * nu] + r[nu + nao * mu]);
    }
    out[e] = v;
  }
}

// out[mu, i] = w[i] C[mu, i]
__global__ void scale_cols_kernel(const double* __restrict__ C, int64_t nao, int64_t nmo, const double* __restrict__ w,
                                  double* __restrict__ out) {
  const int64_t total = nao * nmo;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x)
    out[e] = C[e] * w[e / nao];
}

// y[i] += x[i]
__global__ void axpy_kernel(const double* __restrict__ x, int64_t n, double* __restrict__ y) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    y[e] += x[e];
}

unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 65536)); }

struct AtomTables {
  DevBuf aoloc, z, xyz;
  int natm = 0;
  AtomTables(const Cint& mol, hipStream_t s) {
    HH_REQUIRE(mol.atoms_contiguous(), "Hessian: the basis shells must be grouped by atom");
    natm = mol.natm();
    std::vector<int64_t> loc(2 * natm);
    std::vector<double> zz(natm), xx(3 * natm);
    for (int a = 0; a < natm; ++a) {
      loc[2 * a] = mol.atom_ao0(a);
      loc[2 * a + 1] = mol.atom_ao1(a);
      zz[a] = mol.atom_charge(a);
      for (int d = 0; d < 3; ++d) xx[3 * a + d] = mol.atom_xyz(a)[d];
    }
    aoloc.reserve(sizeof(int64_t) * loc.size());
    z.reserve(sizeof(double) * natm);
    xyz.reserve(sizeof(double) * 3 * natm);
    HH_HIP(hipMemcpyAsync(aoloc.ptr, loc.data(), sizeof(int64_t) * loc.size(), hipMemcpyHostToDevice, s));
    HH_HIP(hipMemcpyAsync(z.ptr, zz.data(), sizeof(double) * natm, hipMemcpyHostToDevice, s));
    HH_HIP(hipMemcpyAsync(xyz.ptr, xx.data(), sizeof(double) * 3 * natm, hipMemcpyHostToDevice, s));
  }
};

void reduce(const double* X, int64_t nao, int nc, const double* M, const AtomTables& at, double* R, hipStream_t s) {
  hipLaunchKernelGGL(block_reduce_kernel, dim3((unsigned)(at.natm * at.natm), (unsigned)nc), dim3(256), 0, s, X, nao,
                     nc, 9, M, at.aoloc.as<int64_t>(), at.natm, R);
  HH_HIP(hipGetLastError());
}

}  // namespace

void nuc_repl_hess(const Cint& mol, double* out, hipStream_t s) {
  HH_REQUIRE(out, "nuc_repl_hess: null output");
  AtomTables at(mol, s);
  const int64_t total = 9 * (int64_t)at.natm * at.natm;
  hipLaunchKernelGGL(nuc_repl_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, s, at.z.as<double>(),
                     at.xyz.as<double>(), at.natm, out);
  HH_HIP(hipGetLastError());
  HH_HIP(hipStreamSynchronize(s));  // the tables are freed on return
}

void ovlp_hess(const Cint& mol, const double* dme0, double* out, hipStream_t s) {
  HH_REQUIRE(dme0 && out, "ovlp_hess: null argument");
  AtomTables at(mol, s);
  const int64_t nao = mol.nao(), n2 = nao * nao, natm = at.natm;
  DevBuf buf;
  double* s2 = static_cast<double*>(buf.reserve(sizeof(double) * (9 * n2 + 2 * 9 * natm * natm)));
  double* Raa = s2 + 9 * n2;
  double* Rab = Raa + 9 * natm * natm;
  mol.int1e(HH_INT1E_OVLP, HH_INT_IPIP1, s2, s);
  reduce(s2, nao, 1, dme0, at, Raa, s);
  mol.int1e(HH_INT1E_OVLP, HH_INT_IPVIP1, s2, s);
  reduce(s2, nao, 1, dme0, at, Rab, s);
  const int64_t total = 9 * natm * natm;
  hipLaunchKernelGGL(ovlp_assemble_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, s, Raa, Rab,
                     (int)natm, out);
  HH_HIP(hipGetLastError());
  HH_HIP(hipStreamSynchronize(s));
}

void hcore_hess(const Cint& mol, const double* dm0, double* out, hipStream_t s) {
  HH_REQUIRE(dm0 && out, "hcore_hess: null argument");
  AtomTables at(mol, s);
  const int64_t nao = mol.nao(), n2 = nao * nao, natm = at.natm;
  const int64_t n9 = 9 * natm * natm;
  DevBuf ints, tmp, red;
  double* X = static_cast<double*>(ints.reserve(sizeof(double) * 9 * n2 * natm));
  double* Y = static_cast<double*>(tmp.reserve(sizeof(double) * 9 * n2));
  double* R = static_cast<double*>(red.reserve(sizeof(double) * (2 * n9 + 2 * n9 * natm)));
  double* Rhaa = R;
  double* Rhab = Rhaa + n9;
  double* Rraa = Rhab + n9;
  double* Rrab = Rraa + n9 * natm;
  // h_aa = ipipkin + ipipnuc, h_ab = ipkinip + ipnucip (X as scratch for the nuclear term)
  auto accumulate = [&](const double* src, double* dst, int64_t n) {
    hipLaunchKernelGGL(axpy_kernel, dim3(grid_of(n)), dim3(256), 0, s, src, n, dst);
    HH_HIP(hipGetLastError());
  };
  mol.int1e(HH_INT1E_KIN, HH_INT_IPIP1, Y, s);
  mol.int1e_nuc(HH_INT_IPIP1, -1, X, s);
  accumulate(X, Y, 9 * n2);
  reduce(Y, nao, 1, dm0, at, Rhaa, s);
  mol.int1e(HH_INT1E_KIN, HH_INT_IPVIP1, Y, s);
  mol.int1e_nuc(HH_INT_IPVIP1, -1, X, s);
  accumulate(X, Y, 9 * n2);
  reduce(Y, nao, 1, dm0, at, Rhab, s);
  // rinv at every nucleus, weighted by -Z_C after the reduction
  std::vector<double> w(natm);
  for (int c = 0; c < natm; ++c) w[c] = -mol.atom_charge(c);
  DevBuf dw;
  double* wd = static_cast<double*>(dw.reserve(sizeof(double) * natm));
  HH_HIP(hipMemcpyAsync(wd, w.data(), sizeof(double) * natm, hipMemcpyHostToDevice, s));
  for (int kind : {HH_INT_IPIP1, HH_INT_IPVIP1}) {
    double* Rr = kind == HH_INT_IPIP1 ? Rraa : Rrab;
    mol.int1e_rinv_all(kind, X, s);
    reduce(X, nao, (int)natm, dm0, at, Rr, s);
    // scale the block of nucleus c by -Z_c: Rr[(q, A, B), c] as an [n9, natm] matrix times diag
    hipLaunchKernelGGL(scale_cols_kernel, dim3(grid_of(n9 * natm)), dim3(256), 0, s, Rr, n9, natm, wd, Rr);
    HH_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(hcore_assemble_kernel, dim3((unsigned)ceil_div(n9, 256)), dim3(256), 0, s, Rhaa, Rhab, Rraa,
                     Rrab, (int)natm, out);
  HH_HIP(hipGetLastError());
  HH_HIP(hipStreamSynchronize(s));
}

void hcore_deriv1(const Cint& mol, double* out, hipStream_t s) {
  HH_REQUIRE(out, "hcore_deriv1: null output");
  AtomTables at(mol, s);
  const int64_t nao = mol.nao(), n2 = nao * nao, natm = at.natm;
  DevBuf a, b, c;
  double* ip = static_cast<double*>(a.reserve(sizeof(double) * 3 * n2));
  double* tmp = static_cast<double*>(b.reserve(sizeof(double) * 3 * n2));
  double* rinv = static_cast<double*>(c.reserve(sizeof(double) * 3 * n2 * natm));
  mol.int1e(HH_INT1E_KIN, HH_INT_IP1, ip, s);
  mol.int1e_nuc(HH_INT_IP1, -1, tmp, s);
  hipLaunchKernelGGL(axpy_kernel, dim3(grid_of(3 * n2)), dim3(256), 0, s, tmp, 3 * n2, ip);  // ip += tmp
  HH_HIP(hipGetLastError());
  mol.int1e_rinv_all(HH_INT_IP1, rinv, s);
  hipLaunchKernelGGL(deriv1_kernel, dim3(grid_of(3 * n2 * natm)), dim3(256), 0, s, ip, rinv, at.z.as<double>(),
                     at.aoloc.as<int64_t>(), nao, (int)natm, out);
  HH_HIP(hipGetLastError());
  HH_HIP(hipStreamSynchronize(s));
}

void ovlp_deriv1(const Cint& mol, double* out, hipStream_t s) {
  HH_REQUIRE(out, "ovlp_deriv1: null output");
  AtomTables at(mol, s);
  const int64_t nao = mol.nao(), n2 = nao * nao, natm = at.natm;
  DevBuf a;
  double* ip = static_cast<double*>(a.reserve(sizeof(double) * 3 * n2));
  mol.int1e(HH_INT1E_OVLP, HH_INT_IP1, ip, s);
  hipLaunchKernelGGL(deriv1_kernel, dim3(grid_of(3 * n2 * natm)), dim3(256), 0, s, ip, nullptr, at.z.as<double>(),
                     at.aoloc.as<int64_t>(), nao, (int)natm, out);
  HH_HIP(hipGetLastError());
  HH_HIP(hipStreamSynchronize(s));
}

void weighted_density(const double* C, int64_t nao, int64_t nmo, const double* w, double* out, hipStream_t s) {
  HH_REQUIRE(C && w && out && nao > 0 && nmo > 0, "weighted_density: invalid argument");
  DevBuf wb, cw, gws;
  double* wd = static_cast<double*>(wb.reserve(sizeof(double) * nmo));
  double* Cw = static_cast<double*>(cw.reserve(sizeof(double) * nao * nmo));
  HH_HIP(hipMemcpyAsync(wd, w, sizeof(double) * nmo, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(scale_cols_kernel, dim3(grid_of(nao * nmo)), dim3(256), 0, s, C, nao, nmo, wd, Cw);
  HH_HIP(hipGetLastError());
  GemmArgs g;  // out = Cw C^T
  g.M = nao;
  g.N = nao;
  g.K = nmo;
  g.A = Cw;
  g.lda = nao;
  g.a_kmajor = false;
  g.B = C;
  g.ldb = nao;
  g.b_kmajor = false;
  g.C = out;
  g.ldc = nao;
  gemm<double, double>(g, PREC_F64, s, gws);
  HH_HIP(hipStreamSynchronize(s));
}

}  // namespace hh
