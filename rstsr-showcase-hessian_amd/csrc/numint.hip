// Grid contractions of the RKS CPKS response and the XC effective-kernel transform.
//
// The reference loops grid chunks (384 points) x perturbation sets with per-set GEMMs and
// Mutex-guarded accumulation (pure_eval_rho.rs:216-334, pure_xcpot.rs:370-465).  Here every
// perturbation is a column of one GEMM, per grid batch:
//   aob[g, i, c]     = sum_mu ao[g, mu, c] C_o[mu, i]                (c < ncomp)
//   akb[g, (i,a), c] = sum_mu ao[g, mu, c] bra[mu, (i,a)]            (matrix cores)
//   rho1[g, x, a], wv[g, y, a] = w_g sum_x fxc[g, x, y] rho1[g, x, a], and the weighted
//   back-contraction operands S_c[g, (i,a)]                           (one pass, FP64)
//   out[mu, (i,a)]  += sum_c sum_g ao[g, mu, c] S_c[g, (i,a)]         (matrix cores)
// so rho1 never leaves the batch and accumulation is a GEMM beta, not a lock.
#include "numint.hpp"

#include "elementwise.hpp"

namespace hh {

namespace {
unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 65535)); }

// rho1 of one (g, a) pair from the half-transformed AOs.
template <int DEN>
__device__ __forceinline__ void rho1_point(const double* aob, const double* akb, int64_t gb, int64_t g, int64_t nocc,
                                           int64_t ncol, int64_t a, double* r) {
  const int64_t sc_b = gb * nocc, sc_k = gb * ncol;
  for (int x = 0; x < 5; ++x) r[x] = 0.0;
  for (int64_t i = 0; i < nocc; ++i) {
    const double b0 = aob[g + gb * i];
    const double k0 = akb[g + gb * (i + nocc * a)];
    r[0] += b0 * k0;
    if (DEN != DEN_RHO) {
#pragma unroll
      for (int t = 1; t < 4; ++t) {
        const double bt = aob[g + gb * i + sc_b * t];
        const double kt = akb[g + gb * (i + nocc * a) + sc_k * t];
        r[t] += b0 * kt + bt * k0;
        if (DEN == DEN_TAU) r[4] += 0.5 * bt * kt;
      }
    }
  }
}

// wv[y] = w sum_x fxc[g, x, y] rho1[x]
template <int DEN>
__device__ __forceinline__ void weighted_kernel(const double* fxc, int64_t ld_fxc, int64_t g, double w, const double* r,
                                                double* wv) {
  constexpr int NV = DEN == DEN_RHO ? 1 : (DEN == DEN_SIGMA ? 4 : 5);
#pragma unroll
  for (int y = 0; y < NV; ++y) {
    double v = 0.0;
#pragma unroll
    for (int x = 0; x < NV; ++x) v += fxc[g + ld_fxc * (x + NV * y)] * r[x];
    wv[y] = w * v;
  }
}

// write S_c[g, (i,a)] (into `S`, which may alias akb: each (g, a) owns its entries)
template <int DEN>
__device__ __forceinline__ void write_s(const double* aob, double* S, int64_t gb, int64_t g, int64_t nocc,
                                        int64_t ncol, int64_t a, const double* wv) {
  const int64_t sc_b = gb * nocc, sc_k = gb * ncol;
  for (int64_t i = 0; i < nocc; ++i) {
    const double b0 = aob[g + gb * i];
    const int64_t o = g + gb * (i + nocc * a);
    if (DEN == DEN_RHO) {
      S[o] = b0 * wv[0];
    } else {
      const double b1 = aob[g + gb * i + sc_b], b2 = aob[g + gb * i + 2 * sc_b], b3 = aob[g + gb * i + 3 * sc_b];
      S[o] = b0 * wv[0] + b1 * wv[1] + b2 * wv[2] + b3 * wv[3];
      const double h = (DEN == DEN_TAU) ? 0.5 * wv[4] : 0.0;
      S[o + sc_k] = b0 * wv[1] + h * b1;
      S[o + 2 * sc_k] = b0 * wv[2] + h * b2;
      S[o + 3 * sc_k] = b0 * wv[3] + h * b3;
    }
  }
}

template <int DEN>
__global__ void rho1_kernel(const double* aob, const double* akb, int64_t gb, int64_t nocc, int64_t nprop,
                            double* rho1, int64_t ng, int64_t g0) {
  constexpr int NV = DEN == DEN_RHO ? 1 : (DEN == DEN_SIGMA ? 4 : 5);
  const int64_t total = gb * nprop, ncol = nocc * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = e % gb, a = e / gb;
    double r[5];
    rho1_point<DEN>(aob, akb, gb, g, nocc, ncol, a, r);
#pragma unroll
    for (int x = 0; x < NV; ++x) rho1[(g0 + g) + ng * (x + NV * a)] = r[x];
  }
}

// fused: rho1 from (aob, akb), then S into akb
template <int DEN>
__global__ void rho1_wv_s_kernel(const double* aob, double* akb, const double* fxc, const double* w, int64_t gb,
                                 int64_t nocc, int64_t nprop, int64_t ng, int64_t g0) {
  const int64_t total = gb * nprop, ncol = nocc * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = e % gb, a = e / gb;
    double r[5], wv[5];
    rho1_point<DEN>(aob, akb, gb, g, nocc, ncol, a, r);
    weighted_kernel<DEN>(fxc + g0, ng, g, w[g0 + g], r, wv);
    write_s<DEN>(aob, akb, gb, g, nocc, ncol, a, wv);
  }
}

// S from a given rho1 [ng, nvar, nprop]
template <int DEN>
__global__ void wv_s_kernel(const double* aob, const double* rho1, double* S, const double* fxc, const double* w,
                            int64_t gb, int64_t nocc, int64_t nprop, int64_t ng, int64_t g0) {
  constexpr int NV = DEN == DEN_RHO ? 1 : (DEN == DEN_SIGMA ? 4 : 5);
  const int64_t total = gb * nprop, ncol = nocc * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = e % gb, a = e / gb;
    double r[5], wv[5];
#pragma unroll
    for (int x = 0; x < NV; ++x) r[x] = rho1[(g0 + g) + ng * (x + NV * a)];
    weighted_kernel<DEN>(fxc + g0, ng, g, w[g0 + g], r, wv);
    write_s<DEN>(aob, S, gb, g, nocc, ncol, a, wv);
  }
}

// ---- xceff transform (xc_deriv.rs:305-428, unpolarised) ----
template <int DEN>
__global__ void xceff_kernel(int order, const double* rho, int64_t ld_rho, const double* xc, int64_t ld_xc, int64_t ng,
                             double* out) {
  constexpr int NV = DEN == DEN_SIGMA ? 4 : 5;
  constexpr int XL = DEN == DEN_SIGMA ? 2 : 3;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x) {
    double c[NV];
    int var[NV];
    c[0] = 1.0;
    var[0] = 0;
#pragma unroll
    for (int t = 1; t < 4; ++t) {
      c[t] = 2.0 * rho[g + ld_rho * t];
      var[t] = 1;
    }
    if (NV == 5) {
      c[NV - 1] = 1.0;
      var[NV - 1] = 2;
    }
    if (order == 1) {
#pragma unroll
      for (int x = 0; x < NV; ++x) out[g + ng * x] = c[x] * xc[g + ld_xc * (1 + var[x])];
    } else {
      const double fsig = xc[g + ld_xc * 2];
#pragma unroll
      for (int y = 0; y < NV; ++y)
#pragma unroll
        for (int x = 0; x < NV; ++x) {
          const int p = var[x] < var[y] ? var[x] : var[y], q = var[x] < var[y] ? var[y] : var[x];
          const int k = p * XL - p * (p - 1) / 2 + (q - p);  // combinations_with_replacement index
          double v = c[x] * c[y] * xc[g + ld_xc * (1 + XL + k)];
          if (x == y && x >= 1 && x <= 3) v += 2.0 * fsig;
          out[g + ng * (x + NV * y)] = v;
        }
    }
  }
}

__global__ void xceff_lda_kernel(int order, const double* xc, int64_t ld_xc, int64_t ng, double* out) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x)
    out[g] = xc[g + ld_xc * order];
}

}  // namespace

Numint::Numint(const double* ao, int64_t ng, int64_t nao, int64_t ncomp, const double* weights, hipStream_t s)
    : ng_(ng), nao_(nao), ncomp_(ncomp), stream_(s) {
  HH_REQUIRE(ao && weights && ng > 0 && nao > 0 && (ncomp == 1 || ncomp >= 4), "numint: bad AO data");
  ao_.reserve(sizeof(double) * ng * nao * ncomp);
  HH_HIP(hipMemcpyAsync(ao_.ptr, ao, sizeof(double) * ng * nao * ncomp, hipMemcpyDeviceToDevice, s));
  w_.reserve(sizeof(double) * ng);
  HH_HIP(hipMemcpyAsync(w_.ptr, weights, sizeof(double) * ng, hipMemcpyDeviceToDevice, s));
}

int64_t Numint::pick_batch(int64_t ncol) const {
  if (batch > 0) return std::min(batch, ng_);
  const int64_t per_point = 8 * (ncol + 64) * 4;
  int64_t gb = (int64_t)(1LL << 30) / std::max<int64_t>(per_point, 1);
  gb = std::max<int64_t>(256, (gb / 256) * 256);
  return std::min(gb, ng_);
}

#define HH_DEN_DISPATCH(den, KERNEL, ...)                                   \
  do {                                                                      \
    if ((den) == DEN_RHO) hipLaunchKernelGGL(KERNEL<DEN_RHO>, __VA_ARGS__); \
    else if ((den) == DEN_SIGMA) hipLaunchKernelGGL(KERNEL<DEN_SIGMA>, __VA_ARGS__); \
    else hipLaunchKernelGGL(KERNEL<DEN_TAU>, __VA_ARGS__);                  \
    HH_HIP(hipGetLastError());                                              \
  } while (0)

namespace {
// aob[g, i, c] and akb[g, (i,a), c] for one grid batch
void half_transform(const double* ao, int64_t ng, int64_t nao, int64_t g0, int64_t gb, int ncomp, const double* B,
                    int64_t ncolB, double* out, Prec prec, hipStream_t s, DevBuf& ws) {
  GemmArgs g;
  g.M = gb; g.N = ncolB; g.K = nao;
  g.A = ao + g0; g.lda = ng; g.a_kmajor = false; g.sA0 = ng * nao;
  g.B = B; g.ldb = nao; g.b_kmajor = true; g.sB0 = 0;
  g.C = out; g.ldc = gb; g.sC0 = gb * ncolB;
  g.nb0 = ncomp;
  gemm<double, double>(g, prec, s, ws);
}

// out[mu, n] (+)= alpha sum_c sum_g ao[g0+g, mu, c] S_c[g, n]
void back_contract(const double* ao, int64_t ng, int64_t nao, int64_t g0, int64_t gb, int ncomp, const double* S,
                   int64_t ncol, double* out, double alpha, bool accumulate, Prec prec, hipStream_t s, DevBuf& ws) {
  for (int c = 0; c < ncomp; ++c) {
    GemmArgs g;
    g.M = nao; g.N = ncol; g.K = gb;
    g.A = ao + g0 + ng * nao * c; g.lda = ng; g.a_kmajor = true;
    g.B = S + gb * ncol * c; g.ldb = gb; g.b_kmajor = true;
    g.C = out; g.ldc = nao;
    g.alpha = alpha;
    g.beta = (accumulate || c > 0) ? 1.0 : 0.0;
    gemm<double, double>(g, prec, s, ws);
  }
}
}  // namespace

void Numint::rho1_one_bra_mult_ket(int den, const double* bra, int64_t nocc, const double* kets, int64_t nprop,
                                   double* rho1, Prec prec) {
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "numint: den_type must be RHO, SIGMA or TAU");
  const int nc = den_ncomp(den);
  HH_REQUIRE(ncomp_ >= nc, "numint: AO component dimension insufficient");
  const int64_t ncol = nocc * nprop;
  const int64_t gbmax = pick_batch(ncol);
  double* aob = static_cast<double*>(aob_.reserve(sizeof(double) * gbmax * nocc * nc));
  double* akb = static_cast<double*>(akb_.reserve(sizeof(double) * gbmax * ncol * nc));
  for (int64_t g0 = 0; g0 < ng_; g0 += gbmax) {
    const int64_t gb = std::min(gbmax, ng_ - g0);
    half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, bra, nocc, aob, prec, stream_, ws_);
    half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, kets, ncol, akb, prec, stream_, ws_);
    HH_DEN_DISPATCH(den, rho1_kernel, dim3(nblk(gb * nprop)), dim3(256), 0, stream_, aob, akb, gb, nocc, nprop, rho1,
                    ng_, g0);
  }
}

void Numint::fxc_pot_bra_trans(int den, const double* fxc_eff, const double* rho1, int64_t nprop, const double* bra,
                               int64_t nocc, double* out, Prec prec, double alpha, bool accumulate) {
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "numint: den_type must be RHO, SIGMA or TAU");
  const int nc = den_ncomp(den);
  HH_REQUIRE(ncomp_ >= nc, "numint: AO component dimension insufficient");
  const int64_t ncol = nocc * nprop;
  const int64_t gbmax = pick_batch(ncol);
  double* aob = static_cast<double*>(aob_.reserve(sizeof(double) * gbmax * nocc * nc));
  double* S = static_cast<double*>(akb_.reserve(sizeof(double) * gbmax * ncol * nc));
  for (int64_t g0 = 0; g0 < ng_; g0 += gbmax) {
    const int64_t gb = std::min(gbmax, ng_ - g0);
    half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, bra, nocc, aob, prec, stream_, ws_);
    HH_DEN_DISPATCH(den, wv_s_kernel, dim3(nblk(gb * nprop)), dim3(256), 0, stream_, aob, rho1, S, fxc_eff,
                    w_.as<double>(), gb, nocc, nprop, ng_, g0);
    back_contract(ao_.as<double>(), ng_, nao_, g0, gb, nc, S, ncol, out, alpha, accumulate || g0 > 0, prec, stream_,
                  ws_);
  }
}

void Numint::rks_response_bra(int den, const double* fxc_eff, const double* mo1_bra, int64_t nprop,
                              const double* mocc, int64_t nocc, double* out, Prec prec, bool accumulate) {
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "numint: den_type must be RHO, SIGMA or TAU");
  const int nc = den_ncomp(den);
  HH_REQUIRE(ncomp_ >= nc, "numint: AO component dimension insufficient");
  const int64_t ncol = nocc * nprop;
  const int64_t gbmax = pick_batch(ncol);
  double* aob = static_cast<double*>(aob_.reserve(sizeof(double) * gbmax * nocc * nc));
  double* akb = static_cast<double*>(akb_.reserve(sizeof(double) * gbmax * ncol * nc));
  for (int64_t g0 = 0; g0 < ng_; g0 += gbmax) {
    const int64_t gb = std::min(gbmax, ng_ - g0);
    half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, mocc, nocc, aob, prec, stream_, ws_);
    half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, mo1_bra, ncol, akb, prec, stream_, ws_);
    HH_DEN_DISPATCH(den, rho1_wv_s_kernel, dim3(nblk(gb * nprop)), dim3(256), 0, stream_, aob, akb, fxc_eff,
                    w_.as<double>(), gb, nocc, nprop, ng_, g0);
    // the x4 closed-shell factor of hess_rks.rs:603
    back_contract(ao_.as<double>(), ng_, nao_, g0, gb, nc, akb, ncol, out, 4.0, accumulate || g0 > 0, prec, stream_,
                  ws_);
  }
}

void xceff_transform(int den, int order, const double* rho, int64_t ld_rho, const double* xc_val, int64_t ld_xc,
                     int64_t ng, double* out, hipStream_t s) {
  HH_REQUIRE(order == 1 || order == 2, "xceff: order must be 1 (vxc) or 2 (fxc)");
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "xceff: den_type must be RHO, SIGMA or TAU");
  if (ng == 0) return;
  if (den == DEN_RHO) {
    hipLaunchKernelGGL(xceff_lda_kernel, dim3(nblk(ng)), dim3(256), 0, s, order, xc_val, ld_xc, ng, out);
  } else if (den == DEN_SIGMA) {
    hipLaunchKernelGGL(xceff_kernel<DEN_SIGMA>, dim3(nblk(ng)), dim3(256), 0, s, order, rho, ld_rho, xc_val, ld_xc, ng,
                       out);
  } else {
    hipLaunchKernelGGL(xceff_kernel<DEN_TAU>, dim3(nblk(ng)), dim3(256), 0, s, order, rho, ld_rho, xc_val, ld_xc, ng,
                       out);
  }
  HH_HIP(hipGetLastError());
}

}  // namespace hh
