// Dimensionless CPHF driver (reference: src/hessian/rscf.rs).
//
// The dimensionless equation U + resp(U)/(e_a - e_i) = -B/(e_a - e_i) is solved for the vir-occ
// block with the occ-occ block frozen at -S1/2 (rscf.rs:88-99, 237-263).  All tensors stay on
// the device; perturbations are the trailing `nprop` dimension ([nmo, nocc, nprop]).
#include "cphf.hpp"

#include <cmath>

#include "elementwise.hpp"

namespace hh {

namespace {
unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 8192)); }

// rhs[m, i, a]: vir rows -(f1 - s1 e_i)/(e_a - e_i + shift); occ rows -s1/2
__global__ void dimless_rhs_kernel(const double* f1, const double* s1, const double* eocc, const double* evir,
                                   int64_t nmo, int64_t nocc, int64_t nprop, double shift, double* rhs) {
  const int64_t total = nmo * nocc * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = e % nmo, i = (e / nmo) % nocc;
    if (m >= nocc) rhs[e] = -(f1[e] - s1[e] * eocc[i]) / (evir[m - nocc] - eocc[i] + shift);
    else rhs[e] = -0.5 * s1[e];
  }
}

// resp[vir] = (resp - shift mo1)[vir] / (e_ai + shift); resp[occ] = 0
__global__ void dimless_scale_kernel(double* resp, const double* mo1, const double* eocc, const double* evir,
                                     int64_t nmo, int64_t nocc, int64_t nprop, double shift) {
  const int64_t total = nmo * nocc * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = e % nmo, i = (e / nmo) % nocc;
    if (m >= nocc) {
      double r = resp[e];
      if (shift != 0.0) r -= shift * mo1[e];
      resp[e] = r / (evir[m - nocc] - eocc[i] + shift);
    } else {
      resp[e] = 0.0;
    }
  }
}

// b1 = f1 - s1 e_i + resp (in place in resp);  mo1_out[vir] = -b1/e_ai, mo1_out[occ] = mo1[occ];
// mo_e1[j, i] = b1[j, i] + mo1[j, i] (e_j - e_i)   (rscf.rs:343-349)
__global__ void finalize_kernel(const double* f1, const double* s1, const double* mo1, double* b1, const double* eocc,
                                const double* evir, int64_t nmo, int64_t nocc, int64_t nprop, double* mo1_out,
                                double* mo_e1) {
  const int64_t total = nmo * nocc * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = e % nmo, i = (e / nmo) % nocc, a = e / (nmo * nocc);
    const double b = f1[e] - s1[e] * eocc[i] + b1[e];
    if (m >= nocc) {
      mo1_out[e] = -b / (evir[m - nocc] - eocc[i]);
    } else {
      mo1_out[e] = mo1[e];
      mo_e1[m + nocc * (i + nocc * a)] = b + mo1[e] * (eocc[m] - eocc[i]);
    }
  }
}

__global__ void scale_by_eocc_kernel(const double* s1, const double* eocc, int64_t nmo, int64_t nocc, int64_t nprop,
                                     double* out) {
  const int64_t total = nmo * nocc * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x)
    out[e] = s1[e] * eocc[(e / nmo) % nocc];
}

// copy s1mo[occ rows] -> compact [nocc, nocc, nprop]
__global__ void occ_rows_kernel(const double* s1, int64_t nmo, int64_t nocc, int64_t nprop, double* out) {
  const int64_t total = nocc * nocc * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e % nocc, rest = e / nocc;
    out[e] = s1[j + nmo * rest];
  }
}

// de[s, t, B, A] from G[(t,A), (s,B)] (ld 3 natm) for B <= A; mirrored for B > A
__global__ void assemble_hess_kernel(const double* G, int64_t natm, double* de) {
  const int64_t n3 = 3 * natm;
  const int64_t total = 9 * natm * natm;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = e % 3, t = (e / 3) % 3, B = (e / 9) % natm, A = e / (9 * natm);
    double v;
    if (B <= A) v = G[(t + 3 * A) + n3 * (s + 3 * B)];
    else v = G[(s + 3 * B) + n3 * (t + 3 * A)];  // de[s,t,B,A] = de[t,s,A,B] with A < B
    de[e] = v;
  }
}

// y = d * x on vir rows (occ rows: y = x)
__global__ void apply_diag_kernel(const double* x, const double* d, int64_t nmo, int64_t nocc, int64_t nprop,
                                  double* y) {
  const int64_t nn = nmo * nocc, total = nn * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = e % nmo;
    y[e] = (m >= nocc) ? d[e % nn] * x[e] : x[e];
  }
}

// ay = (d - 1) y + ay  on vir rows
__global__ void precond_fix_kernel(const double* y, const double* d, int64_t nmo, int64_t nocc, int64_t nprop,
                                   double* ay) {
  const int64_t nn = nmo * nocc, total = nn * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = e % nmo;
    if (m >= nocc) ay[e] += (d[e % nn] - 1.0) * y[e];
  }
}

// r = b - x - ax; colnorm[a] = ||r[:, a]||^2 (one block per column)
__global__ void residual_norm_kernel(const double* b, const double* x, const double* ax, double* r, int64_t n,
                                     double* colnorm) {
  __shared__ double red[256];
  const int64_t a = blockIdx.x;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const int64_t e = i + n * a;
    const double v = b[e] - x[e] - ax[e];
    r[e] = v;
    acc += v * v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) colnorm[a] = red[0];
}

// r[:, j] *= 1/||r_j|| (or 0 when ||r_j|| <= floor); sc[j] = the applied factor
__global__ void scale_columns_kernel(double* r, int64_t n, int64_t ncol, const double* coln2, double floor,
                                     double* sc) {
  const int64_t total = n * ncol;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / n;
    const double nr = sqrt(coln2[j]);
    const double f = nr > floor ? 1.0 / nr : 0.0;
    r[e] *= f;
    if (e % n == 0) sc[j] = f;
  }
}

// y[:, j] += x[:, j] / sc[j]  (sc[j] = 0: column untouched)
__global__ void add_unscaled_kernel(double* y, const double* x, int64_t n, int64_t ncol, const double* sc) {
  const int64_t total = n * ncol;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const double f = sc[e / n];
    if (f != 0.0) y[e] += x[e] / f;
  }
}

__global__ void axpy_kernel(double* y, const double* x, double alpha, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    y[e] += alpha * x[e];
}

// raw features (before standardisation) for sample (a, i)
__global__ void raw_features_kernel(const double* pair, const double* r, const double* eocc, const double* evir,
                                    int64_t nmo, int64_t nocc, int64_t nprop, float* x) {
  const int64_t nvir = nmo - nocc, ns = nvir * nocc;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ns; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = e % nvir, i = e / nvir, m = nocc + a;
    double rr = 0.0;
    for (int64_t p = 0; p < nprop; ++p) {
      const double v = r[m + nmo * (i + nocc * p)];
      rr += v * v;
    }
    x[e] = (float)log(fmax(evir[a] - eocc[i], 1e-8));
    x[e + ns] = (float)log(pair[m + nmo * i] + 1e-30);
    x[e + 2 * ns] = (float)log(sqrt(rr / (double)nprop) + 1e-30);
  }
}

// standardise each of the 3 feature columns in place (one workgroup per column, FP64 sums)
__global__ __launch_bounds__(256) void standardise_kernel(float* x, int64_t ns) {
  __shared__ double a[256], b[256];
  float* col = x + ns * blockIdx.x;
  double sa = 0.0, sb = 0.0;
  for (int64_t i = threadIdx.x; i < ns; i += blockDim.x) {
    sa += col[i];
    sb += (double)col[i] * col[i];
  }
  a[threadIdx.x] = sa;
  b[threadIdx.x] = sb;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      a[threadIdx.x] += a[threadIdx.x + st];
      b[threadIdx.x] += b[threadIdx.x + st];
    }
    __syncthreads();
  }
  const double mean = a[0] / (double)ns;
  const double var = fmax(b[0] / (double)ns - mean * mean, 0.0);
  const double inv = var > 1e-20 ? 1.0 / sqrt(var) : 1.0;
  for (int64_t i = threadIdx.x; i < ns; i += blockDim.x) col[i] = (float)((col[i] - mean) * inv);
}

__global__ void lsq_target_kernel(const double* r, const double* xs, int64_t nmo, int64_t nocc, int64_t nprop,
                                  float* t, float* w) {
  const int64_t nvir = nmo - nocc, ns = nvir * nocc;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ns; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = e % nvir, i = e / nvir, m = nocc + a;
    double rx = 0.0, rr = 0.0;
    for (int64_t p = 0; p < nprop; ++p) {
      const int64_t o = m + nmo * (i + nocc * p);
      rx += r[o] * xs[o];
      rr += r[o] * r[o];
    }
    const double d = rr > 0.0 ? rx / rr : 1.0;
    t[e] = (float)log(fmin(fmax(d, 0.05), 20.0));
    w[e] = (float)rr;
  }
}

__global__ void precond_diag_kernel(const float* y, int64_t nmo, int64_t nocc, double* diag) {
  const int64_t nvir = nmo - nocc, total = nmo * nocc;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = e % nmo, i = e / nmo;
    diag[e] = (m >= nocc) ? exp(fmin(fmax((double)y[(m - nocc) + nvir * i], -3.0), 3.0)) : 1.0;
  }
}

}  // namespace

Cphf::Cphf(const double* mo_coeff, int64_t nao, int64_t nmo, const double* mo_occ, const double* mo_energy,
           const CphfConfig& cfg, hipStream_t s)
    : nao_(nao), nmo_(nmo), cfg_(cfg), stream_(s) {
  HH_REQUIRE(mo_coeff && mo_occ && mo_energy && nao > 0 && nmo > 0, "cphf: null or empty MO data");
  std::vector<double> eo, ev;
  std::vector<int64_t> order;
  for (int64_t j = 0; j < nmo; ++j)
    if (mo_occ[j] > 0.0) eo.push_back(mo_energy[j]);
  for (int64_t j = 0; j < nmo; ++j)
    if (!(mo_occ[j] > 0.0)) ev.push_back(mo_energy[j]);
  // the reference slices occ = [0, nocc), vir = [nocc, nmo) (rscf.rs:172-177): occupied orbitals
  // must come first
  nocc_ = (int64_t)eo.size();
  HH_REQUIRE(nocc_ > 0 && nocc_ < nmo, "cphf: need both occupied and virtual orbitals");
  for (int64_t j = 0; j < nocc_; ++j) HH_REQUIRE(mo_occ[j] > 0.0, "cphf: occupied orbitals must come first");
  c_.reserve(sizeof(double) * nao * nmo);
  HH_HIP(hipMemcpyAsync(c_.ptr, mo_coeff, sizeof(double) * nao * nmo, hipMemcpyDeviceToDevice, s));
  eocc_.reserve(sizeof(double) * nocc_);
  evir_.reserve(sizeof(double) * (nmo - nocc_));
  HH_HIP(hipMemcpyAsync(eocc_.ptr, eo.data(), sizeof(double) * nocc_, hipMemcpyHostToDevice, s));
  HH_HIP(hipMemcpyAsync(evir_.ptr, ev.data(), sizeof(double) * (nmo - nocc_), hipMemcpyHostToDevice, s));
  HH_HIP(hipStreamSynchronize(s));
}

void Cphf::dimless_rhs(const double* f1mo, const double* s1mo, int64_t nprop, double* rhs) {
  hipLaunchKernelGGL(dimless_rhs_kernel, dim3(nblk(nmo_ * nocc_ * nprop)), dim3(256), 0, stream_, f1mo, s1mo,
                     eocc_.as<double>(), evir_.as<double>(), nmo_, nocc_, nprop, cfg_.level_shift, rhs);
  HH_HIP(hipGetLastError());
}

void Cphf::response_mo(const double* mo1, int64_t nprop, double* resp, Prec prec) {
  HH_REQUIRE(!comps_.empty(), "cphf: no electron-interaction component added");
  const int64_t ncol = nocc_ * nprop;
  double* ubra = static_cast<double*>(ubra_.reserve(sizeof(double) * nao_ * ncol));
  double* rbra = static_cast<double*>(rbra_.reserve(sizeof(double) * nao_ * ncol));
  GemmArgs g;  // ubra = C mo1
  g.M = nao_; g.N = ncol; g.K = nmo_;
  g.A = c_.ptr; g.lda = nao_; g.a_kmajor = false;
  g.B = mo1; g.ldb = nmo_; g.b_kmajor = true;
  g.C = ubra; g.ldc = nao_;
  gemm<double, double>(g, PREC_F64, stream_, ws_);
  for (size_t c = 0; c < comps_.size(); ++c) comps_[c]->response_bra(ubra, nprop, rbra, c > 0, prec);
  GemmArgs h;  // resp = C^T rbra
  h.M = nmo_; h.N = ncol; h.K = nao_;
  h.A = c_.ptr; h.lda = nao_; h.a_kmajor = true;
  h.B = rbra; h.ldb = nao_; h.b_kmajor = true;
  h.C = resp; h.ldc = nmo_;
  gemm<double, double>(h, PREC_F64, stream_, ws_);
}

void Cphf::response_dimless(const double* mo1, int64_t nprop, double* resp, Prec prec) {
  response_mo(mo1, nprop, resp, prec);
  hipLaunchKernelGGL(dimless_scale_kernel, dim3(nblk(nmo_ * nocc_ * nprop)), dim3(256), 0, stream_, resp, mo1,
                     eocc_.as<double>(), evir_.as<double>(), nmo_, nocc_, nprop, cfg_.level_shift);
  HH_HIP(hipGetLastError());
}

void Cphf::solve_dimless(const double* rhs, int64_t nprop, double* mo1, Prec prec, CphfStats* st, const double* diag) {
  const int64_t n = nmo_ * nocc_;
  const int64_t blk = n * nprop;
  CphfStats stats;
  double* y = static_cast<double*>(ydiag_.reserve(sizeof(double) * blk * 2));
  double* ytmp = y + blk;
  // aop on the (optionally right-preconditioned) system; `p` is the response precision
  auto make_aop = [&](Prec p) {
    return [this, p, diag, y, ytmp](const double* x, int64_t nb, double* ax) {
      if (!diag) {
        response_dimless(x, nb, ax, p);
        return;
      }
      // (I + A) D x = (I + A') x with A' x = (D - 1) x + A (D x)
      hipLaunchKernelGGL(apply_diag_kernel, dim3(nblk(nmo_ * nocc_ * nb)), dim3(256), 0, stream_, x, diag, nmo_,
                         nocc_, nb, ytmp);
      response_dimless(ytmp, nb, ax, p);
      hipLaunchKernelGGL(precond_fix_kernel, dim3(nblk(nmo_ * nocc_ * nb)), dim3(256), 0, stream_, x, diag, nmo_,
                         nocc_, nb, ax);
      HH_HIP(hipGetLastError());
    };
  };
  auto unprecond = [&](const double* yv, double* out, int64_t nb) {
    if (diag) {
      hipLaunchKernelGGL(apply_diag_kernel, dim3(nblk(n * nb)), dim3(256), 0, stream_, yv, diag, nmo_, nocc_, nb, out);
      HH_HIP(hipGetLastError());
    } else {
      HH_HIP(hipMemcpyAsync(out, yv, sizeof(double) * n * nb, hipMemcpyDeviceToDevice, stream_));
    }
  };
  auto accumulate = [&](const KrylovStats& k) {
    stats.krylov.cycles += k.cycles;
    stats.krylov.restarts += k.restarts;
    stats.krylov.aop_calls += k.aop_calls;
    stats.krylov.max_innerprod = k.max_innerprod;
  };
  if (prec == PREC_F64) {
    KrylovStats ks;
    krylov_block(make_aop(PREC_F64), rhs, nullptr, n, nprop, cfg_.tol, cfg_.max_cycle, cfg_.max_space, cfg_.lindep,
                 y, &ks, stream_);
    accumulate(ks);
    unprecond(y, mo1, nprop);
    if (st) *st = stats;
    return;
  }
  // Mixed precision: x_{k+1} = x_k + solve_mixed(r_k), r_k = rhs - (I + A_64) x_k in FP64.
  // Each inner solve runs on the residual with every column rescaled to unit norm (the columns
  // are independent systems): the reference's absolute lindep floor keeps its meaning as the
  // residual shrinks, and the non-normalised Krylov basis does not mix columns whose norms
  // differ by orders of magnitude (that makes the projected system ill-conditioned).  Columns
  // already far below the tolerance are left out.  Inner solves go to about the
  // reduced-precision operator's own accuracy; the outer loop stops when every column's FP64
  // residual 2-norm is below tol.
  double* r = static_cast<double*>(res_.reserve(sizeof(double) * (3 * blk + 2 * nprop)));
  double* ax = r + blk;
  double* dx = ax + blk;
  double* coln = dx + blk;
  double* csc = coln + nprop;
  std::vector<double> hn(nprop);
  auto max_colnorm2 = [&]() {
    HH_HIP(hipMemcpyAsync(hn.data(), coln, sizeof(double) * nprop, hipMemcpyDeviceToHost, stream_));
    HH_HIP(hipStreamSynchronize(stream_));
    double m = 0.0;
    for (double v : hn) m = std::max(m, v);
    return m;
  };
  fill<double>(mo1, blk, 0.0, stream_);
  fill<double>(ax, blk, 0.0, stream_);
  hipLaunchKernelGGL(residual_norm_kernel, dim3((unsigned)nprop), dim3(256), 0, stream_, rhs, mo1, ax, r, n, coln);
  HH_HIP(hipGetLastError());
  double res = max_colnorm2();
  const double op_tol = (prec == PREC_BF16X3) ? 1e-5 : 1e-7;  // inner target, relative to the rescaled rhs
  const int max_sweeps = 10;
  for (int sweep = 0; sweep < max_sweeps && std::sqrt(res) >= cfg_.tol && res > 0.0; ++sweep) {
    const double sc = 1.0 / std::sqrt(res);
    hipLaunchKernelGGL(scale_columns_kernel, dim3(nblk(blk)), dim3(256), 0, stream_, r, n, nprop, coln,
                       0.01 * cfg_.tol, csc);
    HH_HIP(hipGetLastError());
    KrylovStats ks;
    // the trial-vector norm that krylov_block tests sits about an order of magnitude below the
    // true residual, so ask for 1 % of the reduction this sweep needs (and at most 1e-2)
    const double tol_in = std::min(1e-2, std::max(op_tol, 0.01 * cfg_.tol * sc));
    try {
      krylov_block(make_aop(prec), r, nullptr, n, nprop, tol_in, cfg_.max_cycle, cfg_.max_space, cfg_.lindep, y, &ks,
                   stream_);
    } catch (const Error& e) {
      if (e.code != HH_ERR_NOT_CONVERGED) throw;
      // a partially converged inner solve is still a correction; the FP64 residual decides
    }
    accumulate(ks);
    unprecond(y, dx, nprop);
    hipLaunchKernelGGL(add_unscaled_kernel, dim3(nblk(blk)), dim3(256), 0, stream_, mo1, dx, n, nprop, csc);
    HH_HIP(hipGetLastError());
    response_dimless(mo1, nprop, ax, PREC_F64);
    ++stats.krylov.aop_calls;
    hipLaunchKernelGGL(residual_norm_kernel, dim3((unsigned)nprop), dim3(256), 0, stream_, rhs, mo1, ax, r, n, coln);
    HH_HIP(hipGetLastError());
    res = max_colnorm2();
    stats.refinements = sweep + 1;
    if (verbose) {
      fprintf(stderr, "cphf refinement sweep %d: inner cycles %d (tol %.1e), FP64 residual %.3e\n", sweep + 1,
              ks.cycles, tol_in, std::sqrt(res));
      fflush(stderr);
    }
  }
  stats.residual = std::sqrt(res);
  if (st) *st = stats;
  if (!(std::sqrt(res) < cfg_.tol)) {
    char buf[200];
    snprintf(buf, sizeof buf, "cphf mixed-precision refinement did not reach tol %.3e (residual %.3e after %d sweeps)",
             cfg_.tol, std::sqrt(res), stats.refinements);
    throw Error(HH_ERR_NOT_CONVERGED, buf);
  }
}

void Cphf::finalize(const double* f1mo, const double* s1mo, const double* mo1, int64_t nprop, double* mo1_out,
                    double* mo_e1) {
  const int64_t blk = nmo_ * nocc_ * nprop;
  double* b1 = static_cast<double*>(tmp_.reserve(sizeof(double) * blk));
  response_mo(mo1, nprop, b1, PREC_F64);
  hipLaunchKernelGGL(finalize_kernel, dim3(nblk(blk)), dim3(256), 0, stream_, f1mo, s1mo, mo1, b1, eocc_.as<double>(),
                     evir_.as<double>(), nmo_, nocc_, nprop, mo1_out, mo_e1);
  HH_HIP(hipGetLastError());
}

void Cphf::precond_features(const double* pair, const double* r, int64_t nprop, float* x) {
  const int64_t ns = (nmo_ - nocc_) * nocc_;
  hipLaunchKernelGGL(raw_features_kernel, dim3(nblk(ns)), dim3(256), 0, stream_, pair, r, eocc_.as<double>(),
                     evir_.as<double>(), nmo_, nocc_, nprop, x);
  hipLaunchKernelGGL(standardise_kernel, dim3(3), dim3(256), 0, stream_, x, ns);
  HH_HIP(hipGetLastError());
}

void Cphf::precond_target(const double* r, const double* xs, int64_t nprop, float* t, float* w) {
  const int64_t ns = (nmo_ - nocc_) * nocc_;
  hipLaunchKernelGGL(lsq_target_kernel, dim3(nblk(ns)), dim3(256), 0, stream_, r, xs, nmo_, nocc_, nprop, t, w);
  HH_HIP(hipGetLastError());
}

void Cphf::precond_diag(const float* y, double* diag) {
  hipLaunchKernelGGL(precond_diag_kernel, dim3(nblk(nmo_ * nocc_)), dim3(256), 0, stream_, y, nmo_, nocc_, diag);
  HH_HIP(hipGetLastError());
}

void Cphf::cphf_hess(const double* f1mo, const double* s1mo, const double* mo1, const double* mo_e1, int64_t natm,
                     double* de) {
  const int64_t n3 = 3 * natm, nn = nmo_ * nocc_, no2 = nocc_ * nocc_;
  double* G = static_cast<double*>(g1_.reserve(sizeof(double) * n3 * n3));
  double* se = static_cast<double*>(sx_.reserve(sizeof(double) * nn * n3));
  double* soo = static_cast<double*>(g2_.reserve(sizeof(double) * no2 * n3));
  hipLaunchKernelGGL(scale_by_eocc_kernel, dim3(nblk(nn * n3)), dim3(256), 0, stream_, s1mo, eocc_.as<double>(), nmo_,
                     nocc_, n3, se);
  hipLaunchKernelGGL(occ_rows_kernel, dim3(nblk(no2 * n3)), dim3(256), 0, stream_, s1mo, nmo_, nocc_, n3, soo);
  HH_HIP(hipGetLastError());
  // G[(t,A),(s,B)] = 4 f1^T mo1 - 4 (s1 e)^T mo1 - 2 s1oo^T mo_e1
  GemmArgs g;
  g.M = n3; g.N = n3; g.K = nn;
  g.A = f1mo; g.lda = nn; g.a_kmajor = true;
  g.B = mo1; g.ldb = nn; g.b_kmajor = true;
  g.C = G; g.ldc = n3;
  g.alpha = 4.0; g.beta = 0.0;
  gemm<double, double>(g, PREC_F64, stream_, ws_);
  g.A = se; g.alpha = -4.0; g.beta = 1.0;
  gemm<double, double>(g, PREC_F64, stream_, ws_);
  GemmArgs h;
  h.M = n3; h.N = n3; h.K = no2;
  h.A = soo; h.lda = no2; h.a_kmajor = true;
  h.B = mo_e1; h.ldb = no2; h.b_kmajor = true;
  h.C = G; h.ldc = n3;
  h.alpha = -2.0; h.beta = 1.0;
  gemm<double, double>(h, PREC_F64, stream_, ws_);
  hipLaunchKernelGGL(assemble_hess_kernel, dim3(nblk(9 * natm * natm)), dim3(256), 0, stream_, G, natm, de);
  HH_HIP(hipGetLastError());
}

}  // namespace hh
