// Dimensionless CPHF driver (reference: src/hessian/rscf.rs, src/hessian/uscf.rs).
//
// The dimensionless equation U + resp(U)/(e_a - e_i) = -B/(e_a - e_i) is solved for the vir-occ
// block with the occ-occ block frozen at -S1/2 (rscf.rs:88-99, 237-263), per spin for UHF with
// the spin blocks stacked in one Krylov vector (uscf.rs:288-330).  All tensors stay on the
// device; perturbations are the trailing `nprop` dimension ([nmo, nocc, nprop] per spin).
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

// resp[vir] = (resp - shift mo1)[vir] / (e_ai + shift); resp[occ] = 0.  Column a of the
// [nmo, nocc] block starts at ld * a.
// *flag = 1 when a and b hold the same values (one workgroup, fixed order)
__global__ __launch_bounds__(1024) void equal_kernel(const double* a, const double* b, int64_t n, int* flag) {
  __shared__ int bad[16];
  int d = 0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) d |= (a[i] != b[i]);
  for (int o = 32; o > 0; o >>= 1) d |= __shfl_down(d, o, 64);
  if ((threadIdx.x & 63) == 0) bad[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    int any = 0;
    for (int w = 0; w < 16; ++w) any |= bad[w];
    *flag = any ? 0 : 1;
  }
}

__global__ void dimless_scale_kernel(double* resp, const double* mo1, const double* eocc, const double* evir,
                                     int64_t nmo, int64_t nocc, int64_t nprop, int64_t ld, double shift) {
  const int64_t nn = nmo * nocc, total = nn * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = e % nmo, i = (e / nmo) % nocc, at = e % nn + ld * (e / nn);
    if (m >= nocc) {
      double r = resp[at];
      if (shift != 0.0) r -= shift * mo1[at];
      resp[at] = r / (evir[m - nocc] - eocc[i] + shift);
    } else {
      resp[at] = 0.0;
    }
  }
}

// dst[r + ld_dst c] = src[r + ld_src c], r < rows, c < ncol
__global__ void copy_block_kernel(const double* src, int64_t ld_src, double* dst, int64_t ld_dst, int64_t rows,
                                  int64_t ncol) {
  const int64_t total = rows * ncol;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x)
    dst[e % rows + ld_dst * (e / rows)] = src[e % rows + ld_src * (e / rows)];
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

// x[:, j] /= sc[j]  (sc[j] = 0: column zeroed)
__global__ void unscale_columns_kernel(double* x, int64_t n, int64_t ncol, const double* sc) {
  const int64_t total = n * ncol;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const double f = sc[e / n];
    x[e] = f != 0.0 ? x[e] / f : 0.0;
  }
}

// colnorm[j] = sum_i x[i, j]^2 (one workgroup per column, fixed order)
__global__ void colnorm2_kernel(const double* x, int64_t n, double* colnorm) {
  __shared__ double red[256];
  const int64_t a = blockIdx.x;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = x[i + n * a];
    acc += v * v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if ((int)threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0) colnorm[a] = red[0];
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

Cphf::Cphf(int nspin, const double* const* mo_coeff, int64_t nao, const int64_t* nmo, const double* const* mo_occ,
           const double* const* mo_energy, const CphfConfig& cfg, hipStream_t s)
    : nspin_(nspin), nao_(nao), cfg_(cfg), stream_(s) {
  HH_REQUIRE(nspin == 1 || nspin == 2, "cphf: nspin must be 1 or 2");
  HH_REQUIRE(mo_coeff && mo_occ && mo_energy && nmo && nao > 0, "cphf: null MO data");
  for (int t = 0; t < nspin; ++t) {
    Spin& sp = sp_[t];
    HH_REQUIRE(mo_coeff[t] && mo_occ[t] && mo_energy[t] && nmo[t] > 0, "cphf: null or empty MO data");
    sp.nmo = nmo[t];
    std::vector<double> eo, ev;
    for (int64_t j = 0; j < sp.nmo; ++j) (mo_occ[t][j] > 0.0 ? eo : ev).push_back(mo_energy[t][j]);
    // the reference slices occ = [0, nocc), vir = [nocc, nmo) (rscf.rs:172-177, uscf.rs:165-168):
    // occupied orbitals must come first
    sp.nocc = (int64_t)eo.size();
    HH_REQUIRE(sp.nocc > 0 && sp.nocc < sp.nmo, "cphf: need both occupied and virtual orbitals");
    for (int64_t j = 0; j < sp.nocc; ++j) HH_REQUIRE(mo_occ[t][j] > 0.0, "cphf: occupied orbitals must come first");
    sp.off = n_;
    n_ += sp.nmo * sp.nocc;
    sp.c.reserve(sizeof(double) * nao * sp.nmo);
    HH_HIP(hipMemcpyAsync(sp.c.ptr, mo_coeff[t], sizeof(double) * nao * sp.nmo, hipMemcpyDeviceToDevice, s));
    sp.eocc.reserve(sizeof(double) * sp.nocc);
    sp.evir.reserve(sizeof(double) * (sp.nmo - sp.nocc));
    HH_HIP(hipMemcpyAsync(sp.eocc.ptr, eo.data(), sizeof(double) * sp.nocc, hipMemcpyHostToDevice, s));
    HH_HIP(hipMemcpyAsync(sp.evir.ptr, ev.data(), sizeof(double) * (sp.nmo - sp.nocc), hipMemcpyHostToDevice, s));
    HH_HIP(hipStreamSynchronize(s));
  }
}

void Cphf::add_component(ElecResponse* c) {
  HH_REQUIRE(c && c->nspin() == nspin_, "cphf: component spin count differs from the driver's");
  comps_.push_back(c);
  fin_nprop_ = -1;
  drop_graphs();  // a captured operator leaves the new component out
}

// ---- stacked vector <-> per-spin arrays (one spin: the same memory, no copy) ----
const double* Cphf::stack_in(const double* const* v, int64_t nprop, DevBuf& buf) {
  HH_REQUIRE(v, "cphf: null spin array");
  for (int t = 0; t < nspin_; ++t) HH_REQUIRE(v[t], "cphf: null spin array");
  if (nspin_ == 1) return v[0];
  double* flat = static_cast<double*>(buf.reserve(sizeof(double) * n_ * nprop));
  for (int t = 0; t < nspin_; ++t) {
    const int64_t rows = sp_[t].nmo * sp_[t].nocc;
    hipLaunchKernelGGL(copy_block_kernel, dim3(nblk(rows * nprop)), dim3(256), 0, stream_, v[t], rows,
                       flat + sp_[t].off, n_, rows, nprop);
  }
  HH_HIP(hipGetLastError());
  return flat;
}

double* Cphf::stack_out(double* const* v, int64_t nprop, DevBuf& buf) {
  HH_REQUIRE(v, "cphf: null spin array");
  for (int t = 0; t < nspin_; ++t) HH_REQUIRE(v[t], "cphf: null spin array");
  if (nspin_ == 1) return v[0];
  return static_cast<double*>(buf.reserve(sizeof(double) * n_ * nprop));
}

void Cphf::unstack(const double* flat, int64_t nprop, double* const* v) {
  if (nspin_ == 1) return;
  for (int t = 0; t < nspin_; ++t) {
    const int64_t rows = sp_[t].nmo * sp_[t].nocc;
    hipLaunchKernelGGL(copy_block_kernel, dim3(nblk(rows * nprop)), dim3(256), 0, stream_, flat + sp_[t].off, n_,
                       v[t], rows, rows, nprop);
  }
  HH_HIP(hipGetLastError());
}

void Cphf::dimless_rhs(const double* const* f1mo, const double* const* s1mo, int64_t nprop, double* const* rhs) {
  HH_REQUIRE(f1mo && s1mo && rhs, "cphf: null spin array");
  for (int t = 0; t < nspin_; ++t) {
    const Spin& sp = sp_[t];
    hipLaunchKernelGGL(dimless_rhs_kernel, dim3(nblk(sp.nmo * sp.nocc * nprop)), dim3(256), 0, stream_, f1mo[t],
                       s1mo[t], sp.eocc.as<double>(), sp.evir.as<double>(), sp.nmo, sp.nocc, nprop, cfg_.level_shift,
                       rhs[t]);
  }
  HH_HIP(hipGetLastError());
}

void Cphf::response_mo_stacked(const double* x, int64_t nprop, double* resp, Prec prec) {
  HH_REQUIRE(!comps_.empty(), "cphf: no electron-interaction component added");
  const double* ub[2] = {nullptr, nullptr};
  double* rb[2] = {nullptr, nullptr};
  for (int t = 0; t < nspin_; ++t) {
    Spin& sp = sp_[t];
    const int64_t ncol = sp.nocc * nprop;
    double* ubra = static_cast<double*>(sp.ubra.reserve(sizeof(double) * nao_ * ncol));
    rb[t] = static_cast<double*>(sp.rbra.reserve(sizeof(double) * nao_ * ncol));
    ub[t] = ubra;
    GemmArgs g;  // ubra = C mo1 (per perturbation column block when the spins are stacked)
    g.M = nao_; g.K = sp.nmo;
    g.A = sp.c.ptr; g.lda = nao_; g.a_kmajor = false;
    g.B = x + sp.off; g.ldb = sp.nmo; g.b_kmajor = true;
    g.C = ubra; g.ldc = nao_;
    if (nspin_ == 1) {
      g.N = ncol;
    } else {
      g.N = sp.nocc; g.nb0 = nprop; g.sA0 = 0; g.sB0 = n_; g.sC0 = nao_ * sp.nocc;
    }
    gemm<double, double>(g, PREC_F64, stream_, ws_);
  }
  for (size_t c = 0; c < comps_.size(); ++c) comps_[c]->response_bra(ub, nprop, rb, c > 0, prec);
  for (int t = 0; t < nspin_; ++t) {
    Spin& sp = sp_[t];
    GemmArgs h;  // resp = C^T rbra
    h.M = sp.nmo; h.K = nao_;
    h.A = sp.c.ptr; h.lda = nao_; h.a_kmajor = true;
    h.B = rb[t]; h.ldb = nao_; h.b_kmajor = true;
    h.C = resp + sp.off; h.ldc = sp.nmo;
    if (nspin_ == 1) {
      h.N = sp.nocc * nprop;
    } else {
      h.N = sp.nocc; h.nb0 = nprop; h.sA0 = 0; h.sB0 = nao_ * sp.nocc; h.sC0 = n_;
    }
    gemm<double, double>(h, PREC_F64, stream_, ws_);
  }
}

void Cphf::response_dimless_stacked(const double* x, int64_t nprop, double* resp, Prec prec) {
  response_mo_stacked(x, nprop, resp, prec);
  dimless_scale(x, nprop, resp);
}

void Cphf::dimless_scale(const double* x, int64_t nprop, double* resp) {
  for (int t = 0; t < nspin_; ++t) {
    const Spin& sp = sp_[t];
    hipLaunchKernelGGL(dimless_scale_kernel, dim3(nblk(sp.nmo * sp.nocc * nprop)), dim3(256), 0, stream_,
                       resp + sp.off, x + sp.off, sp.eocc.as<double>(), sp.evir.as<double>(), sp.nmo, sp.nocc, nprop,
                       n_, cfg_.level_shift);
  }
  HH_HIP(hipGetLastError());
}

// ---- the response operator, optionally replayed from HIP graphs ----
void Cphf::aop_body(const double* x, int64_t nb, double* ax, Prec p, const double* diag, double* ytmp) {
  if (!diag) {
    response_dimless_stacked(x, nb, ax, p);
    return;
  }
  // (I + A) D x = (I + A') x with A' x = (D - 1) x + A (D x)
  const int64_t nmo0 = sp_[0].nmo, nocc0 = sp_[0].nocc;
  hipLaunchKernelGGL(apply_diag_kernel, dim3(nblk(nmo0 * nocc0 * nb)), dim3(256), 0, stream_, x, diag, nmo0, nocc0,
                     nb, ytmp);
  response_dimless_stacked(ytmp, nb, ax, p);
  hipLaunchKernelGGL(precond_fix_kernel, dim3(nblk(nmo0 * nocc0 * nb)), dim3(256), 0, stream_, x, diag, nmo0, nocc0,
                     nb, ax);
  HH_HIP(hipGetLastError());
}

void Cphf::drop_graphs() {
  for (auto& g : graphs_)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  graphs_.clear();
}

Cphf::~Cphf() { drop_graphs(); }

void Cphf::aop(const double* x, int64_t nb, double* ax, Prec p, const double* diag, double* ytmp) {
  // (the legacy default stream cannot be captured)
  bool graphs = use_graphs && stream_ != nullptr;
  for (ElecResponse* c : comps_) graphs = graphs && c->stream() == stream_;
  if (!graphs) {
    aop_body(x, nb, ax, p, diag, ytmp);
    return;
  }
  if (graph_gen_ != devbuf_generation().load() || graph_state_ != response_state_epoch().load()) {
    // a buffer a graph may refer to was freed, or a component's state changed (re-prepared
    // orbitals: the graphs skip the reduced-precision copies made on the first eager call)
    drop_graphs();
    graph_gen_ = devbuf_generation().load();
    graph_state_ = response_state_epoch().load();
  }
  AopGraph* e = nullptr;
  for (auto& g : graphs_)
    if (g.x == x && g.resp == ax && g.nprop == nb && g.prec == (int)p && g.diag == diag && g.ytmp == ytmp) e = &g;
  if (!e) {
    if (graphs_.size() >= 32) drop_graphs();  // operands that never repeat: keep the cache small
    graphs_.push_back(AopGraph{x, ax, nb, (int)p, diag, ytmp, 0, nullptr});
    e = &graphs_.back();
  }
  if (e->exec) {
    HH_HIP(hipGraphLaunch(e->exec, stream_));
    ++n_launches_;
    return;
  }
  if (e->uses == 0) {  // first call eager: sizes the buffers and tunes the products
    e->uses = 1;
    aop_body(x, nb, ax, p, diag, ytmp);
    if (graph_gen_ != devbuf_generation().load()) {  // it moved buffers other graphs may refer to
      drop_graphs();
      graph_gen_ = devbuf_generation().load();
    }
    return;
  }
  hipGraph_t g = nullptr;
  HH_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  try {
    aop_body(x, nb, ax, p, diag, ytmp);
  } catch (...) {
    (void)hipStreamEndCapture(stream_, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HH_HIP(hipStreamEndCapture(stream_, &g));
  hipGraphExec_t exec = nullptr;
  const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HH_HIP(ie);
  if (graph_gen_ != devbuf_generation().load()) {  // a buffer moved while capturing: not replayable
    (void)hipGraphExecDestroy(exec);
    drop_graphs();
    graph_gen_ = devbuf_generation().load();
    aop_body(x, nb, ax, p, diag, ytmp);
    return;
  }
  e->exec = exec;
  ++n_captures_;
  HH_HIP(hipGraphLaunch(exec, stream_));
  ++n_launches_;
}

void Cphf::response_mo(const double* const* mo1, int64_t nprop, double* const* resp, Prec prec) {
  const double* x = stack_in(mo1, nprop, st_in_);
  double* r = stack_out(resp, nprop, st_out_);
  response_mo_stacked(x, nprop, r, prec);
  unstack(r, nprop, resp);
}

void Cphf::response_dimless(const double* const* mo1, int64_t nprop, double* const* resp, Prec prec) {
  const double* x = stack_in(mo1, nprop, st_in_);
  double* r = stack_out(resp, nprop, st_out_);
  aop(x, nprop, r, prec, nullptr, nullptr);
  unstack(r, nprop, resp);
}

void Cphf::solve_dimless(const double* const* rhs, int64_t nprop, double* const* mo1, Prec prec, CphfStats* st,
                         const double* diag) {
  HH_REQUIRE(!diag || nspin_ == 1, "cphf: the diagonal preconditioner is for one spin");
  const double* b = stack_in(rhs, nprop, st_in_);
  double* x = stack_out(mo1, nprop, st_out_);
  solve_stacked(b, nprop, x, prec, st, diag);
  unstack(x, nprop, mo1);
}

void Cphf::solve_stacked(const double* rhs, int64_t nprop, double* mo1, Prec prec, CphfStats* st,
                         const double* diag) {
  const int64_t n = n_;
  const int64_t blk = n * nprop;
  const int64_t nmo0 = sp_[0].nmo, nocc0 = sp_[0].nocc;
  CphfStats stats;
  fin_nprop_ = -1;
  double* y = static_cast<double*>(ydiag_.reserve(sizeof(double) * blk * 2));
  double* ytmp = y + blk;
  // aop on the (optionally right-preconditioned) system; `p` is the response precision
  (void)nmo0;
  (void)nocc0;
  auto make_aop = [&](Prec p) {
    return [this, p, diag, ytmp](const double* xv, int64_t nb, double* ax) { aop(xv, nb, ax, p, diag, ytmp); };
  };
  auto unprecond = [&](const double* yv, double* out, int64_t nb) {
    if (diag) {
      hipLaunchKernelGGL(apply_diag_kernel, dim3(nblk(n * nb)), dim3(256), 0, stream_, yv, diag, nmo0, nocc0, nb,
                         out);
      HH_HIP(hipGetLastError());
    } else {
      copy_bytes(out, yv, (int64_t)sizeof(double) * n * nb, stream_);
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
  // reduced-precision operator's own accuracy; after every sweep the residual is rebuilt from one
  // FP64 response of the new solution, and the outer loop stops when every column's FP64 residual
  // 2-norm is below tol.
  double* r = static_cast<double*>(res_.reserve(sizeof(double) * (3 * blk + 2 * nprop)));
  double* ax = r + blk;
  double* dx = ax + blk;
  double* coln = dx + blk;
  double* csc = coln + nprop;
  std::vector<double> hn(nprop);
  auto max_colnorm2 = [&](const double* cn) {
    HH_HIP(hipMemcpyAsync(hn.data(), cn, sizeof(double) * nprop, hipMemcpyDeviceToHost, stream_));
    HH_HIP(hipStreamSynchronize(stream_));
    double m = 0.0;
    for (double v : hn) m = std::max(m, v);
    return m;
  };
  fill<double>(mo1, blk, 0.0, stream_);
  fill<double>(ax, blk, 0.0, stream_);
  hipLaunchKernelGGL(residual_norm_kernel, dim3((unsigned)nprop), dim3(256), 0, stream_, rhs, mo1, ax, r, n, coln);
  HH_HIP(hipGetLastError());
  double res = max_colnorm2(coln);
  double* fin_r = static_cast<double*>(fin_r_.reserve(sizeof(double) * blk));
  const double op_tol = (prec == PREC_BF16X3) ? 1e-5 : 1e-7;  // inner target, relative to the rescaled rhs
  // Stagnation fallback: a sweep that gains less than 10x ends the mixed solve, and the
  // reference's plain FP64 block Krylov solves the system from scratch.  A warm start from mo1
  // is no option: krylov_block restarts from the residual, whose trial vectors then sit at the
  // absolute lindep floor at once.
  const int max_sweeps = 10;
  double prev_res = -1.0;
  for (int sweep = 0; sweep < max_sweeps && std::sqrt(res) >= cfg_.tol && res > 0.0; ++sweep) {
    if (prev_res > 0.0 && res > 0.01 * prev_res) {
      if (verbose) {
        fprintf(stderr, "cphf refinement stagnated at %.3e: plain FP64 solve\n", std::sqrt(res));
        fflush(stderr);
      }
      KrylovStats ks;
      krylov_block(make_aop(PREC_F64), rhs, nullptr, n, nprop, cfg_.tol, cfg_.max_cycle, cfg_.max_space,
                   cfg_.lindep, y, &ks, stream_);
      accumulate(ks);
      unprecond(y, mo1, nprop);
      stats.fallbacks = 1;
      // the true FP64 residual of the returned solution (krylov_block's own test is on the
      // trial-vector norms, about an order of magnitude below it)
      response_mo_stacked(mo1, nprop, ax, PREC_F64);
      ++stats.krylov.aop_calls;
      dimless_scale(mo1, nprop, ax);
      hipLaunchKernelGGL(residual_norm_kernel, dim3((unsigned)nprop), dim3(256), 0, stream_, rhs, mo1, ax, r, n,
                         coln);
      HH_HIP(hipGetLastError());
      stats.residual = std::sqrt(max_colnorm2(coln));
      if (st) *st = stats;
      return;
    }
    prev_res = res;
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
    // the correction itself (columns left out: zero), added to mo1
    hipLaunchKernelGGL(unscale_columns_kernel, dim3(nblk(blk)), dim3(256), 0, stream_, dx, n, nprop, csc);
    HH_HIP(hipGetLastError());
    hipLaunchKernelGGL(axpy_kernel, dim3(nblk(blk)), dim3(256), 0, stream_, mo1, dx, 1.0, blk);
    HH_HIP(hipGetLastError());
    // The raw response of mo1 (kept for finalize) is recomputed in FP64 after every sweep, so the
    // residual the outer loop tests is the exact FP64 residual of the returned solution.
    response_mo_stacked(mo1, nprop, fin_r, PREC_F64);
    ++stats.krylov.aop_calls;
    copy_bytes(ax, fin_r, (int64_t)sizeof(double) * blk, stream_);
    dimless_scale(mo1, nprop, ax);
    hipLaunchKernelGGL(residual_norm_kernel, dim3((unsigned)nprop), dim3(256), 0, stream_, rhs, mo1, ax, r, n, coln);
    HH_HIP(hipGetLastError());
    res = max_colnorm2(coln);
    stats.refinements = sweep + 1;
    if (verbose) {
      fprintf(stderr, "cphf refinement sweep %d: inner cycles %d (tol %.1e), FP64 residual %.3e\n", sweep + 1,
              ks.cycles, tol_in, std::sqrt(res));
      fflush(stderr);
    }
  }
  stats.residual = std::sqrt(res);
  if (st) *st = stats;
  const bool ok = std::sqrt(res) < cfg_.tol;
  if (stats.refinements > 0 && ok) {
    copy_bytes(fin_x_.reserve(sizeof(double) * blk), mo1, (int64_t)sizeof(double) * blk, stream_);
    fin_nprop_ = nprop;
    fin_state_ = response_state_epoch().load();
  }
  if (!ok) {
    char buf[200];
    snprintf(buf, sizeof buf, "cphf mixed-precision refinement did not reach tol %.3e (residual %.3e after %d sweeps)",
             cfg_.tol, std::sqrt(res), stats.refinements);
    throw Error(HH_ERR_NOT_CONVERGED, buf);
  }
}

void Cphf::finalize(const double* const* f1mo, const double* const* s1mo, const double* const* mo1, int64_t nprop,
                    double* const* mo1_out, double* const* mo_e1) {
  HH_REQUIRE(f1mo && s1mo && mo1 && mo1_out && mo_e1, "cphf: null spin array");
  // last response in FP64 (per spin, into one scratch block)
  const double* x = stack_in(mo1, nprop, st_in_);
  double* b1 = static_cast<double*>(tmp_.reserve(sizeof(double) * n_ * nprop));
  bool reuse = false;
  if (fin_nprop_ == nprop && fin_state_ == response_state_epoch().load()) {
    int* flag = static_cast<int*>(fin_flag_.reserve(64));
    hipLaunchKernelGGL(equal_kernel, dim3(1), dim3(1024), 0, stream_, x, fin_x_.as<double>(), n_ * nprop, flag);
    HH_HIP(hipGetLastError());
    int h = 0;
    HH_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, stream_));
    HH_HIP(hipStreamSynchronize(stream_));
    reuse = h == 1;
  }
  if (reuse) {
    copy_bytes(b1, fin_r_.ptr, (int64_t)sizeof(double) * n_ * nprop, stream_);
    ++n_fin_reuse_;
  } else {
    response_mo_stacked(x, nprop, b1, PREC_F64);
  }
  double* b1s[2] = {b1, nullptr};
  if (nspin_ == 2) {
    double* aux = static_cast<double*>(st_aux_.reserve(sizeof(double) * n_ * nprop));
    b1s[0] = aux;
    b1s[1] = aux + sp_[0].nmo * sp_[0].nocc * nprop;
    unstack(b1, nprop, b1s);
  }
  for (int t = 0; t < nspin_; ++t) {
    const Spin& sp = sp_[t];
    const int64_t blk = sp.nmo * sp.nocc * nprop;
    hipLaunchKernelGGL(finalize_kernel, dim3(nblk(blk)), dim3(256), 0, stream_, f1mo[t], s1mo[t], mo1[t], b1s[t],
                       sp.eocc.as<double>(), sp.evir.as<double>(), sp.nmo, sp.nocc, nprop, mo1_out[t], mo_e1[t]);
  }
  HH_HIP(hipGetLastError());
}

void Cphf::precond_features(const double* pair, const double* r, int64_t nprop, float* x) {
  HH_REQUIRE(nspin_ == 1, "cphf: the learned preconditioner is for one spin");
  const Spin& sp = sp_[0];
  const int64_t ns = (sp.nmo - sp.nocc) * sp.nocc;
  hipLaunchKernelGGL(raw_features_kernel, dim3(nblk(ns)), dim3(256), 0, stream_, pair, r, sp.eocc.as<double>(),
                     sp.evir.as<double>(), sp.nmo, sp.nocc, nprop, x);
  hipLaunchKernelGGL(standardise_kernel, dim3(3), dim3(256), 0, stream_, x, ns);
  HH_HIP(hipGetLastError());
}

void Cphf::precond_target(const double* r, const double* xs, int64_t nprop, float* t, float* w) {
  HH_REQUIRE(nspin_ == 1, "cphf: the learned preconditioner is for one spin");
  const Spin& sp = sp_[0];
  const int64_t ns = (sp.nmo - sp.nocc) * sp.nocc;
  hipLaunchKernelGGL(lsq_target_kernel, dim3(nblk(ns)), dim3(256), 0, stream_, r, xs, sp.nmo, sp.nocc, nprop, t, w);
  HH_HIP(hipGetLastError());
}

void Cphf::precond_diag(const float* y, double* diag) {
  HH_REQUIRE(nspin_ == 1, "cphf: the learned preconditioner is for one spin");
  const Spin& sp = sp_[0];
  hipLaunchKernelGGL(precond_diag_kernel, dim3(nblk(sp.nmo * sp.nocc)), dim3(256), 0, stream_, y, sp.nmo, sp.nocc,
                     diag);
  HH_HIP(hipGetLastError());
}

void Cphf::cphf_hess(const double* const* f1mo, const double* const* s1mo, const double* const* mo1,
                     const double* const* mo_e1, int64_t natm, double* de) {
  HH_REQUIRE(f1mo && s1mo && mo1 && mo_e1, "cphf: null spin array");
  const int64_t n3 = 3 * natm;
  double* G = static_cast<double*>(g1_.reserve(sizeof(double) * n3 * n3));
  // G[(t,A),(s,B)] = sum_spin f (2 f1^T mo1 - 2 (s1 e)^T mo1 - s1oo^T mo_e1) with f = 2 for one
  // (doubly occupied) spin, 1 per spin for two (rscf.rs:382-388, uscf.rs:425-431)
  const double f = nspin_ == 1 ? 2.0 : 1.0;
  int64_t nn_max = 0, no2_max = 0;  // scratch sized once: no reallocation under queued work
  for (int t = 0; t < nspin_; ++t) {
    nn_max = std::max(nn_max, sp_[t].nmo * sp_[t].nocc);
    no2_max = std::max(no2_max, sp_[t].nocc * sp_[t].nocc);
  }
  double* se = static_cast<double*>(sx_.reserve(sizeof(double) * nn_max * n3));
  double* soo = static_cast<double*>(g2_.reserve(sizeof(double) * no2_max * n3));
  for (int t = 0; t < nspin_; ++t) {
    const Spin& sp = sp_[t];
    const int64_t nn = sp.nmo * sp.nocc, no2 = sp.nocc * sp.nocc;
    hipLaunchKernelGGL(scale_by_eocc_kernel, dim3(nblk(nn * n3)), dim3(256), 0, stream_, s1mo[t],
                       sp.eocc.as<double>(), sp.nmo, sp.nocc, n3, se);
    hipLaunchKernelGGL(occ_rows_kernel, dim3(nblk(no2 * n3)), dim3(256), 0, stream_, s1mo[t], sp.nmo, sp.nocc, n3,
                       soo);
    HH_HIP(hipGetLastError());
    GemmArgs g;
    g.M = n3; g.N = n3; g.K = nn;
    g.A = f1mo[t]; g.lda = nn; g.a_kmajor = true;
    g.B = mo1[t]; g.ldb = nn; g.b_kmajor = true;
    g.C = G; g.ldc = n3;
    g.alpha = 2.0 * f; g.beta = t == 0 ? 0.0 : 1.0;
    gemm<double, double>(g, PREC_F64, stream_, ws_);
    g.A = se; g.alpha = -2.0 * f; g.beta = 1.0;
    gemm<double, double>(g, PREC_F64, stream_, ws_);
    GemmArgs h;
    h.M = n3; h.N = n3; h.K = no2;
    h.A = soo; h.lda = no2; h.a_kmajor = true;
    h.B = mo_e1[t]; h.ldb = no2; h.b_kmajor = true;
    h.C = G; h.ldc = n3;
    h.alpha = -f; h.beta = 1.0;
    gemm<double, double>(h, PREC_F64, stream_, ws_);
  }
  hipLaunchKernelGGL(assemble_hess_kernel, dim3(nblk(9 * natm * natm)), dim3(256), 0, stream_, G, natm, de);
  HH_HIP(hipGetLastError());
}

}  // namespace hh
