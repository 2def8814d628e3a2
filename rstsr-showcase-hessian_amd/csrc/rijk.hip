// RI-JK response on gfx950.
//
// Algebra (reference: hess_r.rs:128-183, per perturbation a, per aux index P):
//   D1_a   = bra_a C_o^T + C_o bra_a^T
//   J_ao   = 2 sum_P B_P tr(B_P D1_a)                       (two GEMMs over packed cderi)
//   K_a    = sum_P B_P D1_a B_P C_o
//          = sum_P B_P [ bra_a O_P + C_o (bra_a^T X_P) ],   X_P = B_P C_o,  O_P = C_o^T X_P
// and the reference's k_bra is -K_a.  The reference loops over a and P with small GEMMs; here
// all perturbations are batched into the GEMM N dimension:
//   W[(k,a),(P,i)] = sum_mu bra[mu,(k,a)] X[mu,(P,i)]                 (one GEMM, K = nao)
//   T[(mu,P),(i,a)] = sum_j bra[mu,j,a] O[j,(P,i)] + sum_k C_o[mu,k] W[(k,a),(P,i)]
//                                                      (one GEMM per a, K = 2 nocc concatenated)
//   K[u,(i,a)]     = sum_{(mu,P)} B[u,(mu,P)] T[(mu,P),(i,a)]         (K = nao naux, the hot GEMM)
// X[mu,(P,i)] = (B_P C_o)[mu,i] and O[j,(P,i)] = (C_o^T B_P C_o)[j,i] depend only on the MO
// coefficients and are built once per prepare_response.  The hot GEMM reads B and T with the
// K index k = mu + kpad P (mu padded to kpad, a multiple of 32): FP64 / FP32 as rows [u][P][mu]
// and [(i,a)][P][mu]; BF16x3 as bf16 hi / lo pairs in K blocks of 32, [k/32][u][k%32] and
// [k/32][(i,a)][k%32], so every K step of the hot GEMM and every row group of the T generation
// touches contiguous memory.  Perturbations are processed in chunks so T stays a bounded buffer.
// Flop count equals the reference's: 2 nao^2 naux nocc nprop + 6 nao naux nocc^2 nprop (+ X/O).
#include "rijk.hpp"

#include "elementwise.hpp"
#include "linalg.hpp"
#include "rijk_kfused.hpp"
#include "rijk_tgen.hpp"

namespace hh {

Rijk::Rijk(const double* cderi, int64_t nao, int64_t naux, double scale_j, double scale_k, int nspin, hipStream_t s)
    : nao_(nao), naux_(naux), nao_tp_(nao * (nao + 1) / 2), scale_j_(scale_j), scale_k_(scale_k), nspin_(nspin),
      stream_(s) {
  HH_REQUIRE(nao > 0 && naux > 0, "nao and naux must be positive");
  HH_REQUIRE(nspin == 1 || nspin == 2, "nspin must be 1 (RHF) or 2 (UHF)");
  HH_REQUIRE(cderi != nullptr, "cderi is null");
  cderi_.reserve(sizeof(double) * nao_tp_ * naux_);
  HH_HIP(hipMemcpyAsync(cderi_.ptr, cderi, sizeof(double) * nao_tp_ * naux_, hipMemcpyDeviceToDevice, stream_));
  bfull_.reserve(sizeof(double) * nao_ * nao_ * naux_);
  unpack_triu<double, double>(cderi_.as<double>(), nao_, nao_tp_, naux_, bfull_.as<double>(), stream_);
}

namespace {
__global__ void diag_min_kernel(const double* L, int64_t n, double* out) {
  __shared__ double red[256];
  double m = 1e300;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) m = fmin(m, L[i + i * n]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// V[:, j] *= f(e_j): mode 0 (eigenvalue floor of decomp_j2c_cd) f = max(e, thr);
// mode 1 (decomp_j2c_eig) f = e^-1/2 for e >= thr, 0 for the discarded eigenvalues.
__global__ void eig_scale_cols_kernel(const double* V, const double* e, int64_t n, double thr, int mode, double* W) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const double ej = e[t / n];
  const double f = mode == 0 ? fmax(ej, thr) : (ej >= thr ? 1.0 / sqrt(ej) : 0.0);
  W[t] = V[t] * f;
}
}  // namespace

namespace {
// A = W V^T  (n x n x n FP64 product on the matrix cores)
void outer_wvt(const double* W, const double* V, int64_t n, double* A, hipStream_t s, DevBuf& ws) {
  GemmArgs g;
  g.M = n, g.N = n, g.K = n;
  g.A = W, g.lda = n, g.a_kmajor = false;  // W[m, k] at m + k n
  g.B = V, g.ldb = n, g.b_kmajor = false;  // V^T[k, j] = V[j, k] at j + k n
  g.C = A, g.ldc = n;
  gemm<double, double>(g, PREC_F64, s, ws);
}
}  // namespace

std::unique_ptr<Rijk> Rijk::from_j3c(const double* j3c, const double* j2c, int64_t nao, int64_t naux, int policy,
                                     double threshold, double scale_j, double scale_k, int nspin, hipStream_t s) {
  HH_REQUIRE(j3c && j2c && nao > 0 && naux > 0, "from_j3c: null or empty input");
  HH_REQUIRE(policy == HH_J2C_CD || policy == HH_J2C_EIG, "from_j3c: policy must be HH_J2C_CD or HH_J2C_EIG");
  const bool has_thr = threshold > 0.0;  // Some(threshold); <= 0 or NaN: None
  const int64_t ntp = nao * (nao + 1) / 2;
  const int64_t nn = naux * naux;
  const dim3 grid((unsigned)((nn + 255) / 256)), block(256);
  DevBuf F, cd, ws, aux;
  double* Fp = static_cast<double*>(F.reserve(sizeof(double) * nn));
  double* cdp = static_cast<double*>(cd.reserve(sizeof(double) * ntp * naux));
  if (policy == HH_J2C_EIG) {
    const double thr = has_thr ? threshold : 1e-13;  // J2C_THRESH (decompose.rs:5)
    double* V = static_cast<double*>(aux.reserve(sizeof(double) * (2 * nn + naux)));
    double* W = V + nn;
    double* e = W + nn;
    HH_HIP(hipMemcpyAsync(V, j2c, sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
    eigh(V, naux, e, s, ws);
    hipLaunchKernelGGL(eig_scale_cols_kernel, grid, block, 0, s, V, e, naux, thr, 1, W);
    HH_HIP(hipGetLastError());
    outer_wvt(W, V, naux, Fp, s, ws);  // J^-1/2
    // cderi = j3c J^-1/2  [ntp, naux]
    GemmArgs g;
    g.M = ntp, g.N = naux, g.K = naux;
    g.A = j3c, g.lda = ntp, g.a_kmajor = false;
    g.B = Fp, g.ldb = naux, g.b_kmajor = true;
    g.C = cdp, g.ldc = ntp;
    gemm<double, double>(g, PREC_F64, s, ws);
  } else {
    HH_HIP(hipMemcpyAsync(Fp, j2c, sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
    bool ok = cholesky(Fp, naux, s, ws);
    double hmin = 0.0;
    if (ok) {
      double* dm = static_cast<double*>(aux.reserve(sizeof(double)));
      hipLaunchKernelGGL(diag_min_kernel, dim3(1), dim3(256), 0, s, Fp, naux, dm);
      HH_HIP(hipMemcpyAsync(&hmin, dm, sizeof(double), hipMemcpyDeviceToHost, s));
      HH_HIP(hipStreamSynchronize(s));
    }
    if (!ok && !has_thr)
      throw Error(HH_ERR_INVALID, "Cholesky decomposition failed for j2c matrix, and no threshold is provided for "
                                  "fallback");
    if (has_thr && !(ok && hmin > threshold)) {
      // eigenvalue floor: j2c <- V max(e, threshold) V^T, then Cholesky again
      double* V = static_cast<double*>(aux.reserve(sizeof(double) * (2 * nn + naux)));
      double* W = V + nn;
      double* e = W + nn;
      HH_HIP(hipMemcpyAsync(V, j2c, sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
      eigh(V, naux, e, s, ws);
      hipLaunchKernelGGL(eig_scale_cols_kernel, grid, block, 0, s, V, e, naux, threshold, 0, W);
      HH_HIP(hipGetLastError());
      outer_wvt(W, V, naux, Fp, s, ws);
      if (!cholesky(Fp, naux, s, ws))
        throw Error(HH_ERR_INVALID, "j2c with its eigenvalues floored at the threshold is not positive definite");
    }
    HH_HIP(hipMemcpyAsync(cdp, j3c, sizeof(double) * ntp * naux, hipMemcpyDeviceToDevice, s));
    trsm_lower(Fp, naux, cdp, ntp, ntp, true, s, ws);
  }
  auto r = std::make_unique<Rijk>(cdp, nao, naux, scale_j, scale_k, nspin, s);
  HH_HIP(hipStreamSynchronize(s));
  DevBuf& keep = policy == HH_J2C_EIG ? r->j2c_linv_ : r->j2c_l_;
  std::swap(keep.ptr, F.ptr);
  std::swap(keep.bytes, F.bytes);
  return r;
}

Rijk::~Rijk() {
  for (hipEvent_t e : events_) (void)hipEventDestroy(e);
}

void Rijk::set_profiling(bool on) { profiling_ = on; }

void Rijk::reset_timing() {
  for (double& t : times_) t = 0.0;
  hot_launches_ = 0;
}

void Rijk::get_timing(double* out) const {
  for (int i = 0; i < T_COUNT; ++i) out[i] = times_[i];
  out[T_COUNT] = (double)hot_launches_;
}

hipEvent_t Rijk::ev(int i) {
  while ((int)events_.size() <= i) {
    hipEvent_t e;
    HH_HIP(hipEventCreate(&e));
    events_.push_back(e);
  }
  return events_[i];
}

int Rijk::mark() {
  const int e = next_event_++;
  HH_HIP(hipEventRecord(ev(e), stream_));
  return e;
}

void Rijk::span(int key, int begin) { pending_.push_back({key, {begin, mark()}}); }

void Rijk::flush_timing() {
  for (auto& p : pending_) {
    HH_HIP(hipEventSynchronize(events_[p.second.second]));
    float ms = 0.f;
    HH_HIP(hipEventElapsedTime(&ms, events_[p.second.first], events_[p.second.second]));
    times_[p.first] += ms;
    if (p.first == T_K_HOT) ++hot_launches_;
  }
  pending_.clear();
  next_event_ = 0;
}

void Rijk::prepare_response(const double* const* mo_coeff, const double* const* mo_occ, const int64_t* nmo) {
  HH_REQUIRE(mo_coeff && mo_occ && nmo, "null argument");
  for (int s = 0; s < nspin_; ++s) {
    HH_REQUIRE(mo_coeff[s] && mo_occ[s] && nmo[s] > 0, "null mo_coeff / mo_occ or nmo <= 0");
    std::vector<int64_t> occ_idx;
    for (int64_t j = 0; j < nmo[s]; ++j)
      if (mo_occ[s][j] > 0.0) occ_idx.push_back(j);
    const int64_t nocc = (int64_t)occ_idx.size();
    HH_REQUIRE(nocc > 0, "no occupied orbitals");
    nocc_[s] = nocc;
    double* mocc = static_cast<double*>(mocc_[s].reserve(sizeof(double) * nao_ * nocc));
    for (int64_t c = 0; c < nocc; ++c)
      HH_HIP(hipMemcpyAsync(mocc + c * nao_, mo_coeff[s] + occ_idx[c] * nao_, sizeof(double) * nao_,
                            hipMemcpyDeviceToDevice, stream_));
    // X_P = B_P C_o  (batched over P): X[mu + nao*(P + naux*i)]
    double* X = static_cast<double*>(x_[s].reserve(sizeof(double) * nao_ * nocc * naux_));
    {
      GemmArgs g;
      g.M = nao_; g.N = nocc; g.K = nao_;
      g.A = bfull_.ptr; g.lda = nao_; g.a_kmajor = false; g.sA0 = nao_ * nao_;
      g.B = mocc; g.ldb = nao_; g.b_kmajor = true; g.sB0 = 0;
      g.C = X; g.ldc = nao_ * naux_; g.sC0 = nao_;
      g.nb0 = naux_;
      gemm<double, double>(g, PREC_F64, stream_, ws_);
    }
    // O[j, (P,i)] = sum_nu C_o[nu, j] X[nu, (P,i)]
    double* O = static_cast<double*>(o_[s].reserve(sizeof(double) * nocc * nocc * naux_));
    {
      GemmArgs g;
      g.M = nocc; g.N = nocc * naux_; g.K = nao_;
      g.A = mocc; g.lda = nao_; g.a_kmajor = true;
      g.B = X; g.ldb = nao_; g.b_kmajor = true;
      g.C = O; g.ldc = nocc;
      gemm<double, double>(g, PREC_F64, stream_, ws_);
    }
  }
  prepared_ = true;
  mixed_ready_ = false;
  xt_ready_ = false;
  kf_ready_ = false;
  res_ready_[0] = res_ready_[1] = false;
  bump_response_state();
}

namespace {
// pair[m, i] = sum_P V[m, (P, i)]^2, V [nmo, naux * nocc]
__global__ void pair_sumsq_kernel(const double* V, int64_t nmo, int64_t naux, int64_t nocc, double* pair) {
  const int64_t total = nmo * nocc;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = e % nmo, i = e / nmo;
    double s = 0.0;
    for (int64_t P = 0; P < naux; ++P) {
      const double v = V[m + nmo * (P + naux * i)];
      s += v * v;
    }
    pair[e] = s;
  }
}
}  // namespace

void Rijk::pair_norms(int s, const double* mo_coeff, int64_t nmo, double* pair) {
  if (!prepared_) throw Error(HH_ERR_STATE, "pair_norms requested before prepare_response");
  HH_REQUIRE(s >= 0 && s < nspin_ && mo_coeff && pair && nmo > 0, "pair_norms: bad argument");
  const int64_t nocc = nocc_[s];
  DevBuf vb;
  double* V = static_cast<double*>(vb.reserve(sizeof(double) * nmo * naux_ * nocc));
  GemmArgs g;  // V[m, (P,i)] = sum_mu C[mu, m] X[mu, (P,i)]
  g.M = nmo; g.N = naux_ * nocc; g.K = nao_;
  g.A = mo_coeff; g.lda = nao_; g.a_kmajor = true;
  g.B = x_[s].ptr; g.ldb = nao_; g.b_kmajor = true;
  g.C = V; g.ldc = nmo;
  gemm<double, double>(g, PREC_F64, stream_, ws_);
  const int64_t total = nmo * nocc;
  hipLaunchKernelGGL(pair_sumsq_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(total, 256), 65535)), dim3(256), 0,
                     stream_, V, nmo, naux_, nocc, pair);
  HH_HIP(hipGetLastError());
  HH_HIP(hipStreamSynchronize(stream_));
}

void Rijk::ensure_mixed() {
  if (mixed_ready_) return;
  if (!cderi32_.ptr) {
    cderi32_.reserve(sizeof(float) * nao_tp_ * naux_);
    cast<double, float>(cderi_.as<double>(), nao_tp_ * naux_, cderi32_.as<float>(), stream_);
    bfull32_.reserve(sizeof(float) * nao_ * nao_ * naux_);
    unpack_triu<double, float>(cderi_.as<double>(), nao_, nao_tp_, naux_, bfull32_.as<float>(), stream_);
  }
  for (int s = 0; s < nspin_; ++s) {
    const int64_t nocc = nocc_[s];
    mocc32_[s].reserve(sizeof(float) * nao_ * nocc);
    cast<double, float>(mocc_[s].as<double>(), nao_ * nocc, mocc32_[s].as<float>(), stream_);
    x32_[s].reserve(sizeof(float) * nao_ * nocc * naux_);
    cast<double, float>(x_[s].as<double>(), nao_ * nocc * naux_, x32_[s].as<float>(), stream_);
    o32_[s].reserve(sizeof(float) * nocc * nocc * naux_);
    cast<double, float>(o_[s].as<double>(), nocc * nocc * naux_, o32_[s].as<float>(), stream_);
  }
  mixed_ready_ = true;
}

void Rijk::build_j_ao(const double* const* bra, int64_t nprop, Prec prec, double* j_ao) {
  // dm1[mu, nu, a] = sum_s sum_i bra_s[mu, i, a] C_s[nu, i]
  double* dm = static_cast<double*>(dm_.reserve(sizeof(double) * nao_ * nao_ * nprop));
  for (int s = 0; s < nspin_; ++s) {
    GemmArgs g;
    g.M = nao_; g.N = nao_; g.K = nocc_[s];
    g.A = bra[s]; g.lda = nao_; g.a_kmajor = false; g.sA0 = nao_ * nocc_[s];
    g.B = mocc_[s].ptr; g.ldb = nao_; g.b_kmajor = false;  // B[i, nu] = C[nu + nao*i]
    g.C = dm; g.ldc = nao_; g.sC0 = nao_ * nao_;
    g.nb0 = nprop;
    g.beta = (s == 0) ? 0.0 : 1.0;
    gemm<double, double>(g, PREC_F64, stream_, ws_);
  }
  // dm_tp = pack_triu_tilde(dm1 + dm1^T)
  double* dmtp = static_cast<double*>(dmtp_.reserve(sizeof(double) * nao_tp_ * nprop));
  pack_triu_tilde(dm, nao_, nprop, true, dmtp, stream_);
  double* itm = static_cast<double*>(itm_.reserve(sizeof(double) * naux_ * nprop));
  double* jtp = static_cast<double*>(jtp_.reserve(sizeof(double) * nao_tp_ * nprop));
  if (prec == PREC_F64) {
    GemmArgs g;  // itm[P, a] = sum_tp cderi[tp, P] dm_tp[tp, a]
    g.M = naux_; g.N = nprop; g.K = nao_tp_;
    g.A = cderi_.ptr; g.lda = nao_tp_; g.a_kmajor = true;
    g.B = dmtp; g.ldb = nao_tp_; g.b_kmajor = true;
    g.C = itm; g.ldc = naux_;
    gemm<double, double>(g, PREC_F64, stream_, ws_);
    GemmArgs h;  // resp_tp[tp, a] = 2 sum_P cderi[tp, P] itm[P, a]
    h.M = nao_tp_; h.N = nprop; h.K = naux_;
    h.A = cderi_.ptr; h.lda = nao_tp_; h.a_kmajor = false;
    h.B = itm; h.ldb = naux_; h.b_kmajor = true;
    h.C = jtp; h.ldc = nao_tp_; h.alpha = 2.0;
    gemm<double, double>(h, PREC_F64, stream_, ws_);
  } else if (prec == PREC_BF16X3) {
    // both products on the NT kernel: itm = cderi^T dm_tp (K = AO pairs), then
    // resp_tp = 2 cderi itm (K = P), split-bf16 operands, FP32 accumulation, FP64 outputs
    const int64_t kp1 = (nao_tp_ + 31) / 32 * 32, kp2 = (naux_ + 31) / 32 * 32;
    if (!j_bf16_ready_) {
      const size_t e1 = sizeof(unsigned short) * (size_t)kp1 * naux_, e2 = sizeof(unsigned short) * (size_t)kp2 * nao_tp_;
      auto* a = static_cast<unsigned short*>(cj_hi_.reserve(e1));
      auto* b = static_cast<unsigned short*>(cj_lo_.reserve(e1));
      split_bf16_tiled<double>(cderi_.as<double>(), nao_tp_, nao_tp_, naux_, kp1, a, b, stream_);
      auto* c = static_cast<unsigned short*>(cjt_hi_.reserve(e2));
      auto* d = static_cast<unsigned short*>(cjt_lo_.reserve(e2));
      split_bf16_tiled_t<double>(cderi_.as<double>(), nao_tp_, nao_tp_, naux_, kp2, c, d, stream_);
      j_bf16_ready_ = true;
    }
    const int64_t e1 = kp1 * nprop, e2 = kp2 * nprop;
    auto* sp = static_cast<unsigned short*>(jsp_.reserve(sizeof(unsigned short) * 2 * (size_t)(e1 + e2)));
    unsigned short *dh = sp, *dl = sp + e1, *ih = sp + 2 * e1, *il = ih + e2;
    split_bf16_tiled<double>(dmtp, nao_tp_, nao_tp_, nprop, kp1, dh, dl, stream_);
    NtArgs g;  // itm[P, a] = sum_tp cderi[tp, P] dm_tp[tp, a]
    g.M = naux_; g.N = nprop; g.K = kp1;
    g.A = cj_hi_.ptr; g.A_lo = cj_lo_.ptr; g.lda = kp1;
    g.B = dh; g.B_lo = dl; g.ldb = kp1;
    g.tiled = true;
    g.C = itm; g.ldc = naux_;
    gemm_nt(g, NT_BF16X3, stream_, ws_);
    split_bf16_tiled<double>(itm, naux_, naux_, nprop, kp2, ih, il, stream_);
    NtArgs h;  // resp_tp[tp, a] = 2 sum_P cderi[tp, P] itm[P, a]
    h.M = nao_tp_; h.N = nprop; h.K = kp2;
    h.A = cjt_hi_.ptr; h.A_lo = cjt_lo_.ptr; h.lda = kp2;
    h.B = ih; h.B_lo = il; h.ldb = kp2;
    h.tiled = true;
    h.C = jtp; h.ldc = nao_tp_; h.alpha = 2.0;
    gemm_nt(h, NT_BF16X3, stream_, ws_);
  } else {
    ensure_mixed();
    float* dmtp32 = static_cast<float*>(bra32_.reserve(sizeof(float) * std::max(nao_tp_, naux_) * nprop));
    cast<double, float>(dmtp, nao_tp_ * nprop, dmtp32, stream_);
    float* itm32 = reinterpret_cast<float*>(itm);  // reuse (naux*nprop floats fit in the f64 buffer)
    GemmArgs g;
    g.M = naux_; g.N = nprop; g.K = nao_tp_;
    g.A = cderi32_.ptr; g.lda = nao_tp_; g.a_kmajor = true;
    g.B = dmtp32; g.ldb = nao_tp_; g.b_kmajor = true;
    g.C = itm32; g.ldc = naux_;
    gemm<float, float>(g, prec, stream_, ws_);
    GemmArgs h;
    h.M = nao_tp_; h.N = nprop; h.K = naux_;
    h.A = cderi32_.ptr; h.lda = nao_tp_; h.a_kmajor = false;
    h.B = itm32; h.ldb = naux_; h.b_kmajor = true;
    h.C = jtp; h.ldc = nao_tp_; h.alpha = 2.0;
    gemm<float, double>(h, prec, stream_, ws_);
  }
  unpack_triu<double, double>(jtp, nao_, nao_tp_, nprop, j_ao, stream_);
}

namespace {
// Bp[u][P][mu] = B_P[u, mu] for mu < nao, 0 for nao <= mu < kpad  (from packed cderi)
template <typename T>
__global__ void pack_hot_a_kernel(const double* __restrict__ cderi, int64_t nao, int64_t naux, int64_t kpad,
                                  T* __restrict__ out) {
  const int64_t total = nao * naux * kpad;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t mu = e % kpad, rest = e / kpad;
    const int64_t P = rest % naux, u = rest / naux;
    double v = 0.0;
    if (mu < nao) {
      const int64_t lo = mu < u ? mu : u, hi = mu < u ? u : mu;
      v = cderi[hi * (hi + 1) / 2 + lo + P * (nao * (nao + 1) / 2)];
    }
    out[e] = (T)v;
  }
}

// BF16x3 hot operand in K blocks of 32 (k = mu + kpad P, kbp = kpad / 32):
// hi/lo[((P kbp + mu / 32) nao + u) 32 + mu % 32] = split(float(B_P[u, mu])), 0 for mu >= nao
__global__ void pack_hot_a_bf16_kernel(const double* __restrict__ cderi, int64_t nao, int64_t naux, int64_t kbp,
                                       unsigned short* __restrict__ hi, unsigned short* __restrict__ lo) {
  const int64_t total = naux * kbp * nao * 32;
  const int64_t ntp = nao * (nao + 1) / 2;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t kl = e % 32, u = (e / 32) % nao, kb = e / (32 * nao);
    const int64_t P = kb / kbp, mu = (kb % kbp) * 32 + kl;
    float x = 0.f;
    if (mu < nao) {
      const int64_t l = mu < u ? mu : u, h = mu < u ? u : mu;
      x = (float)cderi[h * (h + 1) / 2 + l + P * ntp];
    }
    const __bf16 xh = (__bf16)x;
    hi[e] = __builtin_bit_cast(unsigned short, xh);
    lo[e] = __builtin_bit_cast(unsigned short, (__bf16)(x - (float)xh));
  }
}

// zero rows [nao, kpad) of every (n, Pl) segment of T[n][Pl][kpad]
__global__ void zero_pad_rows_kernel(double* T, int64_t nseg, int64_t nao, int64_t kpad) {
  const int64_t w = kpad - nao;
  const int64_t total = nseg * w;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x)
    T[(e / w) * kpad + nao + e % w] = 0.0;
}

unsigned ew_blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 65536)); }
}  // namespace

void Rijk::ensure_packed(Prec prec) {
  kpad_ = ceil_div(nao_, 32) * 32;
  const int64_t n = nao_ * naux_ * kpad_;
  const bool bf16 = (prec == PREC_BF16X3);
  if (prec == PREC_F32 && !bp32_.ptr) {
    bp32_.reserve(sizeof(float) * n);
    hipLaunchKernelGGL((pack_hot_a_kernel<float>), dim3(ew_blocks(n)), dim3(256), 0, stream_, cderi_.as<double>(),
                       nao_, naux_, kpad_, bp32_.as<float>());
    HH_HIP(hipGetLastError());
  }
  if (prec == PREC_F64 && !bp64_.ptr) {
    bp64_.reserve(sizeof(double) * n);
    hipLaunchKernelGGL((pack_hot_a_kernel<double>), dim3(ew_blocks(n)), dim3(256), 0, stream_, cderi_.as<double>(),
                       nao_, naux_, kpad_, bp64_.as<double>());
    HH_HIP(hipGetLastError());
  }
  if (bf16 && !bp_hi_.ptr) {
    bp_hi_.reserve(sizeof(unsigned short) * n);
    bp_lo_.reserve(sizeof(unsigned short) * n);
    hipLaunchKernelGGL(pack_hot_a_bf16_kernel, dim3(ew_blocks(n)), dim3(256), 0, stream_, cderi_.as<double>(), nao_,
                       naux_, kpad_ / 32, bp_hi_.as<unsigned short>(), bp_lo_.as<unsigned short>());
    HH_HIP(hipGetLastError());
  }
}

void Rijk::build_k(int s, const double* bra, int64_t nprop, Prec prec, double alpha, double beta, double* out) {
  const int64_t nocc = nocc_[s];
  const int64_t nP = naux_ * nocc;  // (P, i) columns
  if (prec == PREC_BF16X3 && !tgen_supported(nocc) && !(k_path != 2 && kfused_supported(nao_, nocc))) {
    // neither BF16x3 K algorithm takes this many occupied orbitals (the materialised T tile holds
    // 2 nocc <= 256 rows): the K product runs in FP32 (J stays BF16x3)
    build_k(s, bra, nprop, PREC_F32, alpha, beta, out);
    return;
  }
  const bool f64 = (prec == PREC_F64), bf = (prec == PREC_BF16X3);
  ensure_packed(prec);
  if (!f64) ensure_mixed();
  const size_t esz = f64 ? sizeof(double) : sizeof(float);
  const void* braX = bra;
  if (!f64) {
    float* b32 = static_cast<float*>(bra32_.reserve(sizeof(float) * std::max(nao_ * nocc * nprop, std::max(nao_tp_, naux_) * nprop)));
    cast<double, float>(bra, nao_ * nocc * nprop, b32, stream_);
    braX = b32;
  }
  const void* X = f64 ? x_[s].ptr : x32_[s].ptr;
  const void* O = f64 ? o_[s].ptr : o32_[s].ptr;
  const void* Co = f64 ? mocc_[s].ptr : mocc32_[s].ptr;
  // W[(k + nocc*a) + nocc*nprop*(P + naux*i)] = sum_mu bra[mu,(k,a)] X[mu,(P,i)]
  const int64_t ldw = nocc * nprop;
  void* W = nullptr;
  if (bf) {
    // NT GEMM on K-block tiled bf16 pairs (K = mu padded to kpad), FP32 output
    if (!xt_ready_) {
      for (int t = 0; t < nspin_; ++t) {
        const int64_t nc = naux_ * nocc_[t];
        xt_hi_[t].reserve(sizeof(unsigned short) * kpad_ * nc);
        xt_lo_[t].reserve(sizeof(unsigned short) * kpad_ * nc);
        split_bf16_tiled<float>(x32_[t].as<float>(), nao_, nao_, nc, kpad_, xt_hi_[t].as<unsigned short>(),
                                xt_lo_[t].as<unsigned short>(), stream_);
      }
      xt_ready_ = true;
    }
    unsigned short* bth = static_cast<unsigned short*>(brat_.reserve(2 * sizeof(unsigned short) * kpad_ * ldw));
    unsigned short* btl = bth + kpad_ * ldw;
    split_bf16_tiled<double>(bra, nao_, nao_, ldw, kpad_, bth, btl, stream_);
    NtArgs a;
    a.M = ldw; a.N = nP; a.K = kpad_;
    a.A = bth; a.A_lo = btl; a.B = xt_hi_[s].ptr; a.B_lo = xt_lo_[s].ptr;
    a.lda = a.ldb = kpad_;
    a.tiled = true;
    if (k_path != 2 && kfused_supported(nao_, nocc)) {
      // fused path: the product's epilogue writes W directly as the term-2 operand W'
      const int64_t np = kfused_naux_pad(naux_), wn = np * nocc * ldw;
      unsigned short* wp = static_cast<unsigned short*>(kw_.reserve(2 * sizeof(unsigned short) * wn));
      kfused_zero_w_pad(naux_, nocc, nprop, wp, wp + wn, stream_);
      a.W_hi = wp; a.W_lo = wp + wn;
      a.w_nocc = nocc; a.w_naux = naux_; a.w_naux_pad = np;
      gemm_nt(a, NT_BF16X3, stream_, ws_);
      build_k_fused(s, bra, nprop, wp, alpha, beta, out);
      return;
    }
    W = w_.reserve(esz * ldw * nP);
    a.C_f32 = static_cast<float*>(W); a.ldc = ldw;
    gemm_nt(a, NT_BF16X3, stream_, ws_);
  } else {
    W = w_.reserve(esz * ldw * nP);
    GemmArgs g;
    g.M = ldw; g.N = nP; g.K = nao_;
    g.A = braX; g.lda = nao_; g.a_kmajor = true;
    g.B = X; g.ldb = nao_; g.b_kmajor = true;
    g.C = W; g.ldc = ldw;
    if (f64) gemm<double, double>(g, prec, stream_, ws_);
    else gemm<float, float>(g, prec, stream_, ws_);
  }
  // perturbation chunk: the caller's batch, else equal chunks with T[(i,a)][P][kpad] within ~16 GiB
  // (bf16 hi + lo, FP32 and FP64 elements: 4, 4 and 8 bytes), so every hot launch has one shape
  const int64_t per_a = nocc * naux_ * kpad_;
  int64_t ac = k_chunk;
  if (ac <= 0) {
    ac = std::min(nprop, std::max<int64_t>(1, (int64_t)((16LL << 30) / ((f64 ? 8 : 4) * per_a))));
    const int64_t nchunk = (nprop + ac - 1) / ac;
    ac = (nprop + nchunk - 1) / nchunk;
  }
  ac = std::min(ac, nprop);
  // FP32 / FP64: rows nao..kpad of every (n, P) segment stay zero (the generation GEMM writes
  // mu < nao only), so the buffer is cleared once per layout.  BF16x3: the generation kernel
  // writes every row of the K-block tiled layout, padding included.
  void* T = nullptr;
  unsigned short* th = nullptr;
  unsigned short* tl = nullptr;
  const int64_t tlen = per_a * ac;
  if (bf) {
    th = static_cast<unsigned short*>(tp_hi_.reserve(sizeof(unsigned short) * tlen));
    tl = static_cast<unsigned short*>(tp_lo_.reserve(sizeof(unsigned short) * tlen));
  } else {
    T = tp_.reserve(esz * tlen);
    if (tp_key_[0] != (int64_t)esz || tp_key_[1] != tlen) {
      if (f64) fill<double>(static_cast<double*>(T), tlen, 0.0, stream_);
      else fill<float>(static_cast<float*>(T), tlen, 0.f, stream_);
      tp_key_[0] = (int64_t)esz;
      tp_key_[1] = tlen;
    }
  }
  for (int64_t a0 = 0; a0 < nprop; a0 += ac) {
    const int64_t na = std::min(ac, nprop - a0);
    if (bf) {
      // T[k = mu + kpad P][n = i + nocc al] = bra_a O + C_o W_a as the bf16 pair, K-block tiled
      tgen_bf16x3(static_cast<const float*>(braX), static_cast<const float*>(Co), static_cast<const float*>(O),
                  static_cast<const float*>(W), nao_, nocc, nprop, nP, a0, na, kpad_, th, tl, gp_, stream_);
    } else {
      // T[mu + kpad*(P + naux*(i + nocc*al))] = bra_a O + C_o W_a   (one GEMM per a, K = 2 nocc)
      GemmArgs g;
      g.M = nao_; g.N = nP; g.K = 2 * nocc; g.K1 = nocc;
      g.A = static_cast<const char*>(braX) + esz * (nao_ * nocc * a0); g.lda = nao_; g.a_kmajor = false;
      g.sA0 = nao_ * nocc;
      g.B = O; g.ldb = nocc; g.b_kmajor = true; g.sB0 = 0;
      g.A2 = Co; g.lda2 = nao_; g.sA2_0 = 0;
      g.B2 = static_cast<const char*>(W) + esz * (nocc * a0); g.ldb2 = ldw; g.sB2_0 = nocc;
      g.ldc = kpad_; g.sC0 = per_a;
      g.nb0 = na;
      g.ksplit = 1;
      g.C = T;
      if (f64) gemm<double, double>(g, prec, stream_, ws_);
      else gemm<float, float>(g, prec, stream_, ws_);
    }
    double* outc = out + nao_ * nocc * a0;
    const double bet = beta;
    hipEvent_t e_begin = nullptr, e_end = nullptr;
    if (profiling_) {
      const int e0 = next_event_++, e1 = next_event_++;
      e_begin = ev(e0);
      e_end = ev(e1);
      pending_.push_back({T_K_HOT, {e0, e1}});
    }
    if (prec == PREC_F32) {
      GemmArgs g;
      g.M = nao_; g.N = nocc * na; g.K = naux_ * kpad_;
      g.A = bp32_.ptr; g.lda = naux_ * kpad_; g.a_kmajor = true;
      g.B = T; g.ldb = naux_ * kpad_; g.b_kmajor = true;
      g.C = outc; g.ldc = nao_;
      g.alpha = alpha; g.beta = bet;
      g.ev_begin = e_begin; g.ev_main_end = e_end;
      gemm<float, double>(g, PREC_F32, stream_, ws_);
      continue;
    }
    NtArgs a;
    a.M = nao_; a.N = nocc * na; a.K = naux_ * kpad_;
    a.lda = naux_ * kpad_; a.ldb = naux_ * kpad_;
    a.C = outc; a.ldc = nao_;
    a.alpha = alpha; a.beta = bet;
    a.ev_begin = e_begin; a.ev_main_end = e_end;
    if (bf) {
      a.A = bp_hi_.as<unsigned short>(); a.A_lo = bp_lo_.as<unsigned short>();
      a.B = th; a.B_lo = tl;
      a.tiled = true;
      gemm_nt(a, NT_BF16X3, stream_, ws_);
    } else {
      a.A = bp64_.as<double>();
      a.B = T;
      gemm_nt(a, NT_F64, stream_, ws_);
    }
  }
}

// BF16x3 K with the fused first term (rijk_kfused.hpp): term 1 = sum_P B_P bra_a O_P with its
// operand generated inside the contraction, term 2 = sum_P X_P W_aP as one NT product over
// K = (P, j) with W permuted to rows (i, a).
void Rijk::build_k_fused(int s, const double* bra, int64_t nprop, const unsigned short* wp, double alpha, double beta,
                         double* out) {
  const int64_t nocc = nocc_[s];
  const int64_t nb = kfused_nb(nao_), jp = kfused_jp(nocc), np = kfused_naux_pad(naux_);
  if (!kb_hi_.ptr) {  // per handle
    const int64_t n = kfused_b_len(nao_, naux_);
    kb_hi_.reserve(sizeof(unsigned short) * n);
    kb_lo_.reserve(sizeof(unsigned short) * n);
    kfused_pack_b(cderi_.as<double>(), nao_, naux_, kb_hi_.as<unsigned short>(), kb_lo_.as<unsigned short>(),
                  stream_);
  }
  if (!kf_ready_) {  // per prepare_response
    for (int t = 0; t < nspin_; ++t) {
      const int64_t no = nocc_[t];
      if (!kfused_supported(nao_, no)) continue;
      const int64_t on = kfused_o_len(no, naux_), xn = np * no * nao_;
      ko_[t].reserve(2 * sizeof(unsigned short) * on);
      kx_[t].reserve(2 * sizeof(unsigned short) * xn);
      unsigned short* o = ko_[t].as<unsigned short>();
      unsigned short* x = kx_[t].as<unsigned short>();
      kfused_pack_o(o_[t].as<double>(), no, naux_, o, o + on, stream_);
      kfused_pack_x(x32_[t].as<float>(), nao_, naux_, no, x, x + xn, stream_);
    }
    kf_ready_ = true;
  }
  const int64_t rn = nprop * nb * 32 * jp;
  unsigned short* r = static_cast<unsigned short*>(kr_.reserve(2 * sizeof(unsigned short) * rn));
  kfused_pack_bra(bra, nao_, nocc, nprop, r, r + rn, stream_);
  const int64_t wn = np * nocc * nocc * nprop;
  KfusedArgs k;
  const int64_t on = kfused_o_len(nocc, naux_);
  k.bh = kb_hi_.as<unsigned short>(); k.bl = kb_lo_.as<unsigned short>();
  k.oh = ko_[s].as<unsigned short>(); k.ol = k.oh + on;
  k.rh = r; k.rl = r + rn;
  k.nao = nao_; k.naux = naux_; k.nocc = nocc; k.nprop = nprop;
  k.alpha = alpha; k.beta = beta; k.out = out;
  if (profiling_) {
    const int e0 = next_event_++, e1 = next_event_++;
    k.ev_begin = ev(e0);
    k.ev_main_end = ev(e1);
    pending_.push_back({T_K_HOT, {e0, e1}});
  }
  kfused_term1(k, kf_ws_, stream_);
  // term 2: out += alpha sum_{(P, j)} X'[mu, (P, j)] W'[(i, a), (P, j)]
  const int64_t xn = np * nocc * nao_;
  NtArgs a;
  a.M = nao_; a.N = nocc * nprop; a.K = np * nocc;
  a.A = kx_[s].as<unsigned short>(); a.A_lo = kx_[s].as<unsigned short>() + xn;
  a.B = wp; a.B_lo = wp + wn;
  a.lda = a.ldb = np * nocc;
  a.tiled = true;
  a.C = out; a.ldc = nao_;
  a.alpha = alpha; a.beta = 1.0;
  gemm_nt(a, NT_BF16X3, stream_, ws_);
}

void Rijk::response_bra_separated(const double* const* bra, int64_t nprop, double* j_ao, double* const* k_bra,
                                  Prec prec) {
  if (!prepared_) throw Error(HH_ERR_STATE, "response requested before prepare_response");
  HH_REQUIRE(bra != nullptr && nprop > 0, "bra is null or nprop <= 0");
  for (int s = 0; s < nspin_; ++s) HH_REQUIRE(bra[s] != nullptr, "bra[s] is null");
  const int r0 = profiling_ ? mark() : -1;
  if (j_ao) {
    const int a = profiling_ ? mark() : -1;
    build_j_ao(bra, nprop, prec, j_ao);
    if (profiling_) span(T_J_TOTAL, a);
  }
  if (k_bra)
    for (int s = 0; s < nspin_; ++s) {
      HH_REQUIRE(k_bra[s] != nullptr, "k_bra[s] is null");
      const int a = profiling_ ? mark() : -1;
      build_k(s, bra[s], nprop, prec, -1.0, 0.0, k_bra[s]);
      if (profiling_) span(T_K_TOTAL, a);
    }
  if (profiling_) {
    span(T_RESPONSE, r0);
    flush_timing();
  }
}

void Rijk::get_response_bra(const double* const* bra, int64_t nprop, double* const* resp, Prec prec,
                            bool accumulate) {
  if (!prepared_) throw Error(HH_ERR_STATE, "response requested before prepare_response");
  HH_REQUIRE(bra != nullptr && resp != nullptr && nprop > 0, "null argument or nprop <= 0");
  for (int s = 0; s < nspin_; ++s) HH_REQUIRE(bra[s] && resp[s], "bra[s] / resp[s] is null");
  const bool do_j = scale_j_ != 0.0, do_k = scale_k_ != 0.0;
  const int r0 = profiling_ ? mark() : -1;
  if (response_path == 1 && resident_fits(prec)) {
    apply_resident(bra, nprop, resp, prec, accumulate);
    if (profiling_) {
      span(T_RESPONSE, r0);
      flush_timing();
    }
    return;
  }
  if (do_k)
    for (int s = 0; s < nspin_; ++s) {
      const int a = profiling_ ? mark() : -1;
      build_k(s, bra[s], nprop, prec, -scale_k_, accumulate ? 1.0 : 0.0, resp[s]);
      if (profiling_) span(T_K_TOTAL, a);
    }
  const int aj = (profiling_ && do_j) ? mark() : -1;
  if (do_j) {
    double* jao = static_cast<double*>(jao_.reserve(sizeof(double) * nao_ * nao_ * nprop));
    build_j_ao(bra, nprop, prec, jao);
    const double f = (nspin_ == 1) ? 1.0 : 0.5;  // UHF J x0.5 (hess_u.rs:300-303)
    for (int s = 0; s < nspin_; ++s) {
      GemmArgs g;  // resp_s[:, :, a] += scale_j f J_a C_s
      g.M = nao_; g.N = nocc_[s]; g.K = nao_;
      g.A = jao; g.lda = nao_; g.a_kmajor = false; g.sA0 = nao_ * nao_;
      g.B = mocc_[s].ptr; g.ldb = nao_; g.b_kmajor = true;
      g.C = resp[s]; g.ldc = nao_; g.sC0 = nao_ * nocc_[s];
      g.nb0 = nprop;
      g.alpha = scale_j_ * f;
      g.beta = (do_k || accumulate) ? 1.0 : 0.0;
      gemm<double, double>(g, PREC_F64, stream_, ws_);
    }
  }
  if (profiling_ && do_j) span(T_J_TOTAL, aj);
  if (!do_j && !do_k && !accumulate)
    for (int s = 0; s < nspin_; ++s) fill<double>(resp[s], nao_ * nocc_[s] * nprop, 0.0, stream_);
  if (profiling_) {
    span(T_RESPONSE, r0);
    flush_timing();
  }
}

}  // namespace hh
