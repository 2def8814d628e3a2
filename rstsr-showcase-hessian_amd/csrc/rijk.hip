// RI-JK response on gfx950.
//
// Algebra (reference: hess_r.rs:128-183, per perturbation a, per aux index P):
//   D1_a   = bra_a C_o^T + C_o bra_a^T
//   J_ao   = 2 sum_P B_P tr(B_P D1_a)                       (two GEMMs over packed cderi)
//   K_a    = sum_P B_P D1_a B_P C_o
//          = sum_P B_P [ bra_a O_P + C_o (bra_a^T X_P) ],   X_P = B_P C_o,  O_P = C_o^T X_P
// and the reference's k_bra is -K_a.  The reference loops over a and P with small GEMMs; here
// all perturbations are batched into the GEMM N dimension:
//   W[(j,a),(i,P)] = sum_mu bra[mu,(j,a)] X[mu,(i,P)]                 (M=nocc*nprop, K=nao)
//   T[mu,P,i,a]    = sum_j bra[mu,j,a] O[j,i,P] + sum_k C_o[mu,k] W[(k,a),(i,P)]   (K=nocc each)
//   K[u,(i,a)]     = sum_{(mu,P)} B[u,(mu,P)] T[(mu,P),(i,a)]          (K=nao*naux, the hot GEMM)
// with the aux index processed in chunks so T stays a bounded buffer.  X and O depend only on
// the MO coefficients and are built once per prepare_response.  Flop count equals the
// reference's: 2 nao^2 naux nocc nprop + 6 nao naux nocc^2 nprop (+ X/O setup).
#include "rijk.hpp"

#include "elementwise.hpp"

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
    // X_P = B_P C_o  (batched over P): X[mu + nao*(i + nocc*P)]
    double* X = static_cast<double*>(x_[s].reserve(sizeof(double) * nao_ * nocc * naux_));
    {
      GemmArgs g;
      g.M = nao_; g.N = nocc; g.K = nao_;
      g.A = bfull_.ptr; g.lda = nao_; g.a_kmajor = false; g.sA0 = nao_ * nao_;
      g.B = mocc; g.ldb = nao_; g.b_kmajor = true; g.sB0 = 0;
      g.C = X; g.ldc = nao_; g.sC0 = nao_ * nocc;
      g.nb0 = naux_;
      gemm<double, double>(g, PREC_F64, stream_, ws_);
    }
    // O[j, (i,P)] = sum_nu C_o[nu, j] X[nu, (i,P)]
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

void Rijk::build_k(int s, const double* bra, int64_t nprop, Prec prec, double alpha, double beta, double* out) {
  const int64_t nocc = nocc_[s];
  const int64_t ncol = nocc * nprop;  // (i, a) columns
  const bool f64 = (prec == PREC_F64);
  const size_t esz = f64 ? sizeof(double) : sizeof(float);
  if (!f64) ensure_mixed();
  const void* braX = bra;
  if (!f64) {
    float* b32 = static_cast<float*>(bra32_.reserve(sizeof(float) * std::max(nao_ * ncol, std::max(nao_tp_, naux_) * nprop)));
    cast<double, float>(bra, nao_ * ncol, b32, stream_);
    braX = b32;
  }
  const void* X = f64 ? x_[s].ptr : x32_[s].ptr;
  const void* O = f64 ? o_[s].ptr : o32_[s].ptr;
  const void* C = f64 ? mocc_[s].ptr : mocc32_[s].ptr;
  const void* Bf = f64 ? bfull_.ptr : bfull32_.ptr;

  // W[(j,a),(i,P)] = sum_mu bra[mu,(j,a)] X[mu,(i,P)]
  void* W = w_.reserve(esz * ncol * nocc * naux_);
  {
    GemmArgs g;
    g.M = ncol; g.N = nocc * naux_; g.K = nao_;
    g.A = braX; g.lda = nao_; g.a_kmajor = true;
    g.B = X; g.ldb = nao_; g.b_kmajor = true;
    g.C = W; g.ldc = ncol;
    if (f64) gemm<double, double>(g, prec, stream_, ws_);
    else gemm<float, float>(g, prec, stream_, ws_);
  }
  // aux chunk so that T[mu, P, i, a] stays within ~2 GiB
  int64_t pc = k_chunk;
  if (pc <= 0) pc = std::max<int64_t>(1, (int64_t)((2LL << 30) / (esz * nao_ * ncol)));
  pc = std::min(pc, naux_);
  void* T = t_.reserve(esz * nao_ * pc * ncol);
  for (int64_t p0 = 0; p0 < naux_; p0 += pc) {
    const int64_t np = std::min(pc, naux_ - p0);
    // T[mu + nao*(Pl + np*(i + nocc*a))] = sum_j bra[mu, j, a] O[j, i, p0+Pl]   (batch Pl, a)
    {
      GemmArgs g;
      g.M = nao_; g.N = nocc; g.K = nocc;
      g.A = braX; g.lda = nao_; g.a_kmajor = false; g.sA0 = 0; g.sA1 = nao_ * nocc;
      g.B = static_cast<const char*>(O) + esz * (p0 * nocc * nocc); g.ldb = nocc; g.b_kmajor = true;
      g.sB0 = nocc * nocc; g.sB1 = 0;
      g.C = T; g.ldc = nao_ * np; g.sC0 = nao_; g.sC1 = nao_ * np * nocc;
      g.nb0 = np; g.nb1 = nprop;
      if (f64) gemm<double, double>(g, prec, stream_, ws_);
      else gemm<float, float>(g, prec, stream_, ws_);
    }
    //   += sum_k C_o[mu, k] W[k + nocc*a + ncol*(i + nocc*(p0+Pl))]
    {
      GemmArgs g;
      g.M = nao_; g.N = nocc; g.K = nocc;
      g.A = C; g.lda = nao_; g.a_kmajor = false;
      g.B = static_cast<const char*>(W) + esz * (p0 * ncol * nocc); g.ldb = ncol; g.b_kmajor = true;
      g.sB0 = ncol * nocc; g.sB1 = nocc;
      g.C = T; g.ldc = nao_ * np; g.sC0 = nao_; g.sC1 = nao_ * np * nocc;
      g.nb0 = np; g.nb1 = nprop;
      g.beta = 1.0;
      if (f64) gemm<double, double>(g, prec, stream_, ws_);
      else gemm<float, float>(g, prec, stream_, ws_);
    }
    // out[u, (i,a)] = alpha * sum_{(mu,Pl)} B[u, mu, p0+Pl] T[(mu,Pl), (i,a)] + beta' * out
    {
      GemmArgs g;
      g.M = nao_; g.N = ncol; g.K = nao_ * np;
      g.A = static_cast<const char*>(Bf) + esz * (p0 * nao_ * nao_); g.lda = nao_; g.a_kmajor = false;
      g.B = T; g.ldb = nao_ * np; g.b_kmajor = true;
      g.C = out; g.ldc = nao_;
      g.alpha = alpha;
      g.beta = (p0 == 0) ? beta : 1.0;
      if (profiling_) {
        const int e0 = next_event_++, e1 = next_event_++;
        g.ev_begin = ev(e0);
        g.ev_main_end = ev(e1);
        pending_.push_back({T_K_HOT, {e0, e1}});
      }
      if (f64) gemm<double, double>(g, prec, stream_, ws_);
      else gemm<float, double>(g, prec, stream_, ws_);
    }
  }
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
