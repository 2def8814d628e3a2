// extern "C" boundary of libhesship: argument checks, exception -> status translation.
#include "elementwise.hpp"
#include "gemm.hpp"
#include "rijk.hpp"

namespace hh {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace hh

struct hh_rijk {
  std::unique_ptr<hh::Rijk> impl;
};

using namespace hh;

extern "C" {

const char* hh_last_error(void) { return g_last_error.c_str(); }
const char* hh_version(void) { return "hesship 0.1 (gfx950)"; }

hh_status hh_pack_triu_tilde(const double* dm, int64_t nao, int64_t nmat, double* out, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(nao >= 0 && nmat >= 0, "negative size");
  HH_REQUIRE((dm && out) || nao * nmat == 0, "null pointer");
  pack_triu_tilde(dm, nao, nmat, false, out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_unpack_triu(const double* tp, int64_t nao, int64_t nmat, double* out, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(nao >= 0 && nmat >= 0, "negative size");
  HH_REQUIRE((tp && out) || nao * nmat == 0, "null pointer");
  unpack_triu<double, double>(tp, nao, nao * (nao + 1) / 2, nmat, out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_rijk_create(const double* cderi, int64_t nao, int64_t naux, double scale_j, double scale_k,
                         int32_t nspin, void* stream, hh_rijk** out) {
  HH_ABI_BEGIN
  HH_REQUIRE(out != nullptr, "out is null");
  *out = nullptr;
  auto h = std::make_unique<hh_rijk>();
  h->impl = std::make_unique<Rijk>(cderi, nao, naux, scale_j, scale_k, nspin, (hipStream_t)stream);
  *out = h.release();
  HH_ABI_END
}

hh_status hh_rijk_destroy(hh_rijk* h) {
  HH_ABI_BEGIN
  if (h) {
    HH_HIP(hipStreamSynchronize(h->impl->stream()));
    delete h;
  }
  HH_ABI_END
}

hh_status hh_rijk_prepare_response(hh_rijk* h, const double* const* mo_coeff, const double* const* mo_occ,
                                   const int64_t* nmo) {
  HH_ABI_BEGIN
  HH_REQUIRE(h != nullptr, "handle is null");
  h->impl->prepare_response(mo_coeff, mo_occ, nmo);
  HH_ABI_END
}

hh_status hh_rijk_nocc(const hh_rijk* h, int32_t s, int64_t* nocc) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && nocc, "null argument");
  HH_REQUIRE(s >= 0 && s < h->impl->nspin(), "spin index out of range");
  if (!h->impl->prepared()) throw Error(HH_ERR_STATE, "prepare_response not called");
  *nocc = h->impl->nocc(s);
  HH_ABI_END
}

hh_status hh_rijk_get_response_bra(hh_rijk* h, const double* const* bra, int64_t nprop, double* const* resp,
                                   hh_precision precision) {
  HH_ABI_BEGIN
  HH_REQUIRE(h != nullptr, "handle is null");
  HH_REQUIRE(precision >= HH_PREC_FP64 && precision <= HH_PREC_BF16X3, "unknown precision");
  h->impl->get_response_bra(bra, nprop, resp, (Prec)precision);
  HH_ABI_END
}

hh_status hh_rijk_response_bra_separated(hh_rijk* h, const double* const* bra, int64_t nprop, double* j_ao,
                                         double* const* k_bra, hh_precision precision) {
  HH_ABI_BEGIN
  HH_REQUIRE(h != nullptr, "handle is null");
  HH_REQUIRE(precision >= HH_PREC_FP64 && precision <= HH_PREC_BF16X3, "unknown precision");
  h->impl->response_bra_separated(bra, nprop, j_ao, k_bra, (Prec)precision);
  HH_ABI_END
}

hh_status hh_rijk_set_aux_chunk(hh_rijk* h, int64_t naux_chunk) {
  HH_ABI_BEGIN
  HH_REQUIRE(h != nullptr && naux_chunk >= 0, "null handle or negative chunk");
  h->impl->k_chunk = naux_chunk;
  HH_ABI_END
}

hh_status hh_rijk_set_profiling(hh_rijk* h, int32_t on) {
  HH_ABI_BEGIN
  HH_REQUIRE(h != nullptr, "handle is null");
  h->impl->set_profiling(on != 0);
  HH_ABI_END
}

hh_status hh_rijk_get_timing(const hh_rijk* h, double* out5) {
  HH_ABI_BEGIN
  HH_REQUIRE(h != nullptr && out5 != nullptr, "null argument");
  h->impl->get_timing(out5);
  HH_ABI_END
}

hh_status hh_rijk_reset_timing(hh_rijk* h) {
  HH_ABI_BEGIN
  HH_REQUIRE(h != nullptr, "handle is null");
  h->impl->reset_timing();
  HH_ABI_END
}

hh_status hh_util_gemm(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda, int32_t a_kmajor,
                       const double* B, int64_t ldb, int32_t b_kmajor, double beta, double* C, int64_t ldc,
                       int64_t nb0, int64_t nb1, const int64_t* strides6, hh_precision precision, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(precision >= HH_PREC_FP64 && precision <= HH_PREC_BF16X3, "unknown precision");
  // split-K workspace per calling thread, kept for the life of the process (calls from one
  // thread on different streams must not overlap)
  static thread_local DevBuf* ws = new DevBuf;
  GemmArgs g;
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.a_kmajor = a_kmajor != 0;
  g.B = B; g.ldb = ldb; g.b_kmajor = b_kmajor != 0;
  g.C = C; g.ldc = ldc;
  g.alpha = alpha; g.beta = beta;
  g.nb0 = nb0; g.nb1 = nb1;
  if (strides6) {
    g.sA0 = strides6[0]; g.sA1 = strides6[1]; g.sB0 = strides6[2];
    g.sB1 = strides6[3]; g.sC0 = strides6[4]; g.sC1 = strides6[5];
  }
  gemm<double, double>(g, (Prec)precision, (hipStream_t)stream, *ws);
  HH_ABI_END
}

}  // extern "C"
