// extern "C" boundary of libhesship: argument checks, exception -> status translation.
#include "becke.hpp"
#include "elementwise.hpp"
#include "gemm.hpp"
#include "gemm_nt.hpp"
#include "cphf.hpp"
#include "cphf_batch.hpp"
#include "cint.hpp"
#include "hess_core.hpp"
#include "scf.hpp"
#include "linalg.hpp"
#include "mlp.hpp"
#include "numint.hpp"
#include "rijk_skeleton.hpp"
#include "rijk.hpp"
#include "vib.hpp"
#include "xc_functionals.hpp"

namespace hh {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace hh

struct hh_rijk {
  std::unique_ptr<hh::Rijk> impl;
};

namespace {
// RHF RI-JK as a CPHF electron-interaction component
struct RijkResponse : hh::ElecResponse {
  hh::Rijk* r;
  explicit RijkResponse(hh::Rijk* x) : r(x) {}
  int nspin() const override { return r->nspin(); }
  void response_bra(const double* const* bra, int64_t nprop, double* const* out, bool accumulate,
                    hh::Prec prec) override {
    r->get_response_bra(bra, nprop, out, prec, accumulate);
  }
  hipStream_t stream() const override { return r->stream(); }
};

// RKS grid response as a CPHF component: mocc = the driver's first nocc MO columns
struct NumintResponse : hh::ElecResponse {
  hh::Numint* ni;
  int den;
  const double* fxc;
  const double* mocc;
  int64_t nocc;
  NumintResponse(hh::Numint* n, int d, const double* f, const double* m, int64_t no)
      : ni(n), den(d), fxc(f), mocc(m), nocc(no) {}
  int nspin() const override { return 1; }
  void response_bra(const double* const* bra, int64_t nprop, double* const* out, bool accumulate,
                    hh::Prec prec) override {
    ni->rks_response_bra(den, fxc, bra[0], nprop, mocc, nocc, out[0], prec, accumulate);
  }
  hipStream_t stream() const override { return ni->stream(); }
};

// UKS grid response as a CPHF component: mocc[s] = the driver's first nocc[s] MO columns of spin s
struct UksNumintResponse : hh::ElecResponse {
  hh::Numint* ni;
  int den;
  const double* fxc;
  const double* mocc[2];
  int64_t nocc[2];
  UksNumintResponse(hh::Numint* n, int d, const double* f, const double* ma, const double* mb, int64_t na,
                    int64_t nb)
      : ni(n), den(d), fxc(f), mocc{ma, mb}, nocc{na, nb} {}
  int nspin() const override { return 2; }
  void response_bra(const double* const* bra, int64_t nprop, double* const* out, bool accumulate,
                    hh::Prec prec) override {
    ni->uks_response_bra(den, fxc, bra, nprop, mocc, nocc, out, prec, accumulate);
  }
  hipStream_t stream() const override { return ni->stream(); }
};

hh::CphfConfig cphf_config(double level_shift, double tol, int32_t max_cycle, int32_t max_space, double lindep) {
  HH_REQUIRE(tol > 0.0 && max_cycle > 0 && max_space > 0 && lindep >= 0.0, "invalid CPHF configuration");
  hh::CphfConfig cfg;
  cfg.level_shift = level_shift;
  cfg.tol = tol;
  cfg.max_cycle = max_cycle;
  cfg.max_space = max_space;
  cfg.lindep = lindep;
  return cfg;
}

void check_prec(hh_precision precision) {
  HH_REQUIRE(precision >= HH_PREC_FP64 && precision <= HH_PREC_BF16X3, "unknown precision");
}

void fill_stats(hh_krylov_stats* out, const hh::KrylovStats& k, int refinements, double residual,
                int fp64_fallback = 0) {
  if (!out) return;
  out->fp64_fallback = fp64_fallback;
  out->cycles = k.cycles;
  out->restarts = k.restarts;
  out->aop_calls = k.aop_calls;
  out->refinements = refinements;
  out->max_innerprod = k.max_innerprod;
  out->residual = residual;
}
}  // namespace

struct hh_numint {
  std::unique_ptr<hh::Numint> impl;
};

struct hh_becke {
  std::unique_ptr<hh::Becke> impl;
};

struct hh_mlp {
  std::unique_ptr<hh::DiagMlp> impl;
};

struct hh_cphf {
  std::unique_ptr<hh::Cphf> impl;
  std::vector<std::unique_ptr<hh::ElecResponse>> comps;
};

struct hh_cint {
  std::unique_ptr<hh::Cint> impl;
};
struct hh_cint_pair {
  const hh::Cint* mol;
  const hh::Cint* aux;
};

struct hh_cphf_batch {
  std::unique_ptr<hh::CphfBatch> impl;
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

hh_status hh_rijk_create_from_j3c_decomp(const double* j3c, const double* j2c, int64_t nao, int64_t naux,
                                         int32_t policy, double threshold, double scale_j, double scale_k,
                                         int32_t nspin, void* stream, hh_rijk** out) {
  HH_ABI_BEGIN
  HH_REQUIRE(out != nullptr, "out is null");
  *out = nullptr;
  HH_REQUIRE(nspin == 1 || nspin == 2, "nspin must be 1 (RHF) or 2 (UHF)");
  auto h = std::make_unique<hh_rijk>();
  h->impl = Rijk::from_j3c(j3c, j2c, nao, naux, policy, threshold, scale_j, scale_k, nspin, (hipStream_t)stream);
  *out = h.release();
  HH_ABI_END
}

hh_status hh_rijk_create_from_j3c(const double* j3c, const double* j2c, int64_t nao, int64_t naux, double threshold,
                                  double scale_j, double scale_k, int32_t nspin, void* stream, hh_rijk** out) {
  return hh_rijk_create_from_j3c_decomp(j3c, j2c, nao, naux, HH_J2C_CD, threshold, scale_j, scale_k, nspin, stream,
                                        out);
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

hh_status hh_rijk_set_batch(hh_rijk* h, int64_t nprop_batch) {
  HH_ABI_BEGIN
  HH_REQUIRE(h != nullptr && nprop_batch >= 0, "null handle or negative batch");
  h->impl->k_chunk = nprop_batch;
  bump_response_state();
  HH_ABI_END
}

hh_status hh_rijk_set_response_path(hh_rijk* h, int32_t path, int64_t budget_bytes) {
  HH_ABI_BEGIN
  HH_REQUIRE(h != nullptr && (path == 0 || path == 1), "null handle or unknown response path");
  h->impl->response_path = path;
  if (budget_bytes > 0) h->impl->resident_budget = budget_bytes;
  bump_response_state();
  HH_ABI_END
}

hh_status hh_rijk_set_k_path(hh_rijk* h, int32_t path) {
  HH_ABI_BEGIN
  HH_REQUIRE(h != nullptr && (path == 0 || path == 2), "null handle or unknown K path");
  h->impl->k_path = path;
  bump_response_state();
  HH_ABI_END
}

hh_status hh_rijk_set_profiling(hh_rijk* h, int32_t on) {
  HH_ABI_BEGIN
  HH_REQUIRE(h != nullptr, "handle is null");
  h->impl->set_profiling(on != 0);
  bump_response_state();
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

hh_status hh_krylov_orth_block(double* V, int64_t n, int64_t m, double lindep, double* R, double* norms_out,
                               int64_t* kept, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(V && R && norms_out && kept && n > 0 && m >= 0, "orth_block: bad argument");
  static thread_local DevBuf* ws = new DevBuf;
  std::vector<double> nr;
  const int64_t k = orth_block(V, n, m, lindep, R, 0, nr, (hipStream_t)stream, *ws);
  for (int64_t i = 0; i < k; ++i) norms_out[i] = nr[i];
  *kept = k;
  HH_ABI_END
}

hh_status hh_krylov_block(hh_aop_fn aop, void* ctx, const double* b, const double* x0, int64_t n, int64_t nset,
                          double tol, int32_t max_cycle, int32_t max_space, double lindep, double* x,
                          hh_krylov_stats* stats, void* stream) {
  hh::KrylovStats ks;
  try {
    HH_REQUIRE(aop && b && x, "null argument");
    hipStream_t s = (hipStream_t)stream;
    hh::Aop f = [&](const double* xv, int64_t nb, double* axv) {
      const int32_t rc = aop(ctx, xv, n, nb, axv);
      if (rc != 0) throw hh::Error(HH_ERR_INVALID, "aop callback returned " + std::to_string(rc));
    };
    hh::krylov_block(f, b, x0, n, nset, tol, max_cycle, max_space, lindep, x, &ks, s);
    fill_stats(stats, ks, 0, 0.0);
    return HH_OK;
  } catch (const hh::Error& e) {
    fill_stats(stats, ks, 0, 0.0);
    hh::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    hh::set_last_error(e.what());
    return HH_ERR_INTERNAL;
  }
}

hh_status hh_cphf_create(const double* mo_coeff, int64_t nao, int64_t nmo, const double* mo_occ,
                         const double* mo_energy, double level_shift, double tol, int32_t max_cycle,
                         int32_t max_space, double lindep, void* stream, hh_cphf** out) {
  HH_ABI_BEGIN
  HH_REQUIRE(out != nullptr, "out is null");
  *out = nullptr;
  const CphfConfig cfg = cphf_config(level_shift, tol, max_cycle, max_space, lindep);
  const double* c[1] = {mo_coeff};
  const double* o[1] = {mo_occ};
  const double* e[1] = {mo_energy};
  auto h = std::make_unique<hh_cphf>();
  h->impl = std::make_unique<Cphf>(1, c, nao, &nmo, o, e, cfg, (hipStream_t)stream);
  *out = h.release();
  HH_ABI_END
}

hh_status hh_ucphf_create(const double* const* mo_coeff, int64_t nao, const int64_t* nmo,
                          const double* const* mo_occ, const double* const* mo_energy, double level_shift, double tol,
                          int32_t max_cycle, int32_t max_space, double lindep, void* stream, hh_cphf** out) {
  HH_ABI_BEGIN
  HH_REQUIRE(out != nullptr, "out is null");
  *out = nullptr;
  const CphfConfig cfg = cphf_config(level_shift, tol, max_cycle, max_space, lindep);
  auto h = std::make_unique<hh_cphf>();
  h->impl = std::make_unique<Cphf>(2, mo_coeff, nao, nmo, mo_occ, mo_energy, cfg, (hipStream_t)stream);
  *out = h.release();
  HH_ABI_END
}

hh_status hh_cphf_destroy(hh_cphf* h) {
  HH_ABI_BEGIN
  if (h) {
    HH_HIP(hipStreamSynchronize(h->impl->stream()));
    delete h;
  }
  HH_ABI_END
}

hh_status hh_cphf_set_verbose(hh_cphf* h, int32_t on) {
  HH_ABI_BEGIN
  HH_REQUIRE(h != nullptr, "handle is null");
  h->impl->verbose = on != 0;
  HH_ABI_END
}

hh_status hh_cphf_set_graphs(hh_cphf* h, int32_t on) {
  HH_ABI_BEGIN
  HH_REQUIRE(h != nullptr, "handle is null");
  h->impl->use_graphs = on != 0;
  HH_ABI_END
}

hh_status hh_cphf_graph_stats(hh_cphf* h, int64_t* captures, int64_t* launches) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && captures && launches, "null argument");
  *captures = h->impl->graph_captures();
  *launches = h->impl->graph_launches();
  HH_ABI_END
}

hh_status hh_cphf_finalize_reuses(hh_cphf* h, int64_t* count) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && count, "null argument");
  *count = h->impl->finalize_reuses();
  HH_ABI_END
}

hh_status hh_cphf_add_rijk(hh_cphf* h, hh_rijk* rijk) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && rijk, "null handle");
  HH_REQUIRE(rijk->impl->nspin() == h->impl->nspin(),
             "RI-JK component spin count differs from the driver's (RHF: nspin 1, UHF: nspin 2)");
  if (!rijk->impl->prepared()) throw Error(HH_ERR_STATE, "RI-JK component not prepared (make_response_preparation)");
  for (int t = 0; t < h->impl->nspin(); ++t)
    HH_REQUIRE(rijk->impl->nocc(t) == h->impl->nocc(t), "RI-JK component occupation differs from the driver's");
  h->comps.push_back(std::make_unique<RijkResponse>(rijk->impl.get()));
  h->impl->add_component(h->comps.back().get());
  HH_ABI_END
}

hh_status hh_cint_create(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas, const double* env,
                         int64_t nenv, int32_t cart, hh_cint** out) {
  HH_ABI_BEGIN
  HH_REQUIRE(out, "null output handle");
  auto* h = new hh_cint;
  try {
    h->impl = std::make_unique<Cint>(atm, natm, bas, nbas, env, nenv, cart != 0);
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  HH_ABI_END
}

hh_status hh_cint_destroy(hh_cint* h) {
  HH_ABI_BEGIN
  delete h;
  HH_ABI_END
}

hh_status hh_cint_nao(const hh_cint* h, int64_t* nao, int64_t* ao_loc) {
  HH_ABI_BEGIN
  HH_REQUIRE(h, "null handle");
  if (nao) *nao = h->impl->nao();
  if (ao_loc)
    for (int i = 0; i <= h->impl->nbas(); ++i) ao_loc[i] = h->impl->ao_loc(i);
  HH_ABI_END
}

hh_status hh_cint_eval_gto(const hh_cint* h, int32_t deriv, const double* coords, int64_t ng, double* out,
                           void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(h, "null handle");
  h->impl->eval_gto(deriv, coords, ng, out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_cint_int2c2e(const hh_cint* aux, int32_t kind, double* out, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(aux, "null handle");
  aux->impl->int2c2e(kind, 0, aux->impl->nbas(), out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_cint_int3c2e(const hh_cint* mol, const hh_cint* aux, int32_t kind, int32_t sh0, int32_t sh1,
                          double* out, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(mol && aux, "null handle");
  Cint::int3c2e(*mol->impl, *aux->impl, kind, sh0, sh1, out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_scf_rhf(const hh_cint* mol, const hh_cint* aux, int32_t nocc, double conv_tol, int32_t max_cycle,
                     double* mo_coeff, double* mo_energy, double* energy, int32_t* cycles, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(mol && aux, "null argument");
  const ScfResult r = rhf_scf(*mol->impl, *aux->impl, nocc, conv_tol, max_cycle, mo_coeff, mo_energy,
                              (hipStream_t)stream);
  if (energy) *energy = r.energy;
  if (cycles) *cycles = r.cycles;
  HH_ABI_END
}

hh_status hh_scf_rks(const hh_cint* mol, const hh_cint* aux, hh_numint* grid, int32_t den_type, int32_t nfunc,
                     const int32_t* xc_ids, const double* xc_scales, double exx, int32_t nocc, double conv_tol,
                     int32_t max_cycle, double* mo_coeff, double* mo_energy, double* energy, int32_t* cycles,
                     void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(mol && aux && grid, "null argument");
  HH_REQUIRE(den_type >= HH_DEN_RHO && den_type <= HH_DEN_TAU, "den_type must be RHO, SIGMA or TAU");
  std::vector<int> ids(xc_ids, xc_ids + std::max(0, nfunc));
  ScfXc xc;
  xc.ni = grid->impl.get();
  xc.den = den_type;
  xc.nfunc = nfunc;
  xc.ids = ids.data();
  xc.scales = xc_scales;
  xc.exx = exx;
  const ScfResult r = rhf_scf(*mol->impl, *aux->impl, nocc, conv_tol, max_cycle, mo_coeff, mo_energy,
                              (hipStream_t)stream, &xc);
  if (energy) *energy = r.energy;
  if (cycles) *cycles = r.cycles;
  HH_ABI_END
}

hh_status hh_scf_uhf(const hh_cint* mol, const hh_cint* aux, hh_numint* grid, int32_t den_type, int32_t nfunc,
                     const int32_t* xc_ids, const double* xc_scales, double exx, int32_t nalpha, int32_t nbeta,
                     double conv_tol, int32_t max_cycle, double* const* mo_coeff, double* const* mo_energy,
                     double* energy, int32_t* cycles, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(mol && aux && mo_coeff && mo_energy, "null argument");
  ScfXc xc;
  std::vector<int> ids;
  if (grid) {
    HH_REQUIRE(den_type >= HH_DEN_RHO && den_type <= HH_DEN_TAU, "den_type must be RHO, SIGMA or TAU");
    ids.assign(xc_ids, xc_ids + std::max(0, nfunc));
    xc.ni = grid->impl.get();
    xc.den = den_type;
    xc.nfunc = nfunc;
    xc.ids = ids.data();
    xc.scales = xc_scales;
    xc.exx = exx;
  }
  const ScfResult r = uhf_scf(*mol->impl, *aux->impl, nalpha, nbeta, conv_tol, max_cycle, mo_coeff, mo_energy,
                              (hipStream_t)stream, grid ? &xc : nullptr);
  if (energy) *energy = r.energy;
  if (cycles) *cycles = r.cycles;
  HH_ABI_END
}

hh_status hh_numint_vxc_matrix(hh_numint* h, int32_t den_type, const double* vxc_eff, double* out) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && vxc_eff && out, "null argument");
  h->impl->vxc_matrix(den_type, vxc_eff, out);
  HH_HIP(hipStreamSynchronize(h->impl->stream()));
  HH_ABI_END
}

hh_status hh_hess_nuc_repl(const hh_cint* mol, double* out, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(mol && out, "null argument");
  nuc_repl_hess(*mol->impl, out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_hess_hcore(const hh_cint* mol, const double* dm0, double* out, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(mol && dm0 && out, "null argument");
  hcore_hess(*mol->impl, dm0, out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_hess_ovlp(const hh_cint* mol, const double* dme0, double* out, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(mol && dme0 && out, "null argument");
  ovlp_hess(*mol->impl, dme0, out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_hess_hcore_deriv1(const hh_cint* mol, double* out, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(mol && out, "null argument");
  hcore_deriv1(*mol->impl, out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_hess_ovlp_deriv1(const hh_cint* mol, double* out, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(mol && out, "null argument");
  ovlp_deriv1(*mol->impl, out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_hess_weighted_density(const double* mo_coeff, int64_t nao, int64_t nmo, const double* w, double* out,
                                   void* stream) {
  HH_ABI_BEGIN
  weighted_density(mo_coeff, nao, nmo, w, out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_cint_int1e(const hh_cint* mol, int32_t op, int32_t kind, int32_t atom, double* out, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(mol && out, "null argument");
  const hipStream_t st = (hipStream_t)stream;
  if (op == HH_INT1E_OVLP || op == HH_INT1E_KIN) {
    mol->impl->int1e(op, kind, out, st);
  } else if (op == HH_INT1E_NUC) {
    mol->impl->int1e_nuc(kind, -1, out, st);
  } else {
    HH_REQUIRE(op == HH_INT1E_RINV, "int1e: unknown operator");
    HH_REQUIRE(atom >= 0, "int1e: RINV needs a nucleus index");
    mol->impl->int1e_nuc(kind, atom, out, st);
  }
  HH_ABI_END
}

hh_status hh_cint_int3c2e_s2ij(const hh_cint* mol, const hh_cint* aux, int32_t sh0, int32_t sh1, double* out,
                               void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(mol && aux, "null handle");
  Cint::int3c2e_s2ij(*mol->impl, *aux->impl, sh0, sh1, out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_cint_pair_create(const hh_cint* mol, const hh_cint* aux, hh_cint_pair** out) {
  HH_ABI_BEGIN
  HH_REQUIRE(mol && aux && out, "null argument");
  *out = new hh_cint_pair{mol->impl.get(), aux->impl.get()};
  HH_ABI_END
}

hh_status hh_cint_pair_destroy(hh_cint_pair* p) {
  HH_ABI_BEGIN
  delete p;
  HH_ABI_END
}

int32_t hh_cint_int3c_deriv(void* ctx, int32_t kind, int64_t p0, int64_t p1, double* out, void* stream) {
  HH_ABI_BEGIN
  auto* pr = static_cast<hh_cint_pair*>(ctx);
  HH_REQUIRE(pr && pr->mol && pr->aux, "null integral pair");
  static const int map[6] = {HH_INT_IP1, HH_INT_IP2, HH_INT_IPIP1, HH_INT_IPVIP1, HH_INT_IP1IP2, HH_INT_IPIP2};
  HH_REQUIRE(kind >= 0 && kind < 6, "unknown skeleton integral kind");
  // AO range -> shell range
  int sh0 = -1, sh1 = -1;
  for (int i = 0; i <= pr->aux->nbas(); ++i) {
    if (pr->aux->ao_loc(i) == p0) sh0 = i;
    if (pr->aux->ao_loc(i) == p1) sh1 = i;
  }
  HH_REQUIRE(sh0 >= 0 && sh1 >= sh0, "auxiliary range not shell aligned");
  Cint::int3c2e(*pr->mol, *pr->aux, map[kind], sh0, sh1, out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_cphf_batch_create(double tol, int32_t max_cycle, double lindep, void* stream, hh_cphf_batch** out) {
  HH_ABI_BEGIN
  HH_REQUIRE(out, "null output handle");
  auto* b = new hh_cphf_batch;
  try {
    b->impl = std::make_unique<CphfBatch>(tol, max_cycle, lindep, (hipStream_t)stream);
  } catch (...) {
    delete b;
    throw;
  }
  *out = b;
  HH_ABI_END
}

hh_status hh_cphf_batch_destroy(hh_cphf_batch* b) {
  HH_ABI_BEGIN
  delete b;
  HH_ABI_END
}

hh_status hh_cphf_batch_add(hh_cphf_batch* b, hh_cphf* drv, hh_rijk* rijk, const double* const* rhs, int64_t nprop,
                            double* const* mo1) {
  HH_ABI_BEGIN
  HH_REQUIRE(b && drv && rijk, "null handle");
  b->impl->add(drv->impl.get(), rijk->impl.get(), rhs, nprop, mo1);
  HH_ABI_END
}

hh_status hh_cphf_batch_solve(hh_cphf_batch* b, hh_krylov_stats* stats) {
  HH_ABI_BEGIN
  HH_REQUIRE(b, "null handle");
  std::vector<KrylovStats> st;
  Error err(HH_OK, "");
  bool failed = false;
  try {
    b->impl->solve(&st);
  } catch (const Error& e) {
    if (e.code != HH_ERR_NOT_CONVERGED) throw;
    err = e;
    failed = true;
  }
  if (stats)
    for (size_t i = 0; i < st.size(); ++i) {
      hh_krylov_stats k{};
      k.cycles = st[i].cycles;
      k.restarts = st[i].restarts;
      k.aop_calls = st[i].aop_calls;
      k.max_innerprod = st[i].max_innerprod;
      stats[i] = k;
    }
  if (failed) throw err;
  HH_ABI_END
}

hh_status hh_cphf_batch_info(hh_cphf_batch* b, int64_t* nsys, int64_t* cycles, int64_t* launches_per_cycle) {
  HH_ABI_BEGIN
  HH_REQUIRE(b, "null handle");
  if (nsys) *nsys = (int64_t)b->impl->size();
  if (cycles) *cycles = b->impl->cycles();
  if (launches_per_cycle) *launches_per_cycle = b->impl->launches_per_cycle();
  HH_ABI_END
}

hh_status hh_cphf_batch_times(hh_cphf_batch* b, double* seconds4) {
  HH_ABI_BEGIN
  HH_REQUIRE(b && seconds4, "null argument");
  for (int i = 0; i < 4; ++i) seconds4[i] = b->impl->phase_seconds()[i];
  HH_ABI_END
}

hh_status hh_rijk_skeleton(hh_rijk* h, const double* const* mo_coeff, const int64_t* nmo, const double* const* mo_occ,
                           const double* j2c, const double* j2c_ip1, const double* j2c_ipip1, const double* j2c_ip1ip2,
                           hh_int3c_deriv_fn int3c, void* ctx, int64_t nrange, const int64_t* aux_ranges, int64_t natm,
                           const int64_t* aoslice, const int64_t* auxslice, double* de, double* const* deriv1_bra) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && mo_coeff && nmo && mo_occ, "null argument");
  hh::SkeletonInput in;
  in.mo_coeff = mo_coeff;
  in.nmo = nmo;
  in.mo_occ = mo_occ;
  in.j2c = j2c;
  in.j2c_ip1 = j2c_ip1;
  in.j2c_ipip1 = j2c_ipip1;
  in.j2c_ip1ip2 = j2c_ip1ip2;
  in.int3c = int3c;
  in.ctx = ctx;
  in.nrange = nrange;
  in.aux_ranges = aux_ranges;
  in.natm = natm;
  in.aoslice = aoslice;
  in.auxslice = auxslice;
  hh::rijk_skeleton(*h->impl, in, de, deriv1_bra);
  HH_ABI_END
}

hh_status hh_cphf_add_numint(hh_cphf* h, hh_numint* ni, int32_t den_type, const double* fxc_eff) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && ni && fxc_eff, "null argument");
  HH_REQUIRE(h->impl->nspin() == 1, "the grid response component is restricted (RKS)");
  HH_REQUIRE(den_type >= HH_DEN_RHO && den_type <= HH_DEN_TAU, "den_type must be RHO, SIGMA or TAU");
  HH_REQUIRE(ni->impl->nao() == h->impl->nao(), "numint nao differs from the driver's");
  h->comps.push_back(std::make_unique<NumintResponse>(ni->impl.get(), den_type, fxc_eff, h->impl->mo_coeff(),
                                                      h->impl->nocc()));
  h->impl->add_component(h->comps.back().get());
  HH_ABI_END
}

hh_status hh_ucphf_add_numint(hh_cphf* h, hh_numint* ni, int32_t den_type, const double* fxc_eff) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && ni && fxc_eff, "null argument");
  HH_REQUIRE(h->impl->nspin() == 2, "this entry point is for the unrestricted driver (hh_ucphf_create)");
  HH_REQUIRE(den_type >= HH_DEN_RHO && den_type <= HH_DEN_TAU, "den_type must be RHO, SIGMA or TAU");
  HH_REQUIRE(ni->impl->nao() == h->impl->nao(), "numint nao differs from the driver's");
  h->comps.push_back(std::make_unique<UksNumintResponse>(ni->impl.get(), den_type, fxc_eff, h->impl->mo_coeff(0),
                                                         h->impl->mo_coeff(1), h->impl->nocc(0), h->impl->nocc(1)));
  h->impl->add_component(h->comps.back().get());
  HH_ABI_END
}

hh_status hh_numint_create(const double* ao, int64_t ng, int64_t nao, int64_t ncomp, const double* weights,
                           void* stream, hh_numint** out) {
  HH_ABI_BEGIN
  HH_REQUIRE(out != nullptr, "out is null");
  *out = nullptr;
  auto h = std::make_unique<hh_numint>();
  h->impl = std::make_unique<Numint>(ao, ng, nao, ncomp, weights, (hipStream_t)stream);
  *out = h.release();
  HH_ABI_END
}

hh_status hh_numint_destroy(hh_numint* h) {
  HH_ABI_BEGIN
  if (h) {
    HH_HIP(hipStreamSynchronize(h->impl->stream()));
    delete h;
  }
  HH_ABI_END
}

hh_status hh_numint_set_batch(hh_numint* h, int64_t nbatch) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && nbatch >= 0, "null handle or negative batch");
  h->impl->batch = nbatch;
  bump_response_state();
  HH_ABI_END
}

hh_status hh_numint_set_fused(hh_numint* h, int32_t on) {
  HH_ABI_BEGIN
  HH_REQUIRE(h, "null handle");
  h->impl->fused_kets = on != 0;
  bump_response_state();
  HH_ABI_END
}

hh_status hh_numint_rho1_one_bra_mult_ket(hh_numint* h, int32_t den_type, const double* bra, int64_t nocc,
                                          const double* kets, int64_t nprop, double* rho1, hh_precision precision) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && bra && kets && rho1 && nocc > 0 && nprop > 0, "null argument or empty size");
  HH_REQUIRE(precision >= HH_PREC_FP64 && precision <= HH_PREC_BF16X3, "unknown precision");
  h->impl->rho1_one_bra_mult_ket(den_type, bra, nocc, kets, nprop, rho1, (Prec)precision);
  HH_ABI_END
}

hh_status hh_numint_fxc_pot_bra_trans(hh_numint* h, int32_t den_type, const double* fxc_eff, const double* rho1,
                                      int64_t nprop, const double* bra, int64_t nocc, double* out,
                                      hh_precision precision) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && fxc_eff && rho1 && bra && out && nocc > 0 && nprop > 0, "null argument or empty size");
  HH_REQUIRE(precision >= HH_PREC_FP64 && precision <= HH_PREC_BF16X3, "unknown precision");
  h->impl->fxc_pot_bra_trans(den_type, fxc_eff, rho1, nprop, bra, nocc, out, (Prec)precision);
  HH_ABI_END
}

hh_status hh_numint_rks_response_bra(hh_numint* h, int32_t den_type, const double* fxc_eff, const double* mo1_bra,
                                     int64_t nprop, const double* mocc, int64_t nocc, double* out,
                                     hh_precision precision) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && fxc_eff && mo1_bra && mocc && out && nocc > 0 && nprop > 0, "null argument or empty size");
  HH_REQUIRE(precision >= HH_PREC_FP64 && precision <= HH_PREC_BF16X3, "unknown precision");
  h->impl->rks_response_bra(den_type, fxc_eff, mo1_bra, nprop, mocc, nocc, out, (Prec)precision);
  HH_ABI_END
}

hh_status hh_numint_uks_fxc_pot_bra_trans(hh_numint* h, int32_t den_type, const double* fxc_eff, const double* rho1,
                                          int64_t nprop, const double* const* bra, const int64_t* nocc,
                                          double* const* out, hh_precision precision) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && fxc_eff && rho1 && bra && nocc && out && nprop > 0, "null argument or empty size");
  for (int s = 0; s < 2; ++s) HH_REQUIRE(bra[s] && out[s] && nocc[s] > 0, "null spin array or empty occupation");
  check_prec(precision);
  h->impl->uks_fxc_pot_bra_trans(den_type, fxc_eff, rho1, nprop, bra, nocc, out, (Prec)precision);
  HH_ABI_END
}

hh_status hh_numint_uks_response_bra(hh_numint* h, int32_t den_type, const double* fxc_eff,
                                     const double* const* mo1_bra, int64_t nprop, const double* const* mocc,
                                     const int64_t* nocc, double* const* out, hh_precision precision) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && fxc_eff && mo1_bra && mocc && nocc && out && nprop > 0, "null argument or empty size");
  for (int s = 0; s < 2; ++s)
    HH_REQUIRE(mo1_bra[s] && mocc[s] && out[s] && nocc[s] > 0, "null spin array or empty occupation");
  check_prec(precision);
  h->impl->uks_response_bra(den_type, fxc_eff, mo1_bra, nprop, mocc, nocc, out, (Prec)precision);
  HH_ABI_END
}

hh_status hh_numint_setup_rho(hh_numint* h, int32_t den_type, int32_t nspin, const double* const* dm0,
                              double* rho) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && dm0 && rho && (nspin == 1 || nspin == 2), "null argument or nspin not 1 / 2");
  for (int s = 0; s < nspin; ++s) HH_REQUIRE(dm0[s], "null dm0");
  h->impl->setup_rho(den_type, nspin, dm0, rho);
  HH_ABI_END
}

hh_status hh_numint_hessian_setup(hh_numint* h, int32_t den_type, int32_t nspin, const double* const* dm0,
                                  int64_t natm, const int64_t* aoslice, const double* vxc_eff, const double* fxc_eff,
                                  double* de_fxc, double* const* de_vxc_diag, double* const* de_vxc_off,
                                  double* const* vmat_ip, double* const* vmat_deriv1) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && dm0 && aoslice && vxc_eff && fxc_eff && de_fxc && de_vxc_diag && de_vxc_off && vmat_ip &&
                 vmat_deriv1 && (nspin == 1 || nspin == 2) && natm > 0,
             "null argument, nspin not 1 / 2 or no atoms");
  Numint::SetupOut o;
  o.de_fxc = de_fxc;
  for (int s = 0; s < nspin; ++s) {
    o.de_vxc_diag[s] = de_vxc_diag[s];
    o.de_vxc_off[s] = de_vxc_off[s];
    o.vmat_ip[s] = vmat_ip[s];
    o.vmat_deriv1[s] = vmat_deriv1[s];
  }
  h->impl->hessian_setup(den_type, nspin, dm0, natm, aoslice, vxc_eff, fxc_eff, o);
  HH_ABI_END
}

hh_status hh_numint_hessian_setup_becke(hh_numint* h, int32_t den_type, int32_t nspin, const double* const* dm0,
                                        int64_t natm, const int64_t* aoslice, const double* exc, const double* vxc_eff,
                                        const double* fxc_eff, hh_becke* becke, const double* grid_coords,
                                        const double* quadrature_weights, const int64_t* atm_quad_split,
                                        int32_t hardness, double* de_xc_skeleton, double* const* vmat_deriv1_grid,
                                        double* de_fxc, double* const* de_vxc_diag, double* const* de_vxc_off,
                                        double* const* vmat_ip, double* const* vmat_deriv1) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && dm0 && aoslice && exc && vxc_eff && fxc_eff && becke && grid_coords && quadrature_weights &&
                 atm_quad_split && de_xc_skeleton && vmat_deriv1_grid && (nspin == 1 || nspin == 2) && natm > 0,
             "null argument, nspin not 1 / 2 or no atoms");
  Numint::SetupOut o;
  o.de_fxc = de_fxc;
  Numint::BeckeOut bo;
  bo.de_xc_skeleton = de_xc_skeleton;
  for (int s = 0; s < nspin; ++s) {
    o.de_vxc_diag[s] = de_vxc_diag ? de_vxc_diag[s] : nullptr;
    o.de_vxc_off[s] = de_vxc_off ? de_vxc_off[s] : nullptr;
    o.vmat_ip[s] = vmat_ip ? vmat_ip[s] : nullptr;
    o.vmat_deriv1[s] = vmat_deriv1 ? vmat_deriv1[s] : nullptr;
    bo.vmat_deriv1_grid[s] = vmat_deriv1_grid[s];
  }
  Numint::BeckeIn bi;
  bi.becke = becke->impl.get();
  bi.coords = grid_coords;
  bi.wquad = quadrature_weights;
  bi.split = atm_quad_split;
  bi.hardness = hardness;
  bi.exc = exc;
  h->impl->hessian_setup_becke(den_type, nspin, dm0, natm, aoslice, vxc_eff, fxc_eff, bi, o, bo);
  HH_ABI_END
}

hh_status hh_becke_create(int64_t natm, const double* atm_coords, const double* adjustment, void* stream,
                          hh_becke** out) {
  HH_ABI_BEGIN
  HH_REQUIRE(out != nullptr, "out is null");
  *out = nullptr;
  auto b = std::make_unique<hh_becke>();
  b->impl = std::make_unique<Becke>(natm, atm_coords, adjustment, (hipStream_t)stream);
  *out = b.release();
  HH_ABI_END
}

hh_status hh_becke_destroy(hh_becke* b) {
  HH_ABI_BEGIN
  if (b) {
    HH_HIP(hipStreamSynchronize(b->impl->stream()));
    delete b;
  }
  HH_ABI_END
}

hh_status hh_becke_partition(hh_becke* b, const double* grid_coords, int64_t ngrids, int32_t by_atom,
                             const int64_t* atm_indices, const double* quadrature_weights, int32_t hardness,
                             int32_t deriv, double* w, double* dw, double* ddw, const double* contract_w,
                             int64_t nset_w, double* c, const double* contract_dw, int64_t nset_dw, double* dc,
                             const double* contract_ddw, int64_t nset_ddw, double* ddc) {
  HH_ABI_BEGIN
  HH_REQUIRE(b != nullptr, "handle is null");
  Becke::Out o;
  o.w = w;
  o.dw = dw;
  o.ddw = ddw;
  o.contract_w = contract_w;
  o.nset_w = nset_w;
  o.c = c;
  o.contract_dw = contract_dw;
  o.nset_dw = nset_dw;
  o.dc = dc;
  o.contract_ddw = contract_ddw;
  o.nset_ddw = nset_ddw;
  o.ddc = ddc;
  b->impl->partition(grid_coords, ngrids, by_atom != 0, atm_indices, quadrature_weights, hardness, deriv, o);
  HH_ABI_END
}

hh_status hh_xceff_transform_polarized(int32_t den_type, int32_t order, const double* rho, int64_t ld_rho,
                                       int64_t ld_spin, const double* xc_val, int64_t ld_xc, int64_t ng, double* out,
                                       void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(ng >= 0 && xc_val && out && (den_type == HH_DEN_RHO || rho), "null argument");
  xceff_transform_pol(den_type, order, rho, ld_rho, ld_spin, xc_val, ld_xc, ng, out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_xceff_libxc_to_xcfun_map(int32_t den_type, int32_t deriv, int64_t* map, int64_t capacity,
                                      int64_t* len) {
  HH_ABI_BEGIN
  HH_REQUIRE(len != nullptr, "len is null");
  const std::vector<int64_t> m = libxc_to_xcfun_map(den_type, deriv);
  *len = (int64_t)m.size();
  if (map) {
    HH_REQUIRE(capacity >= (int64_t)m.size(), "map capacity too small");
    std::copy(m.begin(), m.end(), map);
  }
  HH_ABI_END
}

hh_status hh_xceff_transform(int32_t den_type, int32_t order, const double* rho, int64_t ld_rho, const double* xc_val,
                             int64_t ld_xc, int64_t ng, double* out, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(ng >= 0 && xc_val && out && (den_type == HH_DEN_RHO || rho), "null argument");
  xceff_transform(den_type, order, rho, ld_rho, xc_val, ld_xc, ng, out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_xc_info(int32_t xc_id, int32_t nspin, int32_t deriv, int32_t* den_type, double* exx, int64_t* ncol) {
  HH_ABI_BEGIN
  int fam = 0;
  double x = 0.0;
  HH_REQUIRE(xc_known(xc_id, &fam, &x), "xc: unsupported functional id (LibXC ids 1, 7, 8, 106, 131, 402, 202, 231, 396)");
  HH_REQUIRE(nspin == 1 || nspin == 2, "xc: nspin must be 1 or 2");
  HH_REQUIRE(deriv >= 0 && deriv <= 2, "xc: derivative order 0..2");
  if (den_type) *den_type = fam == 0 ? HH_DEN_RHO : (fam == 1 ? HH_DEN_SIGMA : HH_DEN_TAU);
  if (exx) *exx = x;
  if (ncol) *ncol = xc_ncol(fam, nspin, deriv);
  HH_ABI_END
}

hh_status hh_xc_eval(int32_t xc_id, int32_t nspin, int32_t deriv, const double* rho, int64_t ld_rho, int64_t ld_spin,
                     int64_t ng, double* out, int64_t ld_out, void* stream) {
  HH_ABI_BEGIN
  xc_eval(xc_id, nspin, deriv, rho, ld_rho, ld_spin, ng, 1.0, false, out, ld_out, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_xc_eval_eff(int32_t xc_id, int32_t nspin, int32_t den_type, double scale, const double* rho,
                         int64_t ld_rho, int64_t ld_spin, int64_t ng, int32_t accumulate, double* exc, double* vxc_eff,
                         double* fxc_eff, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(den_type == HH_DEN_RHO || den_type == HH_DEN_SIGMA || den_type == HH_DEN_TAU, "xc: bad den_type");
  const int nvl = den_type == HH_DEN_RHO ? 1 : den_type == HH_DEN_SIGMA ? 4 : 5;
  xc_eval_eff(xc_id, nspin, nvl, rho, ld_rho, ld_spin, ng, scale, accumulate != 0, exc, vxc_eff, fxc_eff,
              (hipStream_t)stream);
  HH_ABI_END
}

}  // extern "C"

namespace {
void need_spins(hh_cphf* h, int nspin) {
  HH_REQUIRE(h != nullptr, "handle is null");
  HH_REQUIRE(h->impl->nspin() == nspin, nspin == 1 ? "this entry point is for the restricted driver (hh_cphf_create)"
                                                    : "this entry point is for the unrestricted driver (hh_ucphf_create)");
}

template <typename T>
void need_arrays(T* const* v, int nspin) {
  HH_REQUIRE(v != nullptr, "null spin array");
  for (int t = 0; t < nspin; ++t) HH_REQUIRE(v[t] != nullptr, "null spin array entry");
}

hh_status solve_common(hh_cphf* h, int nspin, const double* const* rhs, int64_t nprop, double* const* mo1,
                       hh_precision precision, const double* diag, hh_krylov_stats* stats) {
  CphfStats st;
  try {
    need_spins(h, nspin);
    need_arrays(rhs, nspin);
    need_arrays(mo1, nspin);
    HH_REQUIRE(nprop > 0, "nprop <= 0");
    check_prec(precision);
    h->impl->solve_dimless(rhs, nprop, mo1, (Prec)precision, &st, diag);
    fill_stats(stats, st.krylov, st.refinements, st.residual, st.fallbacks);
    return HH_OK;
  } catch (const hh::Error& e) {
    fill_stats(stats, st.krylov, st.refinements, st.residual, st.fallbacks);
    hh::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    hh::set_last_error(e.what());
    return HH_ERR_INTERNAL;
  }
}
}  // namespace

extern "C" {


hh_status hh_cphf_dimless_rhs(hh_cphf* h, const double* f1mo, const double* s1mo, int64_t nprop, double* rhs) {
  HH_ABI_BEGIN
  need_spins(h, 1);
  HH_REQUIRE(f1mo && s1mo && rhs && nprop > 0, "null argument or nprop <= 0");
  h->impl->dimless_rhs(&f1mo, &s1mo, nprop, &rhs);
  HH_ABI_END
}

hh_status hh_cphf_response_mo(hh_cphf* h, const double* mo1, int64_t nprop, double* resp, hh_precision precision) {
  HH_ABI_BEGIN
  need_spins(h, 1);
  HH_REQUIRE(mo1 && resp && nprop > 0, "null argument or nprop <= 0");
  check_prec(precision);
  h->impl->response_mo(&mo1, nprop, &resp, (Prec)precision);
  HH_ABI_END
}

hh_status hh_cphf_response_dimless(hh_cphf* h, const double* mo1, int64_t nprop, double* resp,
                                   hh_precision precision) {
  HH_ABI_BEGIN
  need_spins(h, 1);
  HH_REQUIRE(mo1 && resp && nprop > 0, "null argument or nprop <= 0");
  check_prec(precision);
  h->impl->response_dimless(&mo1, nprop, &resp, (Prec)precision);
  HH_ABI_END
}

hh_status hh_cphf_solve_dimless(hh_cphf* h, const double* rhs, int64_t nprop, double* mo1, hh_precision precision,
                                const double* diag, hh_krylov_stats* stats) {
  return solve_common(h, 1, &rhs, nprop, &mo1, precision, diag, stats);
}

hh_status hh_cphf_finalize(hh_cphf* h, const double* f1mo, const double* s1mo, const double* mo1, int64_t nprop,
                           double* mo1_out, double* mo_e1) {
  HH_ABI_BEGIN
  need_spins(h, 1);
  HH_REQUIRE(f1mo && s1mo && mo1 && mo1_out && mo_e1 && nprop > 0, "null argument or nprop <= 0");
  h->impl->finalize(&f1mo, &s1mo, &mo1, nprop, &mo1_out, &mo_e1);
  HH_ABI_END
}

hh_status hh_cphf_hess(hh_cphf* h, const double* f1mo, const double* s1mo, const double* mo1, const double* mo_e1,
                       int64_t natm, double* de) {
  HH_ABI_BEGIN
  need_spins(h, 1);
  HH_REQUIRE(f1mo && s1mo && mo1 && mo_e1 && de && natm > 0, "null argument or natm <= 0");
  h->impl->cphf_hess(&f1mo, &s1mo, &mo1, &mo_e1, natm, de);
  HH_ABI_END
}

hh_status hh_ucphf_dimless_rhs(hh_cphf* h, const double* const* f1mo, const double* const* s1mo, int64_t nprop,
                               double* const* rhs) {
  HH_ABI_BEGIN
  need_spins(h, 2);
  need_arrays(f1mo, 2); need_arrays(s1mo, 2); need_arrays(rhs, 2);
  HH_REQUIRE(nprop > 0, "nprop <= 0");
  h->impl->dimless_rhs(f1mo, s1mo, nprop, rhs);
  HH_ABI_END
}

hh_status hh_ucphf_response_mo(hh_cphf* h, const double* const* mo1, int64_t nprop, double* const* resp,
                               hh_precision precision) {
  HH_ABI_BEGIN
  need_spins(h, 2);
  need_arrays(mo1, 2); need_arrays(resp, 2);
  HH_REQUIRE(nprop > 0, "nprop <= 0");
  check_prec(precision);
  h->impl->response_mo(mo1, nprop, resp, (Prec)precision);
  HH_ABI_END
}

hh_status hh_ucphf_response_dimless(hh_cphf* h, const double* const* mo1, int64_t nprop, double* const* resp,
                                    hh_precision precision) {
  HH_ABI_BEGIN
  need_spins(h, 2);
  need_arrays(mo1, 2); need_arrays(resp, 2);
  HH_REQUIRE(nprop > 0, "nprop <= 0");
  check_prec(precision);
  h->impl->response_dimless(mo1, nprop, resp, (Prec)precision);
  HH_ABI_END
}

hh_status hh_ucphf_solve_dimless(hh_cphf* h, const double* const* rhs, int64_t nprop, double* const* mo1,
                                 hh_precision precision, hh_krylov_stats* stats) {
  return solve_common(h, 2, rhs, nprop, mo1, precision, nullptr, stats);
}

hh_status hh_ucphf_finalize(hh_cphf* h, const double* const* f1mo, const double* const* s1mo,
                            const double* const* mo1, int64_t nprop, double* const* mo1_out, double* const* mo_e1) {
  HH_ABI_BEGIN
  need_spins(h, 2);
  need_arrays(f1mo, 2); need_arrays(s1mo, 2); need_arrays(mo1, 2); need_arrays(mo1_out, 2); need_arrays(mo_e1, 2);
  HH_REQUIRE(nprop > 0, "nprop <= 0");
  h->impl->finalize(f1mo, s1mo, mo1, nprop, mo1_out, mo_e1);
  HH_ABI_END
}

hh_status hh_ucphf_hess(hh_cphf* h, const double* const* f1mo, const double* const* s1mo, const double* const* mo1,
                        const double* const* mo_e1, int64_t natm, double* de) {
  HH_ABI_BEGIN
  need_spins(h, 2);
  need_arrays(f1mo, 2); need_arrays(s1mo, 2); need_arrays(mo1, 2); need_arrays(mo_e1, 2);
  HH_REQUIRE(de && natm > 0, "null argument or natm <= 0");
  h->impl->cphf_hess(f1mo, s1mo, mo1, mo_e1, natm, de);
  HH_ABI_END
}

hh_status hh_mlp_create(uint64_t seed, void* stream, hh_mlp** out) {
  HH_ABI_BEGIN
  HH_REQUIRE(out != nullptr, "out is null");
  *out = nullptr;
  auto h = std::make_unique<hh_mlp>();
  h->impl = std::make_unique<DiagMlp>(seed, (hipStream_t)stream);
  *out = h.release();
  HH_ABI_END
}

hh_status hh_mlp_destroy(hh_mlp* h) {
  HH_ABI_BEGIN
  if (h) {
    HH_HIP(hipStreamSynchronize(h->impl->stream()));
    delete h;
  }
  HH_ABI_END
}

hh_status hh_mlp_forward(hh_mlp* h, const float* x, int64_t n, float* y) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && n >= 0 && (n == 0 || (x && y)), "invalid argument");
  h->impl->forward(x, n, y);
  HH_ABI_END
}

hh_status hh_mlp_loss_and_grad(hh_mlp* h, const float* x, const float* t, const float* w, int64_t n, float* grad,
                               double* loss) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && x && t && w && grad && loss && n > 0, "invalid argument");
  *loss = h->impl->loss_and_grad(x, t, w, n, grad);
  HH_HIP(hipStreamSynchronize(h->impl->stream()));
  HH_ABI_END
}

hh_status hh_mlp_train(hh_mlp* h, const float* x, const float* t, const float* w, int64_t n, int32_t epochs, float lr,
                       double* loss) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && x && t && w && n > 0 && epochs >= 0, "invalid argument");
  const double l = h->impl->train(x, t, w, n, epochs, lr);
  if (loss) *loss = l;
  HH_ABI_END
}

hh_status hh_mlp_get_params(const hh_mlp* h, float* host_out) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && host_out, "invalid argument");
  h->impl->get_params(host_out);
  HH_ABI_END
}

hh_status hh_mlp_set_params(hh_mlp* h, const float* host_in) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && host_in, "invalid argument");
  h->impl->set_params(host_in);
  HH_ABI_END
}

hh_status hh_rijk_pair_norms(hh_rijk* h, int32_t s, const double* mo_coeff, int64_t nmo, double* pair) {
  HH_ABI_BEGIN
  HH_REQUIRE(h != nullptr, "handle is null");
  h->impl->pair_norms(s, mo_coeff, nmo, pair);
  HH_ABI_END
}

hh_status hh_cphf_precond_features(hh_cphf* h, const double* pair, const double* r, int64_t nprop, float* x) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && pair && r && x && nprop > 0, "invalid argument");
  h->impl->precond_features(pair, r, nprop, x);
  HH_ABI_END
}

hh_status hh_cphf_precond_target(hh_cphf* h, const double* r, const double* xs, int64_t nprop, float* t, float* w) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && r && xs && t && w && nprop > 0, "invalid argument");
  h->impl->precond_target(r, xs, nprop, t, w);
  HH_ABI_END
}

hh_status hh_cphf_precond_diag(hh_cphf* h, const float* y, double* diag) {
  HH_ABI_BEGIN
  HH_REQUIRE(h && y && diag, "invalid argument");
  h->impl->precond_diag(y, diag);
  HH_ABI_END
}

hh_status hh_linalg_solve(double* H, int64_t nd, double* G, int64_t nrhs, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(nd >= 0 && nrhs >= 0 && (nd == 0 || (H && G)), "invalid argument");
  DevBuf ws;
  lu_solve(H, nd, G, nrhs, (hipStream_t)stream, ws);
  HH_HIP(hipStreamSynchronize((hipStream_t)stream));
  HH_ABI_END
}

hh_status hh_linalg_solve_noalloc(double* H, int64_t nd, double* G, int64_t nrhs, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(nd >= 0 && nrhs >= 0 && (nd == 0 || (H && G)), "invalid argument");
  lu_solve_noalloc(H, nd, G, nrhs, (hipStream_t)stream);
  HH_HIP(hipStreamSynchronize((hipStream_t)stream));
  HH_ABI_END
}

hh_status hh_linalg_eigh(double* H, int64_t nd, double* w, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(nd >= 0 && (nd == 0 || (H && w)), "invalid argument");
  if (nd == 0) return HH_OK;
  DevBuf ws;
  eigh(H, nd, w, (hipStream_t)stream, ws);
  HH_ABI_END
}

hh_status hh_linalg_cholesky(double* H, int64_t nd, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(nd >= 0 && (nd == 0 || H), "invalid argument");
  DevBuf ws;
  const bool ok = cholesky(H, nd, (hipStream_t)stream, ws);
  HH_HIP(hipStreamSynchronize((hipStream_t)stream));
  HH_REQUIRE(ok, "cholesky: matrix is not positive definite");
  HH_ABI_END
}

hh_status hh_linalg_trsm_lower(const double* L, int64_t nd, double* X, int64_t n2, int64_t ldx, int32_t side_right,
                               void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(nd >= 0 && n2 >= 0 && ldx >= (side_right ? n2 : nd), "invalid sizes");
  DevBuf ws;
  trsm_lower(L, nd, X, n2, ldx, side_right != 0, (hipStream_t)stream, ws);
  HH_HIP(hipStreamSynchronize((hipStream_t)stream));
  HH_ABI_END
}

hh_status hh_util_gemm_nt(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                          const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int32_t config,
                          int32_t ksplit, void* stream) {
  HH_ABI_BEGIN
  static thread_local DevBuf* ws = new DevBuf;
  NtArgs a;
  a.M = M; a.N = N; a.K = K;
  a.A = A; a.lda = lda; a.B = B; a.ldb = ldb;
  a.C = C; a.ldc = ldc; a.alpha = alpha; a.beta = beta;
  a.ksplit = ksplit;
  if (config < 0) gemm_nt(a, NT_F64, (hipStream_t)stream, *ws);
  else gemm_nt_f64_config(a, config, (hipStream_t)stream, *ws);
  HH_ABI_END
}

hh_status hh_util_gemm_nt_bf16x3(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                                 const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int32_t config,
                                 int32_t ksplit, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(M >= 0 && N >= 0 && K >= 0 && K % 32 == 0 && lda >= K && ldb >= K, "gemm_nt_bf16x3: bad sizes");
  HH_REQUIRE(A && B && C, "null operand");
  static thread_local DevBuf* ws = new DevBuf;
  static thread_local DevBuf* sp = new DevBuf;
  hipStream_t s = (hipStream_t)stream;
  unsigned short* ah = static_cast<unsigned short*>(sp->reserve(2 * sizeof(unsigned short) * K * (M + N) + 64));
  unsigned short* al = ah + K * M;
  unsigned short* bh = al + K * M;
  unsigned short* bl = bh + K * N;
  split_bf16_tiled<double>(A, lda, K, M, K, ah, al, s);
  split_bf16_tiled<double>(B, ldb, K, N, K, bh, bl, s);
  NtArgs a;
  a.M = M; a.N = N; a.K = K;
  a.A = ah; a.A_lo = al; a.lda = K; a.B = bh; a.B_lo = bl; a.ldb = K;
  a.tiled = true;
  a.C = C; a.ldc = ldc; a.alpha = alpha; a.beta = beta;
  a.ksplit = ksplit;
  if (config < 0) gemm_nt(a, NT_BF16X3, s, *ws);
  else gemm_nt_bf16x3_config(a, config, s, *ws);
  HH_ABI_END
}

hh_status hh_util_gemm_nt_bf16x3_choice(int64_t M, int64_t N, int64_t K, int32_t* config, int32_t* ksplit) {
  HH_ABI_BEGIN
  HH_REQUIRE(config && ksplit, "null output");
  int c = -1, k = 0;
  if (!gemm_nt_choice(M, N, K, &c, &k, NT_BF16X3)) c = -1;
  *config = c;
  *ksplit = k;
  HH_ABI_END
}

hh_status hh_util_mfma_peak(int32_t kind, int32_t iters, double* tflops, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(tflops != nullptr, "null output");
  *tflops = mfma_peak_tflops(kind, iters, (hipStream_t)stream);
  HH_ABI_END
}

hh_status hh_util_set_autotune(int32_t enable) {
  HH_ABI_BEGIN
  gemm_nt_set_autotune(enable != 0);
  HH_ABI_END
}

hh_status hh_util_gemm_nt_choice(int64_t M, int64_t N, int64_t K, int32_t* config, int32_t* ksplit) {
  HH_ABI_BEGIN
  HH_REQUIRE(config && ksplit, "null output");
  int c = -1, k = 0;
  if (!gemm_nt_choice(M, N, K, &c, &k)) c = -1;
  *config = c;
  *ksplit = k;
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

hh_status hh_vib_harmonic_analysis(const double* hess, const double* geom, const double* mass, int64_t natm,
                                   int32_t project_trans, int32_t project_rot, const hh_vib_out* out) {
  HH_ABI_BEGIN
  HH_REQUIRE(hess && geom && mass && out && natm >= 1 && natm <= 100000, "invalid argument");
  for (int64_t a = 0; a < natm; ++a) HH_REQUIRE(mass[a] > 0.0, "masses must be positive");
  const vib::Result r = vib::harmonic_analysis(hess, geom, mass, (int)natm, project_trans != 0, project_rot != 0);
  auto put = [&](auto* dst, const auto& src) {
    if (dst)
      for (size_t i = 0; i < src.size(); ++i) dst[i] = src[i];
  };
  put(out->omega, r.omega);
  put(out->imag, r.imag);
  put(out->q, r.q);
  put(out->w, r.w);
  put(out->x, r.x);
  put(out->degeneracy, r.degeneracy);
  put(out->trv, r.trv);
  put(out->mu, r.mu);
  put(out->k, r.k);
  put(out->dq0, r.dq0);
  put(out->qtp0, r.qtp0);
  put(out->xtp0, r.xtp0);
  put(out->theta_vib, r.theta_vib);
  HH_ABI_END
}

hh_status hh_vib_rotation_const(const double* mass, const double* geom_c, int64_t natm, int32_t unit,
                                double* out3) {
  HH_ABI_BEGIN
  HH_REQUIRE(mass && geom_c && out3 && natm >= 1 && (unit == 0 || unit == 1), "invalid argument");
  vib::rotation_const(mass, geom_c, (int)natm, unit, out3);
  HH_ABI_END
}

hh_status hh_vib_rotor_type(const double* rc_ghz, int32_t* rotor) {
  HH_ABI_BEGIN
  HH_REQUIRE(rc_ghz && rotor, "invalid argument");
  *rotor = vib::rotor_type(rc_ghz);
  HH_ABI_END
}

hh_status hh_vib_tr_space(const double* mass, const double* geom, int64_t natm, int32_t space, double* tr,
                          int64_t* ncol) {
  HH_ABI_BEGIN
  HH_REQUIRE(mass && geom && tr && ncol && natm >= 1 && space >= 0 && space <= 3, "invalid argument");
  int nc = 0;
  const std::vector<double> u = vib::tr_space(mass, geom, (int)natm, space, -1, &nc);
  std::copy(u.begin(), u.end(), tr);
  *ncol = nc;
  HH_ABI_END
}

hh_status hh_util_copy(void* dst, const void* src, int64_t bytes, void* stream) {
  HH_ABI_BEGIN
  HH_REQUIRE(bytes >= 0 && (bytes == 0 || (dst && src)), "invalid argument");
  copy_bytes(dst, src, bytes, (hipStream_t)stream);
  HH_ABI_END
}

}  // extern "C"
