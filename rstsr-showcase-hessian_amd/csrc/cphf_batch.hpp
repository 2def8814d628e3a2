// Many CPHF systems of uneven size solved together: one block Krylov per system (the reference's
// krylov_block, src/util/krylov_block.rs:39-253, on each system's dimensionless CPHF equations,
// src/hessian/rscf.rs:278-300 / uscf.rs:278-349), all systems advanced in lockstep so that each
// step of a Krylov cycle is one grouped launch over every system still running.
#pragma once
#include <vector>

#include "cphf.hpp"
#include "rijk.hpp"

namespace hh {

// One grouped FP64 product C = alpha f(sum_k A[m, k] B[k, q]) + beta D[m, q] of the batch.
// A[m, k] at A[m lda + k] (k-major instances) or A[m + k lda]; columns q of B and C address
// two-level strides: B[k, q] at B[k + (q % cib) ldb + (q / cib) ldb2], C[m, q] (and D) at
// C[m + (q % cic) ldc + (q / cic) ldc2].  f: optional division of row m by rowdiv[m], optional
// CPHF dimensionless scaling (rows m < nocc zero, others divided by evir[m - nocc] - eocc[q % cic]),
// optional diagonal addend adddiag[m] where m == q.  ksplit > 1: the K range is cut into chunks
// of kchunk whose FP64 partial sums are reduced in a fixed order by a second grouped launch.
struct GProb {
  int64_t M = 0, N = 0, K = 0;
  const double* A = nullptr;
  int64_t lda = 0;
  const double* B = nullptr;
  int64_t ldb = 0, ldb2 = 0, cib = 1LL << 40;
  double* C = nullptr;
  int64_t ldc = 0, ldc2 = 0, cic = 1LL << 40;
  const double* D = nullptr;
  double alpha = 1.0, beta = 0.0;
  const double* rowdiv = nullptr;
  const double* eocc = nullptr;
  const double* evir = nullptr;
  int64_t nocc = 0;
  const double* adddiag = nullptr;
  // filled by the launcher
  int64_t tiles_m = 0, tiles_n = 0, ksplit = 1, kchunk = 0, item0 = 0, ritem0 = 0;
  double* partial = nullptr;
};

struct NtProb;

class CphfBatch {
 public:
  CphfBatch(double tol, int max_cycle, double lindep, hipStream_t s);
  ~CphfBatch();
  // A system: the CPHF driver's orbitals (mo_coeff, occupations, orbital energies) and the RI-JK
  // handle whose resident response matrix is the system's whole response operator (the handle
  // prepared with the driver's orbitals).  rhs[s], mo1[s]: device [nmo_s, nocc_s, nprop] per
  // spin, as hh_cphf_solve_dimless / hh_ucphf_solve_dimless.  nprop <= 128.
  void add(Cphf* drv, Rijk* rijk, const double* const* rhs, int64_t nprop, double* const* mo1);
  // Solve every system (FP64).  stats[i] per system; throws HH_ERR_NOT_CONVERGED after writing
  // every solution if a system did not converge within max_cycle.
  void solve(std::vector<KrylovStats>* stats);

  int64_t launches_per_cycle() const { return launches_per_cycle_; }
  // wall seconds of the last solve: resident matrix builds, set-up and first block, lockstep
  // cycles, projected solves
  const double* phase_seconds() const { return phase_s_; }
  int64_t cycles() const { return lockstep_cycles_; }
  size_t size() const { return sys_.size(); }

 private:
  struct Sys {
    int nspin = 1;
    int64_t nao = 0, nmo[2] = {0, 0}, nocc[2] = {0, 0}, off[2] = {0, 0}, roff[2] = {0, 0};
    int64_t n = 0, R = 0, kpad = 0, nset = 0;
    const double* C[2] = {nullptr, nullptr};
    const double* eocc[2] = {nullptr, nullptr};
    const double* evir[2] = {nullptr, nullptr};
    const double* m64 = nullptr;
    Rijk* rijk = nullptr;
    double* mo1[2] = {nullptr, nullptr};
    // device state
    DevBuf b, xs, ax, ubra, rbra, xnew, w, coeffs, gram, tm, inner, d1, hbuf, gbuf, lu, luws;
    int64_t cap = 0;  // columns of xs / ax
    // host state
    int64_t nd = 0, nb = 0;
    int cycles = 0;
    bool done = false, converged = false;
    double max_ip = 0.0;
  };
  void grow(Sys& s, int64_t need);
  // the three products of the response operator for the running systems: ax = A x on columns
  // [nd, nd + nb) of xs / ax
  void add_operator(std::vector<GProb>* g1, std::vector<NtProb>* g2, std::vector<GProb>* g3, Sys& s);
  // enqueue a grouped product (instance chosen by a_kmajor); returns the launches made
  int launch(std::vector<GProb>& probs, bool a_kmajor);
  // the matrix products of the operator, one grouped NT launch
  int launch_nt(std::vector<NtProb>& probs);
  // two Gram / LDL passes over xnew [n, nb] of every listed system; the second pass selects
  // pivots with d > sel and writes the selected vectors to xs[:, nd_dst), their norms to
  // inner[nd_dst), and per system (max d, count) to resd
  int orth(const std::vector<Sys*>& ss, const std::vector<int64_t>& nd_dst, double sel);

  double tol_, lindep_;
  int max_cycle_;
  hipStream_t stream_;
  std::vector<Sys*> sys_;
  DevBuf desc_, partial_, resd_, ws_;
  void* pinned_ = nullptr;  // this cycle's descriptors (host side), copied to desc_ at the same offsets
  size_t pinned_bytes_ = 0, arena_off_ = 0;
  int64_t launches_per_cycle_ = 0, lockstep_cycles_ = 0;
  std::vector<hipStream_t> lu_streams_;  // the projected solves' streams (default priority)
  std::vector<hipEvent_t> lu_events_;
  hipEvent_t lu_fork_ = nullptr;
  double phase_s_[4] = {0, 0, 0, 0};
};

}  // namespace hh
