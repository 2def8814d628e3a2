// Many CPHF systems of uneven size solved in lockstep (see cphf_batch.hpp).
//
// Every system keeps the reference's krylov_block semantics (krylov_block.rs:39-197): the
// non-normalised basis, orthogonalisation of each new block against the basis (projection
// coefficients scaled by the stored squared norms) and within itself with the drop rule
// ||v||^2 <= lindep, trial vectors kept only while ||v||^2 > max(lindep, tol^2), convergence when
// every new vector falls below it, and the Galerkin projected solve at the end
// (krylov_block.rs:204-215).  One subspace per system: the batch does not restart (the reference
// with max_space >= max_cycle).  The systems differ in size (rows n, columns nset, basis size),
// so each step of a cycle is one grouped launch over a list of (system, tile) work items:
//   1. ubra_s = C_s x_s            (per spin; x the trial block, columns (i, column))
//   2. rbra   = M ubra             (M the system's resident RI-JK response matrix, [R][kpad])
//   3. ax_s   = scale(C_s^T rbra_s) (the dimensionless CPHF scaling in the epilogue:
//                                    occupied rows zero, virtual rows / (e_a - e_i), rscf.rs:237-263)
//   4. coefficients xs^T ax / inner (split-K, fixed-order FP64 reduction)
//   5. xnew = ax - xs coefficients
//   6. two Gram / LDL^T passes (Gram products, one LDL block per system, V T products); the
//      second pass selects the vectors above the convergence threshold and writes them straight
//      into the next basis slot, with their squared norms.
// The host reads one small array per cycle (per system: largest squared norm, vectors kept) for
// all systems together; everything else stays on the device.
#include "cphf_batch.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>

#include "elementwise.hpp"
#include "krylov_dev.hpp"
#include "linalg.hpp"

namespace hh {

struct NtProb {
  int64_t M, N, K;
  const double* A;
  int64_t lda;
  const double* B;
  int64_t ldb;
  double* C;
  int64_t ldc;
  int64_t tiles_m, item0;
};

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int GBM = 64, GBN = 64, GBK = 16;
constexpr int64_t KSPLIT_CHUNK = 2048;     // K elements per split of a product with few tiles
constexpr int64_t KSPLIT_TILES = 512;      // products with fewer tiles than this are split along K
constexpr size_t ARENA_BYTES = 8u << 20;   // descriptors of one cycle (pinned host + device)
constexpr int64_t MAX_SYSTEMS = 1024;

unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 16384)); }

__device__ __forceinline__ int find_problem(const GProb* __restrict__ probs, int nprob, int64_t item, bool reduce) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t start = reduce ? probs[mid].ritem0 : probs[mid].item0;
    if (start <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// the epilogue value of element (m, q) from the raw sum
__device__ __forceinline__ double epilogue(const GProb& p, int64_t m, int64_t q, int64_t coff, double acc) {
  double v = acc;
  if (p.rowdiv) v /= p.rowdiv[m];
  if (p.eocc) v = (m < p.nocc) ? 0.0 : v / (p.evir[m - p.nocc] - p.eocc[q % p.cic]);
  v *= p.alpha;
  if (p.adddiag && m == q) v += p.adddiag[m];
  if (p.beta != 0.0) v += p.beta * (p.D ? p.D : p.C)[coff];
  return v;
}

// One 64 x 64 tile of one problem over one K chunk.  Layout and MFMA issue order as the generic
// FP64 GEMM (gemm.hip): waves in a 2 x 2 grid of 32 x 32, v_mfma_f64_16x16x4f64 issued as
// (B fragment, A fragment) so each lane accumulates C^T blocks and column-major C is stored by
// 16 consecutive lanes.
template <bool AK>
__global__ __launch_bounds__(256) void ggemm_kernel(const GProb* __restrict__ probs, int nprob) {
  constexpr int BM = GBM, BN = GBN, BK = GBK;
  constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
  constexpr int NA = BM * BK / 256, NB = BN * BK / 256;
  constexpr int LDA = BM + 4, LDB = BN + 4;
  __shared__ double As[BK * LDA];
  __shared__ double Bs[BK * LDB];
  const int64_t item = blockIdx.x;
  const GProb p = probs[find_problem(probs, nprob, item, false)];
  const int64_t r = item - p.item0;
  const int64_t tiles = p.tiles_m * p.tiles_n;
  const int64_t tile = r % tiles, kz = r / tiles;
  const int64_t m0 = (tile % p.tiles_m) * BM, n0 = (tile / p.tiles_m) * BN;
  const int64_t kbeg = kz * p.kchunk, kend = min(p.K, kbeg + p.kchunk);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  // per-thread B columns (fixed over the K loop)
  int64_t boff[NB];
  bool bok[NB];
#pragma unroll
  for (int e = 0; e < NB; ++e) {
    const int64_t q = n0 + (tid + 256 * e) / BK;
    bok[e] = q < p.N;
    boff[e] = bok[e] ? (q % p.cib) * p.ldb + (q / p.cib) * p.ldb2 : 0;
  }
  v4d acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v4d{0, 0, 0, 0};
  double ra[NA], rb[NB];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      const int idx = tid + 256 * e;
      const int mm = AK ? idx / BK : idx % BM;
      const int kk = AK ? idx % BK : idx / BM;
      const int64_t gm = m0 + mm, gk = k0 + kk;
      ra[e] = (gm < p.M && gk < kend) ? (AK ? p.A[gm * p.lda + gk] : p.A[gm + gk * p.lda]) : 0.0;
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int64_t gk = k0 + (tid + 256 * e) % BK;
      rb[e] = (bok[e] && gk < kend) ? p.B[gk + boff[e]] : 0.0;
    }
  };
  if (kbeg < kend) {
    load(kbeg);
    for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
      __syncthreads();
#pragma unroll
      for (int e = 0; e < NA; ++e) {
        const int idx = tid + 256 * e;
        const int mm = AK ? idx / BK : idx % BM;
        const int kk = AK ? idx % BK : idx / BM;
        As[kk * LDA + mm] = ra[e];
      }
#pragma unroll
      for (int e = 0; e < NB; ++e) {
        const int idx = tid + 256 * e;
        Bs[(idx % BK) * LDB + idx / BK] = rb[e];
      }
      __syncthreads();
      if (k0 + BK < kend) load(k0 + BK);
#pragma unroll
      for (int ks = 0; ks < BK / 4; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
        double af[FM], bfr[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i] = As[kr * LDA + wm * WM + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < FN; ++j) bfr[j] = Bs[kr * LDB + wn * WN + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bfr[j], af[i], acc[i][j], 0, 0, 0);
      }
    }
  }
  // acc[i][j][r] = C[mrow0 + 16 i][ncol0 + 16 j + 4 r]
  const int64_t mrow0 = m0 + wm * WM + (lane & 15);
  const int64_t ncol0 = n0 + wn * WN + (lane >> 4);
  if (p.ksplit > 1) {
    double* P = p.partial + kz * p.M * p.N;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int64_t gm = mrow0 + i * 16, gn = ncol0 + j * 16 + 4 * rr;
          if (gm < p.M && gn < p.N) P[gm + gn * p.M] = acc[i][j][rr];
        }
  } else {
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int64_t gn = ncol0 + j * 16 + 4 * rr;
        if (gn >= p.N) continue;
        const int64_t cbase = (gn % p.cic) * p.ldc + (gn / p.cic) * p.ldc2;
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          const int64_t gm = mrow0 + i * 16;
          if (gm < p.M) p.C[gm + cbase] = epilogue(p, gm, gn, gm + cbase, acc[i][j][rr]);
        }
      }
  }
}

// The response matrix products rbra = M ubra of every system in one launch: C[m, n] =
// sum_k A[m lda + k] B[n ldb + k] (both operands K-contiguous, K a multiple of 16, rows 16-byte
// aligned), 208 x 128 tiles (8 waves of 13 x 1 MFMA blocks) so that a trial block of up to 128
// columns reads each row of M once; 16-byte loads, double-buffered LDS with one barrier per K
// step (the FP64 NT product of gemm_nt.hip, with the problem looked up per workgroup).


constexpr int NT_BM = 208, NT_BN = 128, NT_BK = 16;

__global__ __launch_bounds__(512) void gnt_f64_kernel(const NtProb* __restrict__ probs, int nprob) {
  constexpr int FM = 13, BM = NT_BM, BN = NT_BN, BK = NT_BK;
  constexpr int LD = BK + 2;  // doubles per LDS row (144 B)
  constexpr int A_ELEMS = BM * LD, B_ELEMS = BN * LD;
  constexpr int STAGE = A_ELEMS + B_ELEMS;
  constexpr int A_CH = BM * (BK / 2), B_CH = BN * (BK / 2);  // 16-byte chunks
  constexpr int A_PT = (A_CH + 511) / 512, B_PT = (B_CH + 511) / 512;
  __shared__ __attribute__((aligned(16))) double smem[2 * STAGE];
  const int64_t item = blockIdx.x;
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (probs[mid].item0 <= item) lo = mid;
    else hi = mid - 1;
  }
  const NtProb p = probs[lo];
  const int tid = threadIdx.x, lane = tid & 63, wn = tid >> 6;
  const int64_t m0 = (item - p.item0) * BM, n0 = 0;
  v4d acc[FM];
#pragma unroll
  for (int i = 0; i < FM; ++i) acc[i] = v4d{0, 0, 0, 0};
  double2 ra[A_PT], rb[B_PT];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < A_PT; ++e) {
      const int c = tid + 512 * e;
      double2 v = {0.0, 0.0};
      if (c < A_CH) {
        const int r = c / (BK / 2), kk = (c % (BK / 2)) * 2;
        const int64_t gm = m0 + r;
        if (gm < p.M) v = *reinterpret_cast<const double2*>(p.A + gm * p.lda + k0 + kk);
      }
      ra[e] = v;
    }
#pragma unroll
    for (int e = 0; e < B_PT; ++e) {
      const int c = tid + 512 * e;
      double2 v = {0.0, 0.0};
      if (c < B_CH) {
        const int r = c / (BK / 2), kk = (c % (BK / 2)) * 2;
        const int64_t gn = n0 + r;
        if (gn < p.N) v = *reinterpret_cast<const double2*>(p.B + gn * p.ldb + k0 + kk);
      }
      rb[e] = v;
    }
  };
  auto sstore = [&](int buf) {
    double* sm = smem + buf * STAGE;
#pragma unroll
    for (int e = 0; e < A_PT; ++e) {
      const int c = tid + 512 * e;
      if (c < A_CH) *reinterpret_cast<double2*>(sm + (c / (BK / 2)) * LD + (c % (BK / 2)) * 2) = ra[e];
    }
#pragma unroll
    for (int e = 0; e < B_PT; ++e) {
      const int c = tid + 512 * e;
      if (c < B_CH) *reinterpret_cast<double2*>(sm + A_ELEMS + (c / (BK / 2)) * LD + (c % (BK / 2)) * 2) = rb[e];
    }
  };
  auto compute = [&](int buf) {
    const double* sm = smem + buf * STAGE;
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kc = ks * 4 + (lane >> 4);
      const double bf = sm[A_ELEMS + (wn * 16 + (lane & 15)) * LD + kc];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const double af = sm[(i * 16 + (lane & 15)) * LD + kc];
        acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bf, acc[i], 0, 0, 0);
      }
    }
  };
  gload(0);
  sstore(0);
  __syncthreads();
  int buf = 0;
  for (int64_t k0 = 0; k0 < p.K; k0 += BK) {
    const bool more = k0 + BK < p.K;
    if (more) gload(k0 + BK);
    compute(buf);
    if (more) sstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  // f64 result layout D[(l/16) + 4r][l%16]
  const int64_t col = n0 + wn * 16 + (lane & 15);
  const int64_t row0 = m0 + (lane >> 4);
  if (col < p.N) {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gm = row0 + i * 16 + 4 * r;
        if (gm < p.M) p.C[gm + col * p.ldc] = acc[i][r];
      }
  }
}

// Split-K problems: C = epilogue(sum_z partial[z]) in a fixed order; 256 elements per item.
__global__ __launch_bounds__(256) void greduce_kernel(const GProb* __restrict__ probs, int nprob) {
  const int64_t item = blockIdx.x;
  const GProb p = probs[find_problem(probs, nprob, item, true)];
  const int64_t e = (item - p.ritem0) * 256 + threadIdx.x;
  const int64_t MN = p.M * p.N;
  if (p.ksplit <= 1 || e >= MN) return;
  double s = 0.0;
  for (int64_t z = 0; z < p.ksplit; ++z) s += p.partial[z * MN + e];
  const int64_t m = e % p.M, q = e / p.M;
  const int64_t coff = m + (q % p.cic) * p.ldc + (q / p.cic) * p.ldc2;
  p.C[coff] = epilogue(p, m, q, coff, s);
}

struct LdlArg {
  const double* G;
  int m;
  double sel;
  double* Tm;
  double* dout;
  double* resd;  // non-null: resd[0] = largest kept d, resd[1] = selected count
  int* kout;
};

__global__ __launch_bounds__(KR_LDL_T) void ldl_batch_kernel(const LdlArg* __restrict__ args, double lindep) {
  const LdlArg a = args[blockIdx.x];
  __shared__ int k_s;
  __shared__ double mx_s;
  ldl_drop_block(a.G, a.m, lindep, a.sel, a.Tm, a.dout, &k_s, &mx_s);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (a.kout) *a.kout = k_s;
    if (a.resd) {
      a.resd[0] = mx_s;
      a.resd[1] = (double)k_s;
    }
  }
}

// dst[r + ld_dst c] = src[r + ld_src c]
__global__ void copy_cols_kernel(const double* __restrict__ src, int64_t ld_src, double* __restrict__ dst,
                                 int64_t ld_dst, int64_t rows, int64_t ncol) {
  const int64_t total = rows * ncol;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x)
    dst[e % rows + ld_dst * (e / rows)] = src[e % rows + ld_src * (e / rows)];
}

}  // namespace

CphfBatch::CphfBatch(double tol, int max_cycle, double lindep, hipStream_t s)
    : tol_(tol), lindep_(lindep), max_cycle_(max_cycle), stream_(s) {
  HH_REQUIRE(max_cycle >= 0 && tol >= 0.0, "cphf batch: max_cycle and tol must be non-negative");
}

CphfBatch::~CphfBatch() {
  for (Sys* s : sys_) delete s;
  if (pinned_) (void)hipHostFree(pinned_);
  for (hipStream_t st : lu_streams_) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : lu_events_) (void)hipEventDestroy(ev);
  if (lu_fork_) (void)hipEventDestroy(lu_fork_);
}

void CphfBatch::add(Cphf* drv, Rijk* rijk, const double* const* rhs, int64_t nprop, double* const* mo1) {
  HH_REQUIRE(drv && rijk && rhs && mo1, "cphf batch: null argument");
  HH_REQUIRE((int64_t)sys_.size() < MAX_SYSTEMS, "cphf batch: at most 1024 systems");
  HH_REQUIRE(nprop >= 1 && nprop <= KR_GB, "cphf batch: 1 <= nprop <= 128");
  HH_REQUIRE(rijk->nspin() == drv->nspin(), "cphf batch: RI-JK spin count differs from the driver's");
  HH_REQUIRE(rijk->prepared(), "cphf batch: RI-JK handle not prepared (make_response_preparation)");
  HH_REQUIRE(rijk->nao() == drv->nao(), "cphf batch: RI-JK nao differs from the driver's");
  HH_REQUIRE(drv->config().level_shift == 0.0, "cphf batch: level shift not supported");
  auto* s = new Sys;
  s->nspin = drv->nspin();
  s->nao = drv->nao();
  s->nset = nprop;
  s->rijk = rijk;
  int64_t roff = 0;
  for (int t = 0; t < s->nspin; ++t) {
    HH_REQUIRE(rhs[t] && mo1[t], "cphf batch: null spin array");
    HH_REQUIRE(rijk->nocc(t) == drv->nocc(t), "cphf batch: RI-JK occupation differs from the driver's");
    s->nmo[t] = drv->nmo(t);
    s->nocc[t] = drv->nocc(t);
    s->off[t] = drv->spin_offset(t);
    s->roff[t] = roff;
    roff += s->nao * s->nocc[t];
    s->C[t] = drv->mo_coeff(t);
    s->eocc[t] = drv->eocc_dev(t);
    s->evir[t] = drv->evir_dev(t);
    s->mo1[t] = mo1[t];
  }
  s->n = drv->stacked_rows();
  s->R = roff;
  // the stacked right-hand side [n, nset]
  double* b = static_cast<double*>(s->b.reserve(sizeof(double) * s->n * nprop));
  for (int t = 0; t < s->nspin; ++t) {
    const int64_t rows = s->nmo[t] * s->nocc[t];
    // column c of spin t: rhs[t] + rows c -> b + off_t + n c
    hipLaunchKernelGGL(copy_cols_kernel, dim3(nblk(rows * nprop)), dim3(256), 0, stream_, rhs[t], rows,
                       b + s->off[t], s->n, rows, nprop);
  }
  HH_HIP(hipGetLastError());
  sys_.push_back(s);
}

void CphfBatch::grow(Sys& s, int64_t need) {
  if (need <= s.cap) return;
  const int64_t cap = std::max<int64_t>(need, 2 * s.cap);
  const size_t bytes = sizeof(double) * s.n * cap;
  DevBuf xs2, ax2;
  xs2.reserve(bytes);
  ax2.reserve(bytes);
  if (s.cap > 0) {
    copy_bytes(xs2.ptr, s.xs.ptr, (int64_t)sizeof(double) * s.n * s.cap, stream_);
    copy_bytes(ax2.ptr, s.ax.ptr, (int64_t)sizeof(double) * s.n * s.cap, stream_);
    HH_HIP(hipStreamSynchronize(stream_));
  }
  std::swap(s.xs.ptr, xs2.ptr);
  std::swap(s.xs.bytes, xs2.bytes);
  std::swap(s.ax.ptr, ax2.ptr);
  std::swap(s.ax.bytes, ax2.bytes);
  double* inner_old = s.inner.as<double>();
  DevBuf in2;
  in2.reserve(sizeof(double) * cap);
  if (s.cap > 0) {
    copy_bytes(in2.ptr, inner_old, (int64_t)sizeof(double) * s.cap, stream_);
    HH_HIP(hipStreamSynchronize(stream_));
  }
  std::swap(s.inner.ptr, in2.ptr);
  std::swap(s.inner.bytes, in2.bytes);
  s.cap = cap;
  s.coeffs.reserve(sizeof(double) * cap * s.nset);  // the projection coefficients of any cycle
  s.lu.reserve(sizeof(int64_t) * cap + 256);          // pivots and status of the projected solve
}

int CphfBatch::launch(std::vector<GProb>& probs, bool a_kmajor) {
  if (probs.empty()) return 0;
  int64_t items = 0, ritems = 0, pbytes = 0;
  for (GProb& p : probs) {
    HH_REQUIRE(p.M > 0 && p.N > 0 && p.K > 0, "cphf batch: empty product");
    p.tiles_m = ceil_div(p.M, GBM);
    p.tiles_n = ceil_div(p.N, GBN);
    const int64_t tiles = p.tiles_m * p.tiles_n;
    p.ksplit = (tiles < KSPLIT_TILES && p.K > KSPLIT_CHUNK) ? ceil_div(p.K, KSPLIT_CHUNK) : 1;
    p.kchunk = ceil_div(ceil_div(p.K, p.ksplit), GBK) * GBK;
    p.ksplit = ceil_div(p.K, p.kchunk);
    p.item0 = items;
    items += tiles * p.ksplit;
    p.ritem0 = ritems;
    if (p.ksplit > 1) {
      ritems += ceil_div(p.M * p.N, 256);
      pbytes += (int64_t)sizeof(double) * p.ksplit * p.M * p.N;
    }
  }
  HH_REQUIRE(items < (1LL << 31), "cphf batch: grid too large");
  if (pbytes > 0) {
    if ((size_t)pbytes > partial_.bytes) {
      HH_HIP(hipStreamSynchronize(stream_));  // earlier launches of this cycle may still read it
      partial_.reserve((size_t)pbytes);
    }
    int64_t off = 0;
    for (GProb& p : probs)
      if (p.ksplit > 1) {
        p.partial = reinterpret_cast<double*>(static_cast<char*>(partial_.ptr) + off);
        off += (int64_t)sizeof(double) * p.ksplit * p.M * p.N;
      }
  }
  // descriptors: appended to this cycle's pinned arena, copied to the same offset on the device
  const size_t bytes = sizeof(GProb) * probs.size();
  HH_REQUIRE(arena_off_ + bytes <= ARENA_BYTES, "cphf batch: too many products in one cycle");
  memcpy(static_cast<char*>(pinned_) + arena_off_, probs.data(), bytes);
  GProb* dp = reinterpret_cast<GProb*>(static_cast<char*>(desc_.ptr) + arena_off_);
  HH_HIP(hipMemcpyAsync(dp, static_cast<char*>(pinned_) + arena_off_, bytes, hipMemcpyHostToDevice, stream_));
  arena_off_ += (bytes + 255) / 256 * 256;
  const int np = (int)probs.size();
  if (a_kmajor) hipLaunchKernelGGL(ggemm_kernel<true>, dim3((unsigned)items), dim3(256), 0, stream_, dp, np);
  else hipLaunchKernelGGL(ggemm_kernel<false>, dim3((unsigned)items), dim3(256), 0, stream_, dp, np);
  HH_HIP(hipGetLastError());
  int n = 1;
  if (ritems > 0) {
    hipLaunchKernelGGL(greduce_kernel, dim3((unsigned)ritems), dim3(256), 0, stream_, dp, np);
    HH_HIP(hipGetLastError());
    ++n;
  }
  return n;
}

void CphfBatch::add_operator(std::vector<GProb>* g1, std::vector<NtProb>* g2, std::vector<GProb>* g3, Sys& s) {
  const double* x = s.xs.as<double>() + s.nd * s.n;
  double* ax = s.ax.as<double>() + s.nd * s.n;
  double* ubra = s.ubra.as<double>();
  double* rbra = s.rbra.as<double>();
  for (int t = 0; t < s.nspin; ++t) {
    GProb a;  // ubra_t[mu, (i, c)] = sum_m C_t[mu, m] x_t[m, (i, c)]
    a.M = s.nao; a.N = s.nocc[t] * s.nb; a.K = s.nmo[t];
    a.A = s.C[t]; a.lda = s.nao;
    a.B = x + s.off[t]; a.ldb = s.nmo[t]; a.cib = s.nocc[t]; a.ldb2 = s.n;
    a.C = ubra + s.roff[t]; a.ldc = s.nao; a.cic = s.nocc[t]; a.ldc2 = s.kpad;
    g1->push_back(a);
    GProb c;  // ax_t[m, (i, c)] = scale(sum_mu C_t[mu, m] rbra_t[mu, (i, c)])
    c.M = s.nmo[t]; c.N = s.nocc[t] * s.nb; c.K = s.nao;
    c.A = s.C[t]; c.lda = s.nao;
    c.B = rbra + s.roff[t]; c.ldb = s.nao; c.cib = s.nocc[t]; c.ldb2 = s.R;
    c.C = ax + s.off[t]; c.ldc = s.nmo[t]; c.cic = s.nocc[t]; c.ldc2 = s.n;
    c.eocc = s.eocc[t]; c.evir = s.evir[t]; c.nocc = s.nocc[t];
    g3->push_back(c);
  }
  NtProb m{};  // rbra[r, c] = sum_k M[r][k] ubra[k, c] (ubra rows R .. kpad are zero)
  m.M = s.R; m.N = s.nb; m.K = s.kpad;
  m.A = s.m64; m.lda = s.kpad;
  m.B = ubra; m.ldb = s.kpad;
  m.C = rbra; m.ldc = s.R;
  g2->push_back(m);
}

int CphfBatch::launch_nt(std::vector<NtProb>& probs) {
  if (probs.empty()) return 0;
  int64_t items = 0;
  for (NtProb& p : probs) {
    HH_REQUIRE(p.M > 0 && p.N > 0 && p.N <= NT_BN && p.K % NT_BK == 0 && p.lda % 2 == 0 && p.ldb % 2 == 0,
               "cphf batch: matrix product shape");
    p.tiles_m = ceil_div(p.M, NT_BM);
    p.item0 = items;
    items += p.tiles_m;
  }
  const size_t bytes = sizeof(NtProb) * probs.size();
  HH_REQUIRE(arena_off_ + bytes <= ARENA_BYTES, "cphf batch: too many products in one cycle");
  memcpy(static_cast<char*>(pinned_) + arena_off_, probs.data(), bytes);
  NtProb* dp = reinterpret_cast<NtProb*>(static_cast<char*>(desc_.ptr) + arena_off_);
  HH_HIP(hipMemcpyAsync(dp, static_cast<char*>(pinned_) + arena_off_, bytes, hipMemcpyHostToDevice, stream_));
  arena_off_ += (bytes + 255) / 256 * 256;
  hipLaunchKernelGGL(gnt_f64_kernel, dim3((unsigned)items), dim3(512), 0, stream_, dp, (int)probs.size());
  HH_HIP(hipGetLastError());
  return 1;
}

int CphfBatch::orth(const std::vector<Sys*>& ss, const std::vector<int64_t>& nd_dst, double sel) {
  int launches = 0;
  std::vector<GProb> gram, prod;
  std::vector<LdlArg> la;
  auto gram_of = [&](Sys& s, const double* V) {
    GProb g;  // G = V^T V [nb, nb]
    g.M = s.nb; g.N = s.nb; g.K = s.n;
    g.A = V; g.lda = s.n;
    g.B = V; g.ldb = s.n;
    g.C = s.gram.as<double>(); g.ldc = s.nb;
    gram.push_back(g);
  };
  auto vt = [&](Sys& s, const double* V, double* out) {
    GProb g;  // out = V Tm [n, nb]
    g.M = s.n; g.N = s.nb; g.K = s.nb;
    g.A = V; g.lda = s.n;
    g.B = s.tm.as<double>(); g.ldb = s.nb;
    g.C = out; g.ldc = s.n;
    prod.push_back(g);
  };
  auto ldl = [&](std::vector<LdlArg>& v) {
    const size_t bytes = sizeof(LdlArg) * v.size();
    HH_REQUIRE(arena_off_ + bytes <= ARENA_BYTES, "cphf batch: too many products in one cycle");
    memcpy(static_cast<char*>(pinned_) + arena_off_, v.data(), bytes);
    LdlArg* d = reinterpret_cast<LdlArg*>(static_cast<char*>(desc_.ptr) + arena_off_);
    HH_HIP(hipMemcpyAsync(d, static_cast<char*>(pinned_) + arena_off_, bytes, hipMemcpyHostToDevice, stream_));
    arena_off_ += (bytes + 255) / 256 * 256;
    hipLaunchKernelGGL(ldl_batch_kernel, dim3((unsigned)v.size()), dim3(KR_LDL_T), 0, stream_, d, lindep_);
    HH_HIP(hipGetLastError());
    ++launches;
  };
  // pass 1: xnew -> w (kept vectors first, zero columns after)
  for (Sys* s : ss) gram_of(*s, s->xnew.as<double>());
  launches += launch(gram, true);
  for (Sys* s : ss)
    la.push_back(LdlArg{s->gram.as<double>(), (int)s->nb, -__builtin_huge_val(), s->tm.as<double>(),
                        s->d1.as<double>(), nullptr, nullptr});
  ldl(la);
  for (Sys* s : ss) vt(*s, s->xnew.as<double>(), s->w.as<double>());
  launches += launch(prod, false);
  // pass 2: w -> xs[:, nd_dst) (selected vectors), norms -> inner[nd_dst)
  gram.clear();
  prod.clear();
  la.clear();
  for (Sys* s : ss) gram_of(*s, s->w.as<double>());
  launches += launch(gram, true);
  double* resd = resd_.as<double>();
  for (size_t i = 0; i < ss.size(); ++i) {
    Sys* s = ss[i];
    const size_t idx = std::find(sys_.begin(), sys_.end(), s) - sys_.begin();
    la.push_back(LdlArg{s->gram.as<double>(), (int)s->nb, sel, s->tm.as<double>(), s->inner.as<double>() + nd_dst[i],
                        resd + 2 * idx, nullptr});
  }
  ldl(la);
  for (size_t i = 0; i < ss.size(); ++i) vt(*ss[i], ss[i]->w.as<double>(), ss[i]->xs.as<double>() + nd_dst[i] * ss[i]->n);
  launches += launch(prod, false);
  return launches;
}

void CphfBatch::solve(std::vector<KrylovStats>* stats) {
  const int64_t nsys = (int64_t)sys_.size();
  if (stats) stats->assign(nsys, KrylovStats{});
  if (nsys == 0) return;
  using clk = std::chrono::steady_clock;
  const auto t_start = clk::now();
  if (!pinned_) {
    HH_HIP(hipHostMalloc(&pinned_, ARENA_BYTES, hipHostMallocDefault));
    pinned_bytes_ = ARENA_BYTES;
  }
  desc_.reserve(ARENA_BYTES);
  resd_.reserve(sizeof(double) * 2 * MAX_SYSTEMS);
  const double conv = std::max(lindep_, tol_ * tol_);
  // the resident matrices, then the per-system buffers
  {
    // largest first, with one set of build intermediates for all (freed once, at the end: with
    // many matrices resident, freeing multi-GB buffers per build costs more than the products)
    std::vector<Sys*> order(sys_.begin(), sys_.end());
    for (Sys* s : order) s->R = s->nao * (s->nocc[0] + (s->nspin > 1 ? s->nocc[1] : 0));
    std::stable_sort(order.begin(), order.end(), [](const Sys* a, const Sys* b) { return a->R > b->R; });
    ResidentScratch scratch;
    for (Sys* s : order) s->m64 = s->rijk->resident_matrix(&s->R, &s->kpad, &scratch);
    HH_HIP(hipStreamSynchronize(stream_));
  }
  const auto t_built = clk::now();
  for (Sys* s : sys_) {
    s->ubra.reserve(sizeof(double) * s->kpad * s->nset);
    fill<double>(s->ubra.as<double>(), s->kpad * s->nset, 0.0, stream_);  // rows R .. kpad stay zero
    s->rbra.reserve(sizeof(double) * s->R * s->nset);
    s->xnew.reserve(sizeof(double) * s->n * s->nset);
    s->w.reserve(sizeof(double) * s->n * s->nset);
    s->gram.reserve(sizeof(double) * s->nset * s->nset);
    s->tm.reserve(sizeof(double) * s->nset * s->nset);
    s->d1.reserve(sizeof(double) * s->nset);
    s->cap = 0;
    grow(*s, std::min<int64_t>(s->nset * (int64_t)(max_cycle_ + 1), s->nset * 17));
    s->nd = 0;
    s->nb = s->nset;
    s->cycles = 0;
    s->done = max_cycle_ == 0;
    s->converged = false;
    s->max_ip = 0.0;
  }
  std::vector<double> hres(2 * nsys);
  auto read_results = [&]() {
    HH_HIP(hipMemcpyAsync(hres.data(), resd_.ptr, sizeof(double) * 2 * nsys, hipMemcpyDeviceToHost, stream_));
    HH_HIP(hipStreamSynchronize(stream_));
    arena_off_ = 0;
  };
  // initial block: the orthogonalised right-hand side (every kept vector) -> xs[:, 0)
  {
    arena_off_ = 0;
    std::vector<Sys*> all(sys_.begin(), sys_.end());
    for (Sys* s : all) copy_bytes(s->xnew.ptr, s->b.ptr, (int64_t)sizeof(double) * s->n * s->nset, stream_);
    orth(all, std::vector<int64_t>(all.size(), 0), -__builtin_huge_val());
    read_results();
    for (int64_t i = 0; i < nsys; ++i) {
      sys_[i]->nb = (int64_t)hres[2 * i + 1];
      if (sys_[i]->nb == 0) sys_[i]->done = true;  // zero right-hand side: x = 0
    }
  }
  const auto t_cyc = clk::now();
  // lockstep cycles
  launches_per_cycle_ = 0;
  lockstep_cycles_ = 0;
  for (;;) {
    std::vector<Sys*> run;
    for (Sys* s : sys_)
      if (!s->done) run.push_back(s);
    if (run.empty()) break;
    arena_off_ = 0;
    int launches = 0;
    std::vector<int64_t> nd_new(run.size());
    for (size_t i = 0; i < run.size(); ++i) {
      nd_new[i] = run[i]->nd + run[i]->nb;
      grow(*run[i], nd_new[i] + run[i]->nset);
    }
    // 1-3: ax = A x on the trial block
    std::vector<GProb> g1, g3, cf, up;
    std::vector<NtProb> g2;
    for (Sys* s : run) add_operator(&g1, &g2, &g3, *s);
    launches += launch(g1, false);
    launches += launch_nt(g2);
    launches += launch(g3, true);
    // 4: coefficients = (xs^T ax) / inner  [nd_new, nb]
    for (size_t i = 0; i < run.size(); ++i) {
      Sys& s = *run[i];
      GProb g;
      g.M = nd_new[i]; g.N = s.nb; g.K = s.n;
      g.A = s.xs.as<double>(); g.lda = s.n;
      g.B = s.ax.as<double>() + s.nd * s.n; g.ldb = s.n;
      g.C = s.coeffs.as<double>(); g.ldc = nd_new[i];
      g.rowdiv = s.inner.as<double>();
      cf.push_back(g);
    }
    launches += launch(cf, true);
    // 5: xnew = ax - xs coefficients
    for (size_t i = 0; i < run.size(); ++i) {
      Sys& s = *run[i];
      GProb g;
      g.M = s.n; g.N = s.nb; g.K = nd_new[i];
      g.A = s.xs.as<double>(); g.lda = s.n;
      g.B = s.coeffs.as<double>(); g.ldb = nd_new[i];
      g.C = s.xnew.as<double>(); g.ldc = s.n;
      g.D = s.ax.as<double>() + s.nd * s.n;
      g.alpha = -1.0; g.beta = 1.0;
      up.push_back(g);
    }
    launches += launch(up, false);
    // 6: orthogonalise, keep the vectors above the threshold as the next block
    launches += orth(run, nd_new, conv);
    read_results();
    ++lockstep_cycles_;
    launches_per_cycle_ = std::max<int64_t>(launches_per_cycle_, launches);
    for (size_t i = 0; i < run.size(); ++i) {
      Sys& s = *run[i];
      const int64_t idx = std::find(sys_.begin(), sys_.end(), run[i]) - sys_.begin();
      s.nd = nd_new[i];
      s.max_ip = hres[2 * idx];
      s.nb = (int64_t)hres[2 * idx + 1];
      ++s.cycles;
      if (s.max_ip < conv) {
        s.converged = true;
        s.done = true;
      } else if (s.nb == 0 || s.cycles >= max_cycle_) {
        s.done = true;
      }
    }
  }
  const auto t_fin = clk::now();
  // projected solve per system: (xs^T ax + diag(inner)) c = xs^T b, x = xs c (krylov_block.rs:204-215)
  arena_off_ = 0;
  std::vector<GProb> hp, gp, xp;
  std::vector<Sys*> solved;
  for (Sys* s : sys_) {
    if (s->nd == 0) continue;
    s->hbuf.reserve(sizeof(double) * s->nd * s->nd);
    s->gbuf.reserve(sizeof(double) * s->nd * s->nset);
    GProb h;
    h.M = s->nd; h.N = s->nd; h.K = s->n;
    h.A = s->xs.as<double>(); h.lda = s->n;
    h.B = s->ax.as<double>(); h.ldb = s->n;
    h.C = s->hbuf.as<double>(); h.ldc = s->nd;
    h.adddiag = s->inner.as<double>();
    hp.push_back(h);
    GProb g;
    g.M = s->nd; g.N = s->nset; g.K = s->n;
    g.A = s->xs.as<double>(); g.lda = s->n;
    g.B = s->b.as<double>(); g.ldb = s->n;
    g.C = s->gbuf.as<double>(); g.ldc = s->nd;
    gp.push_back(g);
    solved.push_back(s);
  }
  launch(hp, true);
  launch(gp, true);
  // the projected systems are independent: their LU solves run side by side on created streams
  // (each a chain of single-workgroup panel steps), with no host synchronisation until all end
  {
    constexpr int NLU = 8;
    if (lu_streams_.empty()) {
      for (int i = 0; i < NLU; ++i) {
        hipStream_t st;
        HH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        lu_streams_.push_back(st);
        hipEvent_t ev;
        HH_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        lu_events_.push_back(ev);
      }
      HH_HIP(hipEventCreateWithFlags(&lu_fork_, hipEventDisableTiming));
    }
    HH_HIP(hipEventRecord(lu_fork_, stream_));
    for (int i = 0; i < NLU; ++i) HH_HIP(hipStreamWaitEvent(lu_streams_[i], lu_fork_, 0));
    for (size_t i = 0; i < solved.size(); ++i) {
      Sys* s = solved[i];
      lu_solve(s->hbuf.as<double>(), s->nd, s->gbuf.as<double>(), s->nset, lu_streams_[i % NLU], s->luws, s->lu.ptr,
               false);
    }
    for (int i = 0; i < NLU; ++i) {
      HH_HIP(hipEventRecord(lu_events_[i], lu_streams_[i]));
      HH_HIP(hipStreamWaitEvent(stream_, lu_events_[i], 0));
    }
  }
  for (Sys* s : solved) {
    GProb x;  // x = xs c, written over xnew's storage ([n, nset])
    x.M = s->n; x.N = s->nset; x.K = s->nd;
    x.A = s->xs.as<double>(); x.lda = s->n;
    x.B = s->gbuf.as<double>(); x.ldb = s->nd;
    x.C = s->xnew.as<double>(); x.ldc = s->n;
    xp.push_back(x);
  }
  launch(xp, false);
  // unstack into the per-spin outputs (systems with no basis: zero)
  bool all_ok = true;
  for (int64_t i = 0; i < nsys; ++i) {
    Sys& s = *sys_[i];
    for (int t = 0; t < s.nspin; ++t) {
      const int64_t rows = s.nmo[t] * s.nocc[t];
      if (s.nd == 0) {
        fill<double>(s.mo1[t], rows * s.nset, 0.0, stream_);
      } else {
        hipLaunchKernelGGL(copy_cols_kernel, dim3(nblk(rows * s.nset)), dim3(256), 0, stream_,
                           s.xnew.as<double>() + s.off[t], s.n, s.mo1[t], rows, rows, s.nset);
        HH_HIP(hipGetLastError());
      }
    }
    if (stats) {
      KrylovStats& k = (*stats)[i];
      k.cycles = s.cycles;
      k.aop_calls = s.cycles;
      k.restarts = 0;
      k.max_innerprod = s.max_ip;
    }
    // a zero right-hand side, or max_cycle 0, returns x = 0 without error (as krylov_block)
    all_ok = all_ok && (s.converged || s.nd == 0);
  }
  HH_HIP(hipStreamSynchronize(stream_));
  for (Sys* s : solved) {
    int bad = 0;
    HH_HIP(hipMemcpy(&bad, static_cast<char*>(s->lu.ptr) + sizeof(int64_t) * s->nd, sizeof(int), hipMemcpyDeviceToHost));
    HH_REQUIRE(bad == 0, "cphf batch: singular projected system");
  }
  const auto t_end = clk::now();
  phase_s_[0] = std::chrono::duration<double>(t_built - t_start).count();
  phase_s_[1] = std::chrono::duration<double>(t_cyc - t_built).count();
  phase_s_[2] = std::chrono::duration<double>(t_fin - t_cyc).count();
  phase_s_[3] = std::chrono::duration<double>(t_end - t_fin).count();
  if (!all_ok) {
    for (int64_t i = 0; i < nsys; ++i)
      if (!sys_[i]->converged && sys_[i]->nd > 0) {
        char buf[200];
        snprintf(buf, sizeof buf, "cphf batch: system %lld did not converge: max(||v||^2) = %.3e >= %.3e after %d cycles",
                 (long long)i, sys_[i]->max_ip, conv, sys_[i]->cycles);
        throw Error(HH_ERR_NOT_CONVERGED, buf);
      }
  }
}

}  // namespace hh
