// Resident RI-JK response operator.
//
// Per perturbation a the RHF response (hess_r.rs:1916-1953: resp = scale_j J_ao mocc + scale_k k_bra,
// k_bra = -K) is linear in bra and, with X_P = B_P C_o and O_P = C_o^T B_P C_o, reads
//   resp[(mu,i), a] = sum_{(nu,j)} M[(mu,i),(nu,j)] bra[(nu,j), a]
//   M = 4 scale_j G - scale_k (V + G~),
//   G[(mu,i),(nu,j)]  = sum_P X_P[mu,i] X_P[nu,j]          (the J term),
//   V[(mu,i),(nu,j)]  = sum_P B_P[mu,nu] O_P[j,i]           (K: sum_P B_P bra_a O_P),
//   G~[(mu,i),(nu,j)] = sum_P X_P[mu,j] X_P[nu,i]           (K: sum_P X_P (bra_a^T X_P)).
// UHF (hess_u.rs:270-313: r_s = 1/2 scale_j J_ao(sum_t dm1_t) mocc_s + scale_k k_bra_s): the rows
// and columns run over both spins, (mu, i, s) with the alpha block first, and
//   M[s, t] = 2 scale_j G^{st} - [s == t] scale_k (V^s + G~^s),  G^{st} = sum_P X^s_P (x) X^t_P:
// J couples the spins, K does not.
// M [R, R] (R = nao sum_s nocc_s) depends only on the orbitals: it is built once per
// prepare_response (products of 2 R^2 naux flop) and kept in HBM, after which each response call
// is one product of 2 R^2 nprop flop per spin block of rows that streams M once (HBM-bound).  The
// per-call work falls from O(nao^2 naux nocc nprop) to O(nao^2 nocc^2 nprop); the memory is R^2
// elements (c.f. 288 GB of HBM).
//
// Build: G is one NT product of X arranged as Xr[(mu,i,s)][P] with itself (M = N = R, K = naux);
// V^s needs only the packed triangles of its two symmetric index pairs,
// Vtp[(mu<=nu), (i<=j)] = sum_P B_P[mu,nu] O^s_P[i,j], a quarter of the full product's flops, and
// is unpacked over (mu, nu) before the assembly, which reads all terms coalesced (kk fastest).
#include "elementwise.hpp"
#include "gemm_nt.hpp"
#include "rijk.hpp"

namespace hh {

namespace {

unsigned rblocks(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 65536)); }

// packed index of the pair (i, j) of a symmetric index pair: hi (hi + 1) / 2 + lo
__device__ __forceinline__ int64_t tp_index(int64_t i, int64_t j) {
  const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
  return hi * (hi + 1) / 2 + lo;
}

// O2t[tp(i,j) + ntp P] = O[j + nocc (P + naux i)] for i <= j (O_P is symmetric), ntp = nocc (nocc + 1) / 2
__global__ void pack_o_tp_kernel(const double* __restrict__ O, int64_t nocc, int64_t naux, double* __restrict__ O2t) {
  const int64_t ntp = nocc * (nocc + 1) / 2;
  const int64_t total = ntp * naux;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = e % ntp, P = e / ntp;
    int64_t j = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (j * (j + 1) / 2 > t) --j;
    while ((j + 1) * (j + 2) / 2 <= t) ++j;
    const int64_t i = t - j * (j + 1) / 2;
    O2t[e] = O[j + nocc * (P + naux * i)];
  }
}

// Xr[(mu + nao i) ldk + P] = X[mu + nao (P + naux i)], zero for naux <= P < ldk: the [nao, naux]
// slice of each i transposed through a 64 x 64 LDS tile (grid: mu tiles, P tiles, i)
__global__ __launch_bounds__(256) void transpose_x_kernel(const double* __restrict__ X, int64_t nao, int64_t naux,
                                                          int64_t ldk, double* __restrict__ Xr) {
  __shared__ double t[64][65];
  const int64_t mu0 = (int64_t)blockIdx.x * 64, P0 = (int64_t)blockIdx.y * 64, i = blockIdx.z;
  const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
  const double* src = X + nao * naux * i;
  for (int pp = ty; pp < 64; pp += 4) {
    const int64_t P = P0 + pp, mu = mu0 + tx;
    t[pp][tx] = (mu < nao && P < naux) ? src[mu + nao * P] : 0.0;
  }
  __syncthreads();
  for (int mm = ty; mm < 64; mm += 4) {
    const int64_t mu = mu0 + mm, P = P0 + tx;
    if (mu < nao && P < ldk) Xr[(mu + nao * i) * ldk + P] = t[tx][mm];
  }
}

// M[r][kk] (FP64 row-major [R][kpad], zero for kk >= R): r = (mu, i, s), kk = (nu, j, t) with the
// alpha block first (Ra rows); cg G[kk + R r] - [s == t] cv (Vu_s[(nu + nao mu) + nao^2 tp(i, j)]
// + G[(nu + nao i + off_s) + R (mu + nao j + off_s)]) (Vu_s: the (mu, nu)-unpacked Vtp of spin s).
// One thread per output element, kk fastest.
__global__ void assemble_m_kernel(const double* __restrict__ G, const double* __restrict__ Vu0,
                                  const double* __restrict__ Vu1, int64_t nao, int64_t Ra, int64_t R, int64_t kpad,
                                  double cg, double cv, double* __restrict__ m64) {
  const int64_t n2 = nao * nao;
  const int64_t total = R * kpad;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t kk = e % kpad, r = e / kpad;
    double v = 0.0;
    if (kk < R) {
      const int s = r >= Ra, t = kk >= Ra;
      const int64_t off_s = s ? Ra : 0, off_t = t ? Ra : 0;
      const int64_t rl = r - off_s, kl = kk - off_t;
      v = cg * G[kk + R * r];
      if (s == t) {
        const int64_t mu = rl % nao, i = rl / nao, nu = kl % nao, j = kl / nao;
        const double gt = G[(nu + nao * i + off_s) + R * (mu + nao * j + off_s)];
        const double vv = (s ? Vu1 : Vu0)[(nu + nao * mu) + n2 * tp_index(i, j)];
        v -= cv * (vv + gt);
      }
    }
    m64[e] = v;
  }
}

// BF16x3 copy of the FP64 matrix: hi / lo [(kk / 32) R + r][32] from m64[r kpad + kk]; one thread
// per 8 consecutive kk (a 16-byte store per array), threads r-fastest
__global__ void split_m_kernel(const double* __restrict__ m64, int64_t R, int64_t kpad, unsigned short* __restrict__ hi,
                               unsigned short* __restrict__ lo) {
  const int64_t total = R * (kpad / 8);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e % R, k0 = 8 * (e / R);
    const double* src = m64 + r * kpad + k0;
    unsigned short h[8], l[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float x = (float)src[q];
      const __bf16 b = (__bf16)x;
      h[q] = __builtin_bit_cast(unsigned short, b);
      l[q] = __builtin_bit_cast(unsigned short, (__bf16)(x - (float)b));
    }
    const int64_t o = ((k0 / 32) * R + r) * 32 + k0 % 32;
    *reinterpret_cast<uint4*>(hi + o) = *reinterpret_cast<const uint4*>(h);
    *reinterpret_cast<uint4*>(lo + o) = *reinterpret_cast<const uint4*>(l);
  }
}

// dst[a kpad + k] = bra_0[k + Ra a] (k < Ra), bra_1[(k - Ra) + Rb a] (Ra <= k < Ra + Rb), 0 up to kpad:
// the spins' bra stacked along K
__global__ void stack_rows_kernel(const double* __restrict__ b0, const double* __restrict__ b1, int64_t Ra, int64_t Rb,
                                  int64_t kpad, int64_t n, double* __restrict__ dst) {
  const int64_t total = kpad * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = e % kpad, a = e / kpad;
    dst[e] = k < Ra ? b0[k + Ra * a] : (k < Ra + Rb ? b1[(k - Ra) + Rb * a] : 0.0);
  }
}

}  // namespace

bool Rijk::resident_fits(Prec prec) const {
  if (!prepared_ || prec == PREC_F32) return false;
  double R = 0.0;
  for (int t = 0; t < nspin_; ++t) R += (double)nao_ * (double)nocc_[t];
  // the FP64 matrix, the two FP64 product intermediates while it is assembled, and the BF16x3 copy
  const double bytes = 3.0 * 8.0 * R * R + (prec == PREC_BF16X3 ? 4.0 * R * R : 0.0);
  return bytes <= (double)resident_budget;
}

void Rijk::build_resident(Prec prec, ResidentScratch* scratch) {
  HH_REQUIRE(prec == PREC_F64 || prec == PREC_BF16X3, "resident response: FP64 or BF16x3");
  const int64_t Ra = nao_ * nocc_[0], R = Ra + (nspin_ > 1 ? nao_ * nocc_[1] : 0), kpad = ceil_div(R, 32) * 32;
  if (!res_ready_[0]) {
    // FP64 matrix: the products in FP64 (the matrix is exact to FP64 reassociation).  The
    // intermediates live in the caller's scratch (kept for the next build) or in local buffers
    // freed on return.
    const int64_t ldk = ceil_div(naux_, 32) * 32;
    ResidentScratch local;
    ResidentScratch& w = scratch ? *scratch : local;
    double* g = static_cast<double*>(w.g.reserve(sizeof(double) * R * R));
    // G[(mu,i,s),(nu,j,t)] = sum_P Xr[(mu,i,s)][P] Xr[(nu,j,t)][P]: one NT product, K = naux padded to 32
    {
      double* xr = static_cast<double*>(w.xr.reserve(sizeof(double) * R * ldk));
      for (int t = 0; t < nspin_; ++t)
        hipLaunchKernelGGL(transpose_x_kernel,
                           dim3((unsigned)ceil_div(nao_, 64), (unsigned)ceil_div(ldk, 64), (unsigned)nocc_[t]),
                           dim3(256), 0, stream_, x_[t].as<double>(), nao_, naux_, ldk, xr + (t ? Ra : 0) * ldk);
      HH_HIP(hipGetLastError());
      NtArgs a;
      a.M = R; a.N = R; a.K = ldk;
      a.A = xr; a.B = xr; a.lda = a.ldb = ldk;
      a.C = g; a.ldc = R;
      a.ksplit = 1;
      gemm_nt(a, NT_F64, stream_, ws_);
    }
    // per spin: Vtp[(mu<=nu), (i<=j)] = sum_P cderi[(mu,nu), P] O2t[P, (i,j)], then unpacked over (mu, nu)
    double* vp[2] = {nullptr, nullptr};
    const int64_t ntp0 = nocc_[0] * (nocc_[0] + 1) / 2;
    const int64_t ntp1 = nspin_ > 1 ? nocc_[1] * (nocc_[1] + 1) / 2 : 0;
    double* vbase = static_cast<double*>(w.vp.reserve(sizeof(double) * nao_ * nao_ * (ntp0 + ntp1)));
    for (int t = 0; t < nspin_; ++t) {
      const int64_t nocc = nocc_[t], ntp = nocc * (nocc + 1) / 2;
      vp[t] = vbase + (t ? nao_ * nao_ * ntp0 : 0);
      double* o2 = static_cast<double*>(w.o2t.reserve(sizeof(double) * ntp * naux_));
      double* vtp = static_cast<double*>(w.vtp.reserve(sizeof(double) * nao_tp_ * ntp));
      hipLaunchKernelGGL(pack_o_tp_kernel, dim3(rblocks(ntp * naux_)), dim3(256), 0, stream_, o_[t].as<double>(), nocc,
                         naux_, o2);
      HH_HIP(hipGetLastError());
      GemmArgs a;
      a.M = nao_tp_; a.N = ntp; a.K = naux_;
      a.A = cderi_.ptr; a.lda = nao_tp_; a.a_kmajor = false;
      a.B = o2; a.ldb = ntp; a.b_kmajor = false;
      a.C = vtp; a.ldc = nao_tp_;
      gemm<double, double>(a, PREC_F64, stream_, ws_);
      unpack_triu<double, double>(vtp, nao_, nao_tp_, ntp, vp[t], stream_);
    }
    res64_.reserve(sizeof(double) * R * kpad);
    // J weight: 4 sj (RHF, doubly occupied) or 2 sj (UHF: 1/2 of J from both spins' dm1)
    const double cg = (nspin_ == 1 ? 4.0 : 2.0) * scale_j_;
    hipLaunchKernelGGL(assemble_m_kernel, dim3(rblocks(R * kpad)), dim3(256), 0, stream_, g, vp[0],
                       nspin_ > 1 ? vp[1] : vp[0], nao_, Ra, R, kpad, cg, scale_k_, res64_.as<double>());
    HH_HIP(hipGetLastError());
    HH_HIP(hipStreamSynchronize(stream_));  // local intermediates are freed on return
    res_ready_[0] = true;
  }
  if (prec == PREC_BF16X3 && !res_ready_[1]) {
    res_hi_.reserve(sizeof(unsigned short) * R * kpad);
    res_lo_.reserve(sizeof(unsigned short) * R * kpad);
    hipLaunchKernelGGL(split_m_kernel, dim3(rblocks(R * (kpad / 8))), dim3(256), 0, stream_, res64_.as<double>(), R,
                       kpad, res_hi_.as<unsigned short>(), res_lo_.as<unsigned short>());
    HH_HIP(hipGetLastError());
    res_ready_[1] = true;
  }
}

const double* Rijk::resident_matrix(int64_t* R, int64_t* kpad, ResidentScratch* scratch) {
  HH_REQUIRE(prepared_, "resident response matrix: make_response_preparation first");
  build_resident(PREC_F64, scratch);
  const int64_t r = nao_ * (nocc_[0] + (nspin_ > 1 ? nocc_[1] : 0));
  if (R) *R = r;
  if (kpad) *kpad = ceil_div(r, 32) * 32;
  return res64_.as<double>();
}

void Rijk::apply_resident(const double* const* bra, int64_t nprop, double* const* resp, Prec prec,
                          bool accumulate) {
  const bool bf = prec == PREC_BF16X3;
  build_resident(prec);
  const int64_t Ra = nao_ * nocc_[0], Rb = nspin_ > 1 ? nao_ * nocc_[1] : 0, R = Ra + Rb;
  const int64_t kpad = ceil_div(R, 32) * 32;
  // the spins' bra stacked along K (FP64, zero-padded to kpad)
  const size_t b64 = sizeof(double) * kpad * nprop;
  double* b = static_cast<double*>(res_b_.reserve(b64 + (bf ? 2 * sizeof(unsigned short) * kpad * nprop : 0)));
  hipLaunchKernelGGL(stack_rows_kernel, dim3(rblocks(kpad * nprop)), dim3(256), 0, stream_, bra[0],
                     nspin_ > 1 ? bra[1] : bra[0], Ra, Rb, kpad, nprop, b);
  HH_HIP(hipGetLastError());
  unsigned short* bh = nullptr;
  if (bf) {
    bh = reinterpret_cast<unsigned short*>(reinterpret_cast<char*>(b) + b64);
    split_bf16_tiled<double>(b, kpad, kpad, nprop, kpad, bh, bh + kpad * nprop, stream_);
  }
  const int e0 = profiling_ ? mark() : -1;
  // one product per spin block of rows of M
  for (int s = 0; s < nspin_; ++s) {
    const int64_t off = s ? Ra : 0, Rs = s ? Rb : Ra;
    NtArgs a;
    a.M = Rs; a.N = nprop; a.K = kpad;
    a.C = resp[s]; a.ldc = Rs;
    a.alpha = 1.0; a.beta = accumulate ? 1.0 : 0.0;
    if (bf) {
      a.A = res_hi_.as<unsigned short>() + off * 32; a.A_lo = res_lo_.as<unsigned short>() + off * 32;
      a.a_rows = R;
      a.B = bh; a.B_lo = bh + kpad * nprop;
      a.lda = a.ldb = kpad;
      a.tiled = true;
      gemm_nt(a, NT_BF16X3, stream_, ws_);
    } else {
      a.A = res64_.as<double>() + off * kpad; a.lda = kpad;
      a.B = b; a.ldb = kpad;
      gemm_nt(a, NT_F64, stream_, ws_);
    }
  }
  if (profiling_) pending_.push_back({T_K_HOT, {e0, mark()}});
}

}  // namespace hh
