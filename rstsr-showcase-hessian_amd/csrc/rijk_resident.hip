// Resident RHF RI-JK response operator.
//
// Per perturbation a the RHF response (hess_r.rs:1916-1953: resp = scale_j J_ao mocc + scale_k k_bra,
// k_bra = -K) is linear in bra and, with X_P = B_P C_o and O_P = C_o^T B_P C_o, reads
//   resp[(mu,i), a] = sum_{(nu,j)} M[(mu,i),(nu,j)] bra[(nu,j), a]
//   M = 4 scale_j G - scale_k (V + G~),
//   G[(mu,i),(nu,j)]  = sum_P X_P[mu,i] X_P[nu,j]          (the J term),
//   V[(mu,i),(nu,j)]  = sum_P B_P[mu,nu] O_P[j,i]           (K: sum_P B_P bra_a O_P),
//   G~[(mu,i),(nu,j)] = sum_P X_P[mu,j] X_P[nu,i]           (K: sum_P X_P (bra_a^T X_P)).
// M [nao nocc, nao nocc] depends only on the orbitals: it is built once per prepare_response
// (two products of 2 (nao nocc)^2 naux flop) and kept in HBM, after which each response call is
// one product of 2 (nao nocc)^2 nprop flop that streams M once (HBM-bound).  The per-call work
// falls from O(nao^2 naux nocc nprop) to O(nao^2 nocc^2 nprop); the memory is (nao nocc)^2
// elements (c.f. 288 GB of HBM).
//
// Symmetries used by the assembly kernel (coalesced reads of all three terms): G is symmetric,
// and V'[(mu,nu),(j,i)] = V'[(nu,mu),(i,j)] (B_P and O_P are symmetric).
#include "elementwise.hpp"
#include "gemm_nt.hpp"
#include "rijk.hpp"

namespace hh {

namespace {

unsigned rblocks(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 65536)); }

// O2[(j + nocc i) + nocc^2 P] = O[j + nocc (P + naux i)]
template <typename T>
__global__ void permute_o_kernel(const double* __restrict__ O, int64_t nocc, int64_t naux, T* __restrict__ O2) {
  const int64_t total = nocc * nocc * naux;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e % nocc, i = (e / nocc) % nocc, P = e / (nocc * nocc);
    O2[e] = (T)O[j + nocc * (P + naux * i)];
  }
}

// M[(mu,i),(nu,j)] = cg G[r + R kk] - cv (V'[(nu + nao mu) + nao^2 (i + nocc j)] + G[(nu + nao i) + R (mu + nao j)])
// with r = mu + nao i, kk = nu + nao j.  Output BF16: tiled hi / lo [(kk / 32) R + r][32] (kk < kpad,
// zero past R); FP64: row-major [r][kpad].  One thread per output element, kk fastest.
template <typename T, bool BF>
__global__ void assemble_m_kernel(const T* __restrict__ G, const T* __restrict__ Vp, int64_t nao, int64_t nocc,
                                  int64_t kpad, double cg, double cv, unsigned short* __restrict__ hi,
                                  unsigned short* __restrict__ lo, double* __restrict__ m64) {
  const int64_t R = nao * nocc, n2 = nao * nao;
  const int64_t total = R * kpad;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t r, kk;
    if (BF) {
      const int64_t kl = e % 32, rest = e / 32;
      r = rest % R;
      kk = (rest / R) * 32 + kl;
    } else {
      kk = e % kpad;
      r = e / kpad;
    }
    double v = 0.0;
    if (kk < R) {
      const int64_t mu = r % nao, i = r / nao, nu = kk % nao, j = kk / nao;
      const double g = (double)G[kk + R * r];
      const double gt = (double)G[(nu + nao * i) + R * (mu + nao * j)];
      const double vv = (double)Vp[(nu + nao * mu) + n2 * (i + nocc * j)];
      v = cg * g - cv * (vv + gt);
    }
    if (BF) {
      const float x = (float)v;
      const __bf16 h = (__bf16)x;
      hi[e] = __builtin_bit_cast(unsigned short, h);
      lo[e] = __builtin_bit_cast(unsigned short, (__bf16)(x - (float)h));
    } else {
      m64[e] = v;
    }
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

// dst[(a) kpad + k] = src[k + R a] (k < R), 0 for R <= k < kpad
__global__ void pad_rows_kernel(const double* __restrict__ src, int64_t R, int64_t kpad, int64_t n,
                                double* __restrict__ dst) {
  const int64_t total = kpad * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = e % kpad, a = e / kpad;
    dst[e] = k < R ? src[k + R * a] : 0.0;
  }
}

}  // namespace

bool Rijk::resident_fits(Prec prec) const {
  if (nspin_ != 1 || !prepared_ || prec == PREC_F32) return false;
  const double R = (double)nao_ * (double)nocc_[0];
  // the FP64 matrix, the two FP64 product intermediates while it is assembled, and the BF16x3 copy
  const double bytes = 3.0 * 8.0 * R * R + (prec == PREC_BF16X3 ? 4.0 * R * R : 0.0);
  return bytes <= (double)resident_budget;
}

void Rijk::build_resident(Prec prec) {
  HH_REQUIRE(nspin_ == 1, "resident response: RHF handles only");
  HH_REQUIRE(prec == PREC_F64 || prec == PREC_BF16X3, "resident response: FP64 or BF16x3");
  const int64_t nocc = nocc_[0], R = nao_ * nocc, kpad = ceil_div(R, 32) * 32;
  if (!res_ready_[0]) {
    // FP64 matrix: the products in FP64 (the matrix is exact to FP64 reassociation)
    DevBuf G, Vp, O2;
    double* g = static_cast<double*>(G.reserve(sizeof(double) * R * R));
    double* vp = static_cast<double*>(Vp.reserve(sizeof(double) * R * R));
    // G[(mu,i),(nu,j)] = sum_P X[mu,P,i] X[nu,P,j]: one [nao x nao] block per (i, j), batched
    {
      GemmArgs a;
      a.M = nao_; a.N = nao_; a.K = naux_;
      a.A = x_[0].ptr; a.lda = nao_; a.a_kmajor = false;   // A[mu, P] = X[mu + nao P (+ nao naux i)]
      a.B = x_[0].ptr; a.ldb = nao_; a.b_kmajor = false;   // B[P, nu] = X[nu + nao P (+ nao naux j)]
      a.C = g; a.ldc = R;
      a.nb0 = nocc; a.sA0 = nao_ * naux_; a.sB0 = 0; a.sC0 = nao_;
      a.nb1 = nocc; a.sA1 = 0; a.sB1 = nao_ * naux_; a.sC1 = R * nao_;
      gemm<double, double>(a, PREC_F64, stream_, ws_);
    }
    // V'[(mu,nu),(j,i)] = sum_P B[mu,nu,P] O2[P,(j,i)]
    {
      double* o2 = static_cast<double*>(O2.reserve(sizeof(double) * nocc * nocc * naux_));
      hipLaunchKernelGGL(permute_o_kernel<double>, dim3(rblocks(nocc * nocc * naux_)), dim3(256), 0, stream_,
                         o_[0].as<double>(), nocc, naux_, o2);
      HH_HIP(hipGetLastError());
      GemmArgs a;
      a.M = nao_ * nao_; a.N = nocc * nocc; a.K = naux_;
      a.A = bfull_.ptr; a.lda = nao_ * nao_; a.a_kmajor = false;
      a.B = o2; a.ldb = nocc * nocc; a.b_kmajor = false;
      a.C = vp; a.ldc = nao_ * nao_;
      gemm<double, double>(a, PREC_F64, stream_, ws_);
    }
    res64_.reserve(sizeof(double) * R * kpad);
    hipLaunchKernelGGL((assemble_m_kernel<double, false>), dim3(rblocks(R * kpad)), dim3(256), 0, stream_, g, vp, nao_,
                       nocc, kpad, 4.0 * scale_j_, scale_k_, nullptr, nullptr, res64_.as<double>());
    HH_HIP(hipGetLastError());
    HH_HIP(hipStreamSynchronize(stream_));  // G, V' and O2 are freed on return
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

void Rijk::apply_resident(const double* bra, int64_t nprop, double* resp, Prec prec, bool accumulate) {
  const bool bf = prec == PREC_BF16X3;
  build_resident(prec);
  const int64_t R = nao_ * nocc_[0], kpad = ceil_div(R, 32) * 32;
  NtArgs a;
  a.M = R; a.N = nprop; a.K = kpad;
  a.C = resp; a.ldc = R;
  a.alpha = 1.0; a.beta = accumulate ? 1.0 : 0.0;
  if (profiling_) {
    const int e0 = next_event_++, e1 = next_event_++;
    a.ev_begin = ev(e0);
    a.ev_main_end = ev(e1);
    pending_.push_back({T_K_HOT, {e0, e1}});
  }
  if (bf) {
    unsigned short* b = static_cast<unsigned short*>(res_b_.reserve(2 * sizeof(unsigned short) * kpad * nprop));
    split_bf16_tiled<double>(bra, R, R, nprop, kpad, b, b + kpad * nprop, stream_);
    a.A = res_hi_.ptr; a.A_lo = res_lo_.ptr;
    a.B = b; a.B_lo = b + kpad * nprop;
    a.lda = a.ldb = kpad;
    a.tiled = true;
    gemm_nt(a, NT_BF16X3, stream_, ws_);
  } else {
    double* b = static_cast<double*>(res_b_.reserve(sizeof(double) * kpad * nprop));
    hipLaunchKernelGGL(pad_rows_kernel, dim3(rblocks(kpad * nprop)), dim3(256), 0, stream_, bra, R, kpad, nprop, b);
    HH_HIP(hipGetLastError());
    a.A = res64_.ptr; a.lda = kpad;
    a.B = b; a.ldb = kpad;
    gemm_nt(a, NT_F64, stream_, ws_);
  }
}

}  // namespace hh
