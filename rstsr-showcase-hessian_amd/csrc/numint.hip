// Grid contractions of the RKS CPKS response and the XC effective-kernel transform.
//
// The reference loops grid chunks (384 points) x perturbation sets with per-set GEMMs and
// Mutex-guarded accumulation (pure_eval_rho.rs:216-334, pure_xcpot.rs:370-465).  Here every
// perturbation is a column of one GEMM, per grid batch:
//   aob[g, i, c]     = sum_mu ao[g, mu, c] C_o[mu, i]                (c < ncomp)
//   akb[g, (i,a), c] = sum_mu ao[g, mu, c] bra[mu, (i,a)]            (matrix cores)
//   rho1[g, x, a], wv[g, y, a] = w_g sum_x fxc[g, x, y] rho1[g, x, a], and the weighted
//   back-contraction operands S_c[g, (i,a)]                           (one pass, FP64)
//   out[mu, (i,a)]  += sum_c sum_g ao[g, mu, c] S_c[g, (i,a)]         (matrix cores)
// so rho1 never leaves the batch and accumulation is a GEMM beta, not a lock.
#include "numint.hpp"

#include "elementwise.hpp"
#include "gemm_nt.hpp"
#include <hip/hip_bf16.h>

namespace hh {

namespace {
unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 65535)); }

// rho1 of one (g, a) pair from the half-transformed AOs.
template <int DEN, typename T>
__device__ __forceinline__ void rho1_point(const double* aob, const T* akb, int64_t gb, int64_t g, int64_t nocc,
                                           int64_t ncol, int64_t a, double* r) {
  const int64_t sc_b = gb * nocc, sc_k = gb * ncol;
  for (int x = 0; x < 5; ++x) r[x] = 0.0;
  for (int64_t i = 0; i < nocc; ++i) {
    const double b0 = aob[g + gb * i];
    const double k0 = (double)akb[g + gb * (i + nocc * a)];
    r[0] += b0 * k0;
    if (DEN != DEN_RHO) {
#pragma unroll
      for (int t = 1; t < 4; ++t) {
        const double bt = aob[g + gb * i + sc_b * t];
        const double kt = (double)akb[g + gb * (i + nocc * a) + sc_k * t];
        r[t] += b0 * kt + bt * k0;
        if (DEN == DEN_TAU) r[4] += 0.5 * bt * kt;
      }
    }
  }
}

// wv[y] = w sum_x fxc[g, x, y] rho1[x]
template <int DEN>
__device__ __forceinline__ void weighted_kernel(const double* fxc, int64_t ld_fxc, int64_t g, double w, const double* r,
                                                double* wv) {
  constexpr int NV = DEN == DEN_RHO ? 1 : (DEN == DEN_SIGMA ? 4 : 5);
#pragma unroll
  for (int y = 0; y < NV; ++y) {
    double v = 0.0;
#pragma unroll
    for (int x = 0; x < NV; ++x) v += fxc[g + ld_fxc * (x + NV * y)] * r[x];
    wv[y] = w * v;
  }
}

// write S_c[g, (i,a)] (into `S`, which may alias akb: each (g, a) owns its entries)
template <int DEN, typename T>
__device__ __forceinline__ void write_s(const double* aob, T* S, int64_t gb, int64_t g, int64_t nocc,
                                        int64_t ncol, int64_t a, const double* wv) {
  const int64_t sc_b = gb * nocc, sc_k = gb * ncol;
  for (int64_t i = 0; i < nocc; ++i) {
    const double b0 = aob[g + gb * i];
    const int64_t o = g + gb * (i + nocc * a);
    if (DEN == DEN_RHO) {
      S[o] = (T)(b0 * wv[0]);
    } else {
      const double b1 = aob[g + gb * i + sc_b], b2 = aob[g + gb * i + 2 * sc_b], b3 = aob[g + gb * i + 3 * sc_b];
      S[o] = (T)(b0 * wv[0] + b1 * wv[1] + b2 * wv[2] + b3 * wv[3]);
      const double h = (DEN == DEN_TAU) ? 0.5 * wv[4] : 0.0;
      S[o + sc_k] = (T)(b0 * wv[1] + h * b1);
      S[o + 2 * sc_k] = (T)(b0 * wv[2] + h * b2);
      S[o + 3 * sc_k] = (T)(b0 * wv[3] + h * b3);
    }
  }
}

template <int DEN, typename T>
__global__ void rho1_kernel(const double* aob, const void* akb_v, int64_t gb, int64_t nocc, int64_t nprop,
                            double* rho1, int64_t ng, int64_t g0) {
  const T* akb = static_cast<const T*>(akb_v);
  constexpr int NV = DEN == DEN_RHO ? 1 : (DEN == DEN_SIGMA ? 4 : 5);
  const int64_t total = gb * nprop, ncol = nocc * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = e % gb, a = e / gb;
    double r[5];
    rho1_point<DEN>(aob, akb, gb, g, nocc, ncol, a, r);
#pragma unroll
    for (int x = 0; x < NV; ++x) rho1[(g0 + g) + ng * (x + NV * a)] = r[x];
  }
}

// fused: rho1 from (aob, akb), then S into akb
template <int DEN, typename T>
__global__ void rho1_wv_s_kernel(const double* aob, void* akb_v, const double* fxc, const double* w, int64_t gb,
                                 int64_t nocc, int64_t nprop, int64_t ng, int64_t g0) {
  T* akb = static_cast<T*>(akb_v);
  const int64_t total = gb * nprop, ncol = nocc * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = e % gb, a = e / gb;
    double r[5], wv[5];
    rho1_point<DEN>(aob, akb, gb, g, nocc, ncol, a, r);
    weighted_kernel<DEN>(fxc + g0, ng, g, w[g0 + g], r, wv);
    write_s<DEN>(aob, akb, gb, g, nocc, ncol, a, wv);
  }
}

// S from a given rho1 [ng, nvar, nprop]
template <int DEN, typename T>
__global__ void wv_s_kernel(const double* aob, const double* rho1, void* S_v, const double* fxc, const double* w,
                            int64_t gb, int64_t nocc, int64_t nprop, int64_t ng, int64_t g0) {
  T* S = static_cast<T*>(S_v);
  constexpr int NV = DEN == DEN_RHO ? 1 : (DEN == DEN_SIGMA ? 4 : 5);
  const int64_t total = gb * nprop, ncol = nocc * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = e % gb, a = e / gb;
    double r[5], wv[5];
#pragma unroll
    for (int x = 0; x < NV; ++x) r[x] = rho1[(g0 + g) + ng * (x + NV * a)];
    weighted_kernel<DEN>(fxc + g0, ng, g, w[g0 + g], r, wv);
    write_s<DEN>(aob, S, gb, g, nocc, ncol, a, wv);
  }
}

// ---- BF16X3: S written directly as the tiled bf16 hi/lo back-contraction operand ----
// component c of element (g, n) at c * cs + ((g / 32) * ncol + n) * 32 + g % 32 (K = grid point)
__device__ __forceinline__ void put_split(unsigned short* hi, unsigned short* lo, int64_t idx, double v) {
  const float x = (float)v;
  const __bf16 h = (__bf16)x;
  hi[idx] = __builtin_bit_cast(unsigned short, h);
  lo[idx] = __builtin_bit_cast(unsigned short, (__bf16)(x - (float)h));
}

template <int DEN>
__device__ __forceinline__ void write_s_tiled(const double* aob, unsigned short* hi, unsigned short* lo, int64_t cs,
                                              int64_t gb, int64_t g, int64_t nocc, int64_t ncol, int64_t a,
                                              const double* wv) {
  const int64_t sc_b = gb * nocc;
  const int64_t tb = (g >> 5) * ncol * 32 + (g & 31);
  for (int64_t i = 0; i < nocc; ++i) {
    const double b0 = aob[g + gb * i];
    const int64_t t = tb + (i + nocc * a) * 32;
    if (DEN == DEN_RHO) {
      put_split(hi, lo, t, b0 * wv[0]);
    } else {
      const double b1 = aob[g + gb * i + sc_b], b2 = aob[g + gb * i + 2 * sc_b], b3 = aob[g + gb * i + 3 * sc_b];
      const double h = (DEN == DEN_TAU) ? 0.5 * wv[4] : 0.0;
      put_split(hi, lo, t, b0 * wv[0] + b1 * wv[1] + b2 * wv[2] + b3 * wv[3]);
      put_split(hi, lo, t + cs, b0 * wv[1] + h * b1);
      put_split(hi, lo, t + 2 * cs, b0 * wv[2] + h * b2);
      put_split(hi, lo, t + 3 * cs, b0 * wv[3] + h * b3);
    }
  }
}

template <int DEN>
__global__ void rho1_wv_s_tiled_kernel(const double* aob, const float* akb, unsigned short* hi, unsigned short* lo,
                                       int64_t cs, const double* fxc, const double* w, int64_t gb, int64_t nocc,
                                       int64_t nprop, int64_t ng, int64_t g0) {
  const int64_t total = gb * nprop, ncol = nocc * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = e % gb, a = e / gb;
    double r[5], wv[5];
    rho1_point<DEN, float>(aob, akb, gb, g, nocc, ncol, a, r);
    weighted_kernel<DEN>(fxc + g0, ng, g, w[g0 + g], r, wv);
    write_s_tiled<DEN>(aob, hi, lo, cs, gb, g, nocc, ncol, a, wv);
  }
}

// ---- BF16X3, RKS: the kets product fused with the pointwise pass ----
// One workgroup (2 waves) takes 32 points of a grid batch and NA perturbations with all their
// occupied orbitals (columns n = i + nocc a, 16 NFR at most).  akb[g, n, c] is computed on the
// matrix cores from the tiled AO values (K = mu) and the tiled split kets: the same split-bf16
// products, in the same order per output, as the NT kernel that writes akb for the unfused path
// (one FP32 accumulation over mu, which is what that kernel does at K = naop).  akb stays in
// LDS; rho1, the weighted kernel and the tiled S operand follow with the arithmetic of
// rho1_wv_s_tiled_kernel, so S is bit-identical to the unfused path's and the FP32 akb array
// (0.5 GB per batch at 72 perturbations) never reaches HBM.
// workgroup w runs on XCD w % 8: renumber so that consecutive indices share an XCD
__device__ __forceinline__ int64_t ni_xcd_remap(int64_t w, int64_t nwg) {
  const int64_t q = nwg / 8, r = nwg % 8, x = w % 8;
  return x * q + (x < r ? x : r) + w / 8;
}

typedef float v4f_ni __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf_ni __attribute__((ext_vector_type(8)));

template <int DEN, int NFR>
__global__ __launch_bounds__(128) void kets_rho1_s_kernel(const unsigned short* __restrict__ aoth,
                                                          const unsigned short* __restrict__ aotl, int64_t nt,
                                                          int64_t ngp, const unsigned short* __restrict__ kh,
                                                          const unsigned short* __restrict__ kl, int64_t ncol, int nkb,
                                                          const double* __restrict__ aob, const double* __restrict__ fxc,
                                                          const double* __restrict__ w, unsigned short* __restrict__ hi,
                                                          unsigned short* __restrict__ lo, int64_t cs, int64_t gb,
                                                          int64_t nocc, int64_t nprop, int na, int64_t ng, int64_t g0) {
  constexpr int NC = DEN == DEN_RHO ? 1 : 4;
  constexpr int NV = DEN == DEN_RHO ? 1 : (DEN == DEN_SIGMA ? 4 : 5);
  constexpr int LDA = NFR * 16 + 4;  // LDS row stride (floats): conflict-free staging stores
  __shared__ float akbs[NC][32 * LDA];
  __shared__ double wvs[32 * 8 * NV];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // one-dimensional grid, perturbation tiles fastest after an XCD-aware remap: the tiles of one
  // 32-point block run together on one XCD, so its AO values come from HBM once and from L2 after
  const int64_t nwg = gridDim.x, nty = (nprop + na - 1) / na;
  const int64_t lin = ni_xcd_remap(blockIdx.x, nwg);
  const int64_t gt0 = (lin / nty) * 32;  // first point of the tile (batch-local)
  const int64_t a0 = (lin % nty) * na;
  const int nat = (int)min((int64_t)na, nprop - a0);
  const int64_t n0 = nocc * a0;
  const int nct = (int)(nocc * nat);  // columns of this tile
  // kets: wave `wave` owns points gt0 + 16 wave .. + 15
  v4f_ni acc[NFR][NC];
#pragma unroll
  for (int f = 0; f < NFR; ++f)
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[f][c] = v4f_ni{0.f, 0.f, 0.f, 0.f};
  const int64_t grow = g0 + gt0 + 16 * wave + (lane & 15);
  const int kq = 8 * (lane >> 4);
  for (int kb = 0; kb < nkb; ++kb) {
    v8bf_ni bh[NFR], bl[NFR];
#pragma unroll
    for (int f = 0; f < NFR; ++f) {
      const int nl = 16 * f + (lane & 15);
      uint4 h = {0, 0, 0, 0}, l = {0, 0, 0, 0};
      if (nl < nct) {
        const int64_t off = ((int64_t)kb * ncol + n0 + nl) * 32 + kq;
        h = *reinterpret_cast<const uint4*>(kh + off);
        l = *reinterpret_cast<const uint4*>(kl + off);
      }
      bh[f] = __builtin_bit_cast(v8bf_ni, h);
      bl[f] = __builtin_bit_cast(v8bf_ni, l);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int64_t off = c * nt + ((int64_t)kb * ngp + grow) * 32 + kq;
      const v8bf_ni ah = *reinterpret_cast<const v8bf_ni*>(aoth + off);
      const v8bf_ni al = *reinterpret_cast<const v8bf_ni*>(aotl + off);
#pragma unroll
      for (int f = 0; f < NFR; ++f) {
        acc[f][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[f], acc[f][c], 0, 0, 0);
        acc[f][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[f], acc[f][c], 0, 0, 0);
        acc[f][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[f], acc[f][c], 0, 0, 0);
      }
    }
  }
  // stage: the lane holds akb[point 16 wave + 4 (lane / 16) + r][column 16 f + lane % 16]
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int f = 0; f < NFR; ++f)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        akbs[c][(16 * wave + 4 * (lane >> 4) + r) * LDA + 16 * f + (lane & 15)] = acc[f][c][r];
  __syncthreads();
  // rho1 and the weighted kernel per (point, perturbation)
  const int64_t sc_b = gb * nocc;
  for (int pr = tid; pr < 32 * nat; pr += 128) {
    const int gl = pr & 31, al = pr >> 5;
    const int64_t g = gt0 + gl;
    if (g < gb) {
      double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      for (int64_t i = 0; i < nocc; ++i) {
        const double b0 = aob[g + gb * i];
        const int col = (int)(i + nocc * al);
        const double k0 = (double)akbs[0][gl * LDA + col];
        r[0] += b0 * k0;
        if (DEN != DEN_RHO) {
#pragma unroll
          for (int t = 1; t < 4; ++t) {
            const double bt = aob[g + gb * i + sc_b * t];
            const double kt = (double)akbs[t < NC ? t : 0][gl * LDA + col];
            r[t] += b0 * kt + bt * k0;
            if (DEN == DEN_TAU) r[4] += 0.5 * bt * kt;
          }
        }
      }
      double wv[5];
      weighted_kernel<DEN>(fxc + g0, ng, g, w[g0 + g], r, wv);
#pragma unroll
      for (int x = 0; x < NV; ++x) wvs[(gl * 8 + al) * NV + x] = wv[x];
    }
  }
  __syncthreads();
  // tiled S: 4 groups of 32 lanes (one 32-point block: 64 contiguous bytes per column)
  const int gl = tid & 31, q = tid >> 5;
  const int64_t g = gt0 + gl;
  if (g >= gb) return;
  const int64_t tb = (g >> 5) * ncol * 32 + (g & 31);
  for (int p = q; p < nct; p += 4) {
    const int al = (int)(p / nocc);
    const int64_t i = p - (int64_t)al * nocc;
    const double* wv = wvs + (gl * 8 + al) * NV;
    const double b0 = aob[g + gb * i];
    const int64_t t = tb + (n0 + p) * 32;
    if (DEN == DEN_RHO) {
      put_split(hi, lo, t, b0 * wv[0]);
    } else {
      const double b1 = aob[g + gb * i + sc_b], b2 = aob[g + gb * i + 2 * sc_b], b3 = aob[g + gb * i + 3 * sc_b];
      const double h = (DEN == DEN_TAU) ? 0.5 * wv[4] : 0.0;
      put_split(hi, lo, t, b0 * wv[0] + b1 * wv[1] + b2 * wv[2] + b3 * wv[3]);
      put_split(hi, lo, t + cs, b0 * wv[1] + h * b1);
      put_split(hi, lo, t + 2 * cs, b0 * wv[2] + h * b2);
      put_split(hi, lo, t + 3 * cs, b0 * wv[3] + h * b3);
    }
  }
}

template <int DEN>
__global__ void wv_s_tiled_kernel(const double* aob, const double* rho1, unsigned short* hi, unsigned short* lo,
                                  int64_t cs, const double* fxc, const double* w, int64_t gb, int64_t nocc,
                                  int64_t nprop, int64_t ng, int64_t g0) {
  constexpr int NV = DEN == DEN_RHO ? 1 : (DEN == DEN_SIGMA ? 4 : 5);
  const int64_t total = gb * nprop, ncol = nocc * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = e % gb, a = e / gb;
    double r[5], wv[5];
#pragma unroll
    for (int x = 0; x < NV; ++x) r[x] = rho1[(g0 + g) + ng * (x + NV * a)];
    weighted_kernel<DEN>(fxc + g0, ng, g, w[g0 + g], r, wv);
    write_s_tiled<DEN>(aob, hi, lo, cs, gb, g, nocc, ncol, a, wv);
  }
}

// ---- UKS: the kernel couples the spins (fxc_eff [ng, NV, 2, NV, 2]) ----
// w_s[x] = w sum_{y,t} fxc[g, y, t, x, s] r_t[y]
template <int DEN>
__device__ __forceinline__ void weighted_kernel_uks(const double* fxc, int64_t ld, int64_t g, double w,
                                                    const double* ra, const double* rb, double* wa, double* wb) {
  constexpr int NV = DEN == DEN_RHO ? 1 : (DEN == DEN_SIGMA ? 4 : 5);
#pragma unroll
  for (int sp = 0; sp < 2; ++sp)
#pragma unroll
    for (int x = 0; x < NV; ++x) {
      double v = 0.0;
#pragma unroll
      for (int y = 0; y < NV; ++y) {
        v += fxc[g + ld * (y + NV * (0 + 2 * (x + NV * sp)))] * ra[y];
        v += fxc[g + ld * (y + NV * (1 + 2 * (x + NV * sp)))] * rb[y];
      }
      if (sp == 0) wa[x] = w * v;  // (a select of the two arrays would put them in scratch memory)
      else wb[x] = w * v;
    }
}

// fused: rho1 of both spins from (aob_s, akb_s), then S_s into akb_s (both rho1 first: S overwrites)
template <int DEN, typename T>
__global__ __launch_bounds__(256) void uks_rho1_wv_s_kernel(const double* aob_a, void* akb_a_v, const double* aob_b, void* akb_b_v,
                                     const double* fxc, const double* w, int64_t gb, int64_t nocc_a, int64_t nocc_b,
                                     int64_t nprop, int64_t ng, int64_t g0) {
  T* akb_a = static_cast<T*>(akb_a_v);
  T* akb_b = static_cast<T*>(akb_b_v);
  const int64_t total = gb * nprop, ncol_a = nocc_a * nprop, ncol_b = nocc_b * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = e % gb, a = e / gb;
    double ra[5], rb[5], wa[5], wb[5];
    rho1_point<DEN>(aob_a, akb_a, gb, g, nocc_a, ncol_a, a, ra);
    rho1_point<DEN>(aob_b, akb_b, gb, g, nocc_b, ncol_b, a, rb);
    weighted_kernel_uks<DEN>(fxc + g0, ng, g, w[g0 + g], ra, rb, wa, wb);
    write_s<DEN>(aob_a, akb_a, gb, g, nocc_a, ncol_a, a, wa);
    write_s<DEN>(aob_b, akb_b, gb, g, nocc_b, ncol_b, a, wb);
  }
}

// S_s from a given rho1 [ng, NV, 2, nprop]
template <int DEN, typename T>
__global__ __launch_bounds__(256) void uks_wv_s_kernel(const double* aob_a, const double* aob_b, const double* rho1, void* Sa_v, void* Sb_v,
                                const double* fxc, const double* w, int64_t gb, int64_t nocc_a, int64_t nocc_b,
                                int64_t nprop, int64_t ng, int64_t g0) {
  T* Sa = static_cast<T*>(Sa_v);
  T* Sb = static_cast<T*>(Sb_v);
  constexpr int NV = DEN == DEN_RHO ? 1 : (DEN == DEN_SIGMA ? 4 : 5);
  const int64_t total = gb * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = e % gb, a = e / gb;
    double ra[5], rb[5], wa[5], wb[5];
#pragma unroll
    for (int y = 0; y < NV; ++y) {
      ra[y] = rho1[(g0 + g) + ng * (y + NV * (0 + 2 * a))];
      rb[y] = rho1[(g0 + g) + ng * (y + NV * (1 + 2 * a))];
    }
    weighted_kernel_uks<DEN>(fxc + g0, ng, g, w[g0 + g], ra, rb, wa, wb);
    write_s<DEN>(aob_a, Sa, gb, g, nocc_a, nocc_a * nprop, a, wa);
    write_s<DEN>(aob_b, Sb, gb, g, nocc_b, nocc_b * nprop, a, wb);
  }
}

template <int DEN>
__global__ __launch_bounds__(256) void uks_rho1_wv_s_tiled_kernel(const double* aob_a, const float* akb_a, const double* aob_b,
                                           const float* akb_b, unsigned short* hi_a, unsigned short* lo_a,
                                           int64_t cs_a, unsigned short* hi_b, unsigned short* lo_b, int64_t cs_b,
                                           const double* fxc, const double* w, int64_t gb, int64_t nocc_a,
                                           int64_t nocc_b, int64_t nprop, int64_t ng, int64_t g0) {
  const int64_t total = gb * nprop, ncol_a = nocc_a * nprop, ncol_b = nocc_b * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = e % gb, a = e / gb;
    double ra[5], rb[5], wa[5], wb[5];
    rho1_point<DEN, float>(aob_a, akb_a, gb, g, nocc_a, ncol_a, a, ra);
    rho1_point<DEN, float>(aob_b, akb_b, gb, g, nocc_b, ncol_b, a, rb);
    weighted_kernel_uks<DEN>(fxc + g0, ng, g, w[g0 + g], ra, rb, wa, wb);
    write_s_tiled<DEN>(aob_a, hi_a, lo_a, cs_a, gb, g, nocc_a, ncol_a, a, wa);
    write_s_tiled<DEN>(aob_b, hi_b, lo_b, cs_b, gb, g, nocc_b, ncol_b, a, wb);
  }
}

template <int DEN>
__global__ __launch_bounds__(256) void uks_wv_s_tiled_kernel(const double* aob_a, const double* aob_b, const double* rho1,
                                      unsigned short* hi_a, unsigned short* lo_a, int64_t cs_a, unsigned short* hi_b,
                                      unsigned short* lo_b, int64_t cs_b, const double* fxc, const double* w,
                                      int64_t gb, int64_t nocc_a, int64_t nocc_b, int64_t nprop, int64_t ng,
                                      int64_t g0) {
  constexpr int NV = DEN == DEN_RHO ? 1 : (DEN == DEN_SIGMA ? 4 : 5);
  const int64_t total = gb * nprop;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = e % gb, a = e / gb;
    double ra[5], rb[5], wa[5], wb[5];
#pragma unroll
    for (int y = 0; y < NV; ++y) {
      ra[y] = rho1[(g0 + g) + ng * (y + NV * (0 + 2 * a))];
      rb[y] = rho1[(g0 + g) + ng * (y + NV * (1 + 2 * a))];
    }
    weighted_kernel_uks<DEN>(fxc + g0, ng, g, w[g0 + g], ra, rb, wa, wb);
    write_s_tiled<DEN>(aob_a, hi_a, lo_a, cs_a, gb, g, nocc_a, nocc_a * nprop, a, wa);
    write_s_tiled<DEN>(aob_b, hi_b, lo_b, cs_b, gb, g, nocc_b, nocc_b * nprop, a, wb);
  }
}

// ---- xceff transform (xc_deriv.rs:305-428, unpolarised) ----
template <int DEN>
__global__ void xceff_kernel(int order, const double* rho, int64_t ld_rho, const double* xc, int64_t ld_xc, int64_t ng,
                             double* out) {
  constexpr int NV = DEN == DEN_SIGMA ? 4 : 5;
  constexpr int XL = DEN == DEN_SIGMA ? 2 : 3;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x) {
    double c[NV];
    int var[NV];
    c[0] = 1.0;
    var[0] = 0;
#pragma unroll
    for (int t = 1; t < 4; ++t) {
      c[t] = 2.0 * rho[g + ld_rho * t];
      var[t] = 1;
    }
    if (NV == 5) {
      c[NV - 1] = 1.0;
      var[NV - 1] = 2;
    }
    if (order == 1) {
#pragma unroll
      for (int x = 0; x < NV; ++x) out[g + ng * x] = c[x] * xc[g + ld_xc * (1 + var[x])];
    } else {
      const double fsig = xc[g + ld_xc * 2];
#pragma unroll
      for (int y = 0; y < NV; ++y)
#pragma unroll
        for (int x = 0; x < NV; ++x) {
          const int p = var[x] < var[y] ? var[x] : var[y], q = var[x] < var[y] ? var[y] : var[x];
          const int k = p * XL - p * (p - 1) / 2 + (q - p);  // combinations_with_replacement index
          double v = c[x] * c[y] * xc[g + ld_xc * (1 + XL + k)];
          if (x == y && x >= 1 && x <= 3) v += 2.0 * fsig;
          out[g + ng * (x + NV * y)] = v;
        }
    }
  }
}

__global__ void xceff_lda_kernel(int order, const double* xc, int64_t ld_xc, int64_t ng, double* out) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x)
    out[g] = xc[g + ld_xc * order];
}

// ---- spin-polarised transform ----
// Column (x, s) of the Jacobian d(XCFun variable)/d(rho_t variable) has at most two entries:
// rho_s -> variable s; grad_t rho_a -> sigma_aa (2 grad_t rho_a), sigma_ab (grad_t rho_b);
// grad_t rho_b -> sigma_ab (grad_t rho_a), sigma_bb (2 grad_t rho_b); tau_s -> 5 + s.
template <int DEN>
__global__ __launch_bounds__(256) void xceff_pol_kernel(int order, const double* rho, int64_t ld_rho, int64_t ld_spin, const double* xc,
                                 int64_t ld_xc, int64_t ng, double* out) {
  constexpr int NV = DEN == DEN_SIGMA ? 4 : 5;
  constexpr int XL = DEN == DEN_SIGMA ? 5 : 7;
  constexpr int NC = 2 * NV;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x) {
    int k1[NC], k2[NC];
    double c1[NC], c2[NC];
#pragma unroll
    for (int sp = 0; sp < 2; ++sp) {
      k1[NV * sp] = sp; c1[NV * sp] = 1.0; k2[NV * sp] = -1; c2[NV * sp] = 0.0;
#pragma unroll
      for (int t = 1; t < 4; ++t) {
        const double ga = rho[g + ld_rho * t], gbv = rho[g + ld_rho * t + ld_spin];
        const int c = t + NV * sp;
        if (sp == 0) { k1[c] = 2; c1[c] = 2.0 * ga; k2[c] = 3; c2[c] = gbv; }
        else { k1[c] = 3; c1[c] = ga; k2[c] = 4; c2[c] = 2.0 * gbv; }
      }
      if (NV == 5) { k1[4 + NV * sp] = 5 + sp; c1[4 + NV * sp] = 1.0; k2[4 + NV * sp] = -1; c2[4 + NV * sp] = 0.0; }
    }
    if (order == 1) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        double v = c1[c] * xc[g + ld_xc * (1 + k1[c])];
        if (k2[c] >= 0) v += c2[c] * xc[g + ld_xc * (1 + k2[c])];
        out[g + ng * c] = v;
      }
    } else {
      auto f2 = [&](int p, int q) {
        if (p > q) { const int r = p; p = q; q = r; }
        return xc[g + ld_xc * (1 + XL + p * XL - p * (p - 1) / 2 + (q - p))];
      };
      const double fsaa = xc[g + ld_xc * 3], fsab = xc[g + ld_xc * 4], fsbb = xc[g + ld_xc * 5];
#pragma unroll
      for (int cy = 0; cy < NC; ++cy)
#pragma unroll
        for (int cx = 0; cx < NC; ++cx) {
          double v = c1[cx] * c1[cy] * f2(k1[cx], k1[cy]);
          if (k2[cy] >= 0) v += c1[cx] * c2[cy] * f2(k1[cx], k2[cy]);
          if (k2[cx] >= 0) v += c2[cx] * c1[cy] * f2(k2[cx], k1[cy]);
          if (k2[cx] >= 0 && k2[cy] >= 0) v += c2[cx] * c2[cy] * f2(k2[cx], k2[cy]);
          const int x = cx % NV, y = cy % NV, sx = cx / NV, sy = cy / NV;
          if (x == y && x >= 1 && x <= 3)  // d2 sigma / d grad d grad: [[2 f_aa, f_ab], [f_ab, 2 f_bb]]
            v += (sx != sy) ? fsab : (sx == 0 ? 2.0 * fsaa : 2.0 * fsbb);
          out[g + ng * (cx + NC * cy)] = v;
        }
    }
  }
}

// polarised LDA: the raw block; order 2 unfolds (aa, ab, bb) to [s, t]
__global__ void xceff_lda_pol_kernel(int order, const double* xc, int64_t ld_xc, int64_t ng, double* out) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x) {
    if (order == 1) {
      out[g] = xc[g + ld_xc * 1];
      out[g + ng] = xc[g + ld_xc * 2];
    } else {
      for (int t = 0; t < 2; ++t)
        for (int sp = 0; sp < 2; ++sp) out[g + ng * (sp + 2 * t)] = xc[g + ld_xc * (3 + sp + t)];
    }
  }
}

}  // namespace

Numint::Numint(const double* ao, int64_t ng, int64_t nao, int64_t ncomp, const double* weights, hipStream_t s)
    : ng_(ng), nao_(nao), ncomp_(ncomp), stream_(s) {
  HH_REQUIRE(ao && weights && ng > 0 && nao > 0 && (ncomp == 1 || ncomp >= 4), "numint: bad AO data");
  ao_.reserve(sizeof(double) * ng * nao * ncomp);
  HH_HIP(hipMemcpyAsync(ao_.ptr, ao, sizeof(double) * ng * nao * ncomp, hipMemcpyDeviceToDevice, s));
  w_.reserve(sizeof(double) * ng);
  HH_HIP(hipMemcpyAsync(w_.ptr, weights, sizeof(double) * ng, hipMemcpyDeviceToDevice, s));
}

int64_t Numint::pick_batch(int64_t ncol) const {
  if (batch > 0) return std::min(batch, ng_);
  // about 8 GB of per-batch intermediates: few batches, so the back-contraction's split-K
  // partial sums are written a few times per call rather than once per small batch
  const int64_t per_point = 8 * (ncol + 64) * 4;
  int64_t gb = (int64_t)(1LL << 33) / std::max<int64_t>(per_point, 1);
  gb = std::max<int64_t>(256, (gb / 256) * 256);
  return std::min(gb, ng_);
}

#define HH_DEN_DISPATCH(den, KERNEL, ...)                                     \
  do {                                                                        \
    if ((den) == DEN_RHO) hipLaunchKernelGGL(KERNEL<DEN_RHO>, __VA_ARGS__);   \
    else if ((den) == DEN_SIGMA) hipLaunchKernelGGL(KERNEL<DEN_SIGMA>, __VA_ARGS__); \
    else hipLaunchKernelGGL(KERNEL<DEN_TAU>, __VA_ARGS__);                    \
    HH_HIP(hipGetLastError());                                                \
  } while (0)

// launch KERNEL<den, T> with T = float (bf) or double; the T* arguments are cast by the caller
#define HH_DEN_DISPATCH_T(den, bf, KERNEL, ...)                                          \
  do {                                                                                   \
    if (bf) {                                                                            \
      if ((den) == DEN_RHO) hipLaunchKernelGGL((KERNEL<DEN_RHO, float>), __VA_ARGS__);   \
      else if ((den) == DEN_SIGMA) hipLaunchKernelGGL((KERNEL<DEN_SIGMA, float>), __VA_ARGS__); \
      else hipLaunchKernelGGL((KERNEL<DEN_TAU, float>), __VA_ARGS__);                    \
    } else {                                                                             \
      if ((den) == DEN_RHO) hipLaunchKernelGGL((KERNEL<DEN_RHO, double>), __VA_ARGS__);  \
      else if ((den) == DEN_SIGMA) hipLaunchKernelGGL((KERNEL<DEN_SIGMA, double>), __VA_ARGS__); \
      else hipLaunchKernelGGL((KERNEL<DEN_TAU, double>), __VA_ARGS__);                   \
    }                                                                                    \
    HH_HIP(hipGetLastError());                                                           \
  } while (0)

namespace {
// out[g, n, c] = sum_mu ao[g0+g, mu, c] B[mu, n] for one grid batch (c < ncomp)
void half_transform(const double* ao, int64_t ng, int64_t nao, int64_t g0, int64_t gb, int ncomp, const double* B,
                    int64_t ncolB, double* out, Prec prec, hipStream_t s, DevBuf& ws) {
  GemmArgs g;
  g.M = gb; g.N = ncolB; g.K = nao;
  g.A = ao + g0; g.lda = ng; g.a_kmajor = false; g.sA0 = ng * nao;
  g.B = B; g.ldb = nao; g.b_kmajor = true; g.sB0 = 0;
  g.C = out; g.ldc = gb; g.sC0 = gb * ncolB;
  g.nb0 = ncomp;
  gemm<double, double>(g, prec, s, ws);
}

// out[mu, n] (+)= alpha sum_c sum_g ao[g0+g, mu, c] S_c[g, n]
void back_contract(const double* ao, int64_t ng, int64_t nao, int64_t g0, int64_t gb, int ncomp, const double* S,
                   int64_t ncol, double* out, double alpha, bool accumulate, Prec prec, hipStream_t s, DevBuf& ws) {
  for (int c = 0; c < ncomp; ++c) {
    GemmArgs g;
    g.M = nao; g.N = ncol; g.K = gb;
    g.A = ao + g0 + ng * nao * c; g.lda = ng; g.a_kmajor = true;
    g.B = S + gb * ncol * c; g.ldb = gb; g.b_kmajor = true;
    g.C = out; g.ldc = nao;
    g.alpha = alpha;
    g.beta = (accumulate || c > 0) ? 1.0 : 0.0;
    gemm<double, double>(g, prec, s, ws);
  }
}

// AO values of one component as the tiled BF16X3 operand with K = mu (rows = grid points):
// hi/lo[((mu / 32) * ngp + g) * 32 + mu % 32] = split(ao[g + ng mu]), zero for mu >= nao or g >= ng.
// One thread per (K block, g): reads are consecutive in g, each thread writes 64 contiguous bytes.
__global__ void ao_tiled_kmu_kernel(const double* __restrict__ ao, int64_t ng, int64_t nao, int64_t ngp, int64_t nkb,
                                    unsigned short* __restrict__ hi, unsigned short* __restrict__ lo) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= nkb * ngp) return;
  const int64_t g = e % ngp, kb = e / ngp;
  unsigned short h[32], l[32];
#pragma unroll
  for (int kk = 0; kk < 32; ++kk) {
    const int64_t mu = kb * 32 + kk;
    const float x = (g < ng && mu < nao) ? (float)ao[g + ng * mu] : 0.0f;
    const __hip_bfloat16 bh = __float2bfloat16(x);
    const float r = x - __bfloat162float(bh);
    const __hip_bfloat16 bl = __float2bfloat16(r);
    h[kk] = *reinterpret_cast<const unsigned short*>(&bh);
    l[kk] = *reinterpret_cast<const unsigned short*>(&bl);
  }
  uint4* ph = reinterpret_cast<uint4*>(hi + e * 32);
  uint4* pl = reinterpret_cast<uint4*>(lo + e * 32);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    ph[q] = make_uint4(h[8 * q] | (unsigned)h[8 * q + 1] << 16, h[8 * q + 2] | (unsigned)h[8 * q + 3] << 16,
                       h[8 * q + 4] | (unsigned)h[8 * q + 5] << 16, h[8 * q + 6] | (unsigned)h[8 * q + 7] << 16);
    pl[q] = make_uint4(l[8 * q] | (unsigned)l[8 * q + 1] << 16, l[8 * q + 2] | (unsigned)l[8 * q + 3] << 16,
                       l[8 * q + 4] | (unsigned)l[8 * q + 5] << 16, l[8 * q + 6] | (unsigned)l[8 * q + 7] << 16);
  }
}
}  // namespace

template <int DEN>
void launch_kets_rho1_s_den(int nfr, dim3 grid, hipStream_t s, const unsigned short* aoth, const unsigned short* aotl,
                            int64_t nt, int64_t ngp, const unsigned short* kh, const unsigned short* kl, int64_t ncol,
                            int nkb, const double* aob, const double* fxc, const double* w, unsigned short* hi,
                            unsigned short* lo, int64_t cs, int64_t gb, int64_t nocc, int64_t nprop, int na, int64_t ng,
                            int64_t g0) {
  // grid: (point tiles, perturbation tiles) launched as one dimension (see the kernel)
  const dim3 g1((unsigned)((int64_t)grid.x * grid.y));
  HH_REQUIRE((int64_t)grid.x * grid.y < (1LL << 31), "numint: fused grid too large");
#define HH_KRS(N)                                                                                                   \
  case N:                                                                                                           \
    hipLaunchKernelGGL((kets_rho1_s_kernel<DEN, N>), g1, dim3(128), 0, s, aoth, aotl, nt, ngp, kh, kl, ncol, nkb,   \
                       aob, fxc, w, hi, lo, cs, gb, nocc, nprop, na, ng, g0);                                       \
    break;
  switch (nfr) {
    HH_KRS(1) HH_KRS(2) HH_KRS(3) HH_KRS(4) HH_KRS(5) HH_KRS(6) HH_KRS(7) HH_KRS(8)
    default: throw Error(HH_ERR_INVALID, "numint: fused tile too wide");
  }
#undef HH_KRS
  HH_HIP(hipGetLastError());
}

static void launch_kets_rho1_s(int den, int nfr, dim3 grid, hipStream_t s, const unsigned short* aoth,
                        const unsigned short* aotl, int64_t nt, int64_t ngp, const unsigned short* kh,
                        const unsigned short* kl, int64_t ncol, int nkb, const double* aob, const double* fxc,
                        const double* w, unsigned short* hi, unsigned short* lo, int64_t cs, int64_t gb, int64_t nocc,
                        int64_t nprop, int na, int64_t ng, int64_t g0) {
  if (den == DEN_RHO)
    launch_kets_rho1_s_den<DEN_RHO>(nfr, grid, s, aoth, aotl, nt, ngp, kh, kl, ncol, nkb, aob, fxc, w, hi, lo, cs, gb,
                                    nocc, nprop, na, ng, g0);
  else if (den == DEN_SIGMA)
    launch_kets_rho1_s_den<DEN_SIGMA>(nfr, grid, s, aoth, aotl, nt, ngp, kh, kl, ncol, nkb, aob, fxc, w, hi, lo, cs,
                                      gb, nocc, nprop, na, ng, g0);
  else
    launch_kets_rho1_s_den<DEN_TAU>(nfr, grid, s, aoth, aotl, nt, ngp, kh, kl, ncol, nkb, aob, fxc, w, hi, lo, cs, gb,
                                    nocc, nprop, na, ng, g0);
}

// Tiled bf16 hi/lo copies of the AO values (components 0..3) for the BF16X3 grid GEMMs:
// aok_* with K = grid point (back-contraction A operand, rows mu), aot_* with K = mu
// (half-transform A operand, rows = grid points).  Built once per handle.
void Numint::ensure_bf16() {
  if (bf_ready_) return;
  ngp_ = ceil_div(ng_, 32) * 32;
  naop_ = ceil_div(nao_, 32) * 32;
  const int nc = (int)std::min<int64_t>(ncomp_, 4);
  const size_t nk = (size_t)ngp_ * nao_, nt = (size_t)naop_ * ngp_;
  unsigned short* kh = static_cast<unsigned short*>(aok_hi_.reserve(sizeof(unsigned short) * nk * nc));
  unsigned short* kl = static_cast<unsigned short*>(aok_lo_.reserve(sizeof(unsigned short) * nk * nc));
  unsigned short* th = static_cast<unsigned short*>(aot_hi_.reserve(sizeof(unsigned short) * nt * nc));
  unsigned short* tl = static_cast<unsigned short*>(aot_lo_.reserve(sizeof(unsigned short) * nt * nc));
  for (int c = 0; c < nc; ++c) {
    const double* a = ao_.as<double>() + ng_ * nao_ * c;
    split_bf16_tiled<double>(a, ng_, ng_, nao_, ngp_, kh + nk * c, kl + nk * c, stream_);
    const int64_t n = (naop_ / 32) * ngp_;
    hipLaunchKernelGGL(ao_tiled_kmu_kernel, dim3(nblk(n)), dim3(256), 0, stream_, a, ng_, nao_, ngp_, naop_ / 32,
                       th + nt * c, tl + nt * c);
    HH_HIP(hipGetLastError());
  }
  bf_ready_ = true;
}

// BF16X3 half-transform of one batch: out[g, n, c] (FP32) = sum_mu ao[g0+g, mu, c] B[mu, n], with
// B given as its tiled split (K = mu, kpad naop_)
void Numint::kets_bf16(int nc, int64_t g0, int64_t gb, const unsigned short* bh, const unsigned short* bl,
                       int64_t ncol, float* out) {
  const size_t nt = (size_t)naop_ * ngp_;
  for (int c = 0; c < nc; ++c) {
    NtArgs a;
    a.M = gb; a.N = ncol; a.K = naop_;
    a.A = aot_hi_.as<unsigned short>() + nt * c + g0 * 32;
    a.A_lo = aot_lo_.as<unsigned short>() + nt * c + g0 * 32;
    a.a_rows = ngp_;
    a.B = bh; a.B_lo = bl;
    a.C_f32 = out + gb * ncol * c; a.ldc = gb;
    a.tiled = true;
    gemm_nt(a, NT_BF16X3, stream_, ws_);
  }
}

// Tiled bf16 hi/lo buffers for S of one batch ([nc][gbp / 32][ncol][32] each); rows of the last K
// block beyond gb are cleared (the pointwise kernels write rows < gb).
void Numint::s_tiled(int nc, int64_t gb, int64_t ncol, DevBuf& hi, DevBuf& lo, unsigned short** h,
                     unsigned short** l) {
  const int64_t gbp = ceil_div(gb, 32) * 32, cs = gbp * ncol;
  *h = static_cast<unsigned short*>(hi.reserve(sizeof(unsigned short) * cs * nc));
  *l = static_cast<unsigned short*>(lo.reserve(sizeof(unsigned short) * cs * nc));
  if (gb % 32 != 0)
    for (int c = 0; c < nc; ++c) {
      fill<unsigned short>(*h + cs * c + (gbp - 32) * ncol, 32 * ncol, 0, stream_);
      fill<unsigned short>(*l + cs * c + (gbp - 32) * ncol, 32 * ncol, 0, stream_);
    }
}

// BF16X3 back-contraction of one batch: out[mu, n] (+)= alpha sum_c sum_g ao[g0+g, mu, c] S_c[g, n],
// S given tiled (s_tiled layout)
void Numint::back_bf16(int nc, int64_t g0, int64_t gb, const unsigned short* sh, const unsigned short* sl,
                       int64_t ncol, double* out, double alpha, bool accumulate) {
  const int64_t gbp = ceil_div(gb, 32) * 32, cs = gbp * ncol;
  const size_t nk = (size_t)ngp_ * nao_;
  for (int c = 0; c < nc; ++c) {
    NtArgs a;
    a.M = nao_; a.N = ncol; a.K = gbp;
    a.A = aok_hi_.as<unsigned short>() + nk * c + (g0 / 32) * nao_ * 32;
    a.A_lo = aok_lo_.as<unsigned short>() + nk * c + (g0 / 32) * nao_ * 32;
    a.B = sh + cs * c; a.B_lo = sl + cs * c;
    a.C = out; a.ldc = nao_;
    a.alpha = alpha;
    a.beta = (accumulate || c > 0) ? 1.0 : 0.0;
    a.tiled = true;
    gemm_nt(a, NT_BF16X3, stream_, ws_);
  }
}

// tiled split of a [nao, ncol] operand (K = mu) into the given buffers
void Numint::split_kets(const double* B, int64_t ncol, DevBuf& hi, DevBuf& lo) {
  unsigned short* h = static_cast<unsigned short*>(hi.reserve(sizeof(unsigned short) * naop_ * ncol));
  unsigned short* l = static_cast<unsigned short*>(lo.reserve(sizeof(unsigned short) * naop_ * ncol));
  split_bf16_tiled<double>(B, nao_, nao_, ncol, naop_, h, l, stream_);
}

// Batch size: from memory, or the user's; a multiple of 32 for the BF16X3 path (K blocks).
int64_t Numint::batch_for(int64_t ncol, bool bf) const {
  int64_t gb = pick_batch(ncol);
  if (bf) gb = std::min<int64_t>(ceil_div(gb, 32) * 32, ceil_div(ng_, 32) * 32);
  return gb;
}

void Numint::rho1_one_bra_mult_ket(int den, const double* bra, int64_t nocc, const double* kets, int64_t nprop,
                                   double* rho1, Prec prec) {
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "numint: den_type must be RHO, SIGMA or TAU");
  const int nc = den_ncomp(den);
  HH_REQUIRE(ncomp_ >= nc, "numint: AO component dimension insufficient");
  const bool bf = prec == PREC_BF16X3;
  const int64_t ncol = nocc * nprop;
  const int64_t gbmax = batch_for(ncol, bf);
  const size_t es = bf ? sizeof(float) : sizeof(double);
  double* aob = static_cast<double*>(aob_.reserve(sizeof(double) * gbmax * nocc * nc));
  void* akb = akb_.reserve(es * gbmax * ncol * nc);
  if (bf) {
    ensure_bf16();
    split_kets(kets, ncol, kh_, kl_);
  }
  for (int64_t g0 = 0; g0 < ng_; g0 += gbmax) {
    const int64_t gb = std::min(gbmax, ng_ - g0);
    half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, bra, nocc, aob, PREC_F64, stream_, ws_);
    if (bf) kets_bf16(nc, g0, gb, kh_.as<unsigned short>(), kl_.as<unsigned short>(), ncol, static_cast<float*>(akb));
    else half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, kets, ncol, static_cast<double*>(akb), prec, stream_, ws_);
    HH_DEN_DISPATCH_T(den, bf, rho1_kernel, dim3(nblk(gb * nprop)), dim3(256), 0, stream_, aob, akb, gb, nocc, nprop,
                      rho1, ng_, g0);
  }
}

void Numint::fxc_pot_bra_trans(int den, const double* fxc_eff, const double* rho1, int64_t nprop, const double* bra,
                               int64_t nocc, double* out, Prec prec, double alpha, bool accumulate) {
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "numint: den_type must be RHO, SIGMA or TAU");
  const int nc = den_ncomp(den);
  HH_REQUIRE(ncomp_ >= nc, "numint: AO component dimension insufficient");
  const bool bf = prec == PREC_BF16X3;
  const int64_t ncol = nocc * nprop;
  const int64_t gbmax = batch_for(ncol, bf);
  const size_t es = bf ? sizeof(float) : sizeof(double);
  double* aob = static_cast<double*>(aob_.reserve(sizeof(double) * gbmax * nocc * nc));
  void* S = bf ? nullptr : akb_.reserve(es * gbmax * ncol * nc);
  if (bf) ensure_bf16();
  for (int64_t g0 = 0; g0 < ng_; g0 += gbmax) {
    const int64_t gb = std::min(gbmax, ng_ - g0);
    half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, bra, nocc, aob, PREC_F64, stream_, ws_);
    if (bf) {
      unsigned short *h, *l;
      s_tiled(nc, gb, ncol, sh_, sl_, &h, &l);
      HH_DEN_DISPATCH(den, wv_s_tiled_kernel, dim3(nblk(gb * nprop)), dim3(256), 0, stream_, aob, rho1, h, l,
                      ceil_div(gb, 32) * 32 * ncol, fxc_eff, w_.as<double>(), gb, nocc, nprop, ng_, g0);
      back_bf16(nc, g0, gb, h, l, ncol, out, alpha, accumulate || g0 > 0);
    } else {
      HH_DEN_DISPATCH_T(den, false, wv_s_kernel, dim3(nblk(gb * nprop)), dim3(256), 0, stream_, aob, rho1, S, fxc_eff,
                        w_.as<double>(), gb, nocc, nprop, ng_, g0);
      back_contract(ao_.as<double>(), ng_, nao_, g0, gb, nc, static_cast<double*>(S), ncol, out, alpha,
                    accumulate || g0 > 0, prec, stream_, ws_);
    }
  }
}

namespace {
// Exchange-correlation potential operand of one grid batch: S_c[g, nu] with
// V[mu, nu] = sum_c sum_g ao_c[g, mu] S_c[g, nu] = int phi_mu v_rho phi_nu + v_t (phi_mu d_t phi_nu +
// d_t phi_mu phi_nu) + v_tau / 2 grad phi_mu . grad phi_nu  (wv = w vxc_eff)
template <int DEN>
__global__ void vxc_s_kernel(const double* __restrict__ ao, int64_t ng, int64_t nao, int64_t g0, int64_t gb,
                             const double* __restrict__ vxc, const double* __restrict__ w, double* __restrict__ S) {
  const int64_t total = gb * nao;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = e % gb, nu = e / gb;
    const int64_t gg = g0 + g;
    const double wg = w[gg];
    const double phi = ao[gg + ng * nu];
    if (DEN == DEN_RHO) {
      S[e] = wg * vxc[gg] * phi;
    } else {
      double d[3], wv[4];
      for (int t = 0; t < 3; ++t) d[t] = ao[gg + ng * (nu + nao * (t + 1))];
      for (int x = 0; x < 4; ++x) wv[x] = wg * vxc[gg + ng * x];
      const double htau = DEN == DEN_TAU ? 0.5 * wg * vxc[gg + ng * 4] : 0.0;
      S[e] = wv[0] * phi + wv[1] * d[0] + wv[2] * d[1] + wv[3] * d[2];
      for (int t = 0; t < 3; ++t) S[e + gb * nao * (t + 1)] = wv[t + 1] * phi + htau * d[t];
    }
  }
}

// part[b] = sum over the block's points of w exc rho
__global__ __launch_bounds__(256) void xc_energy_kernel(const double* __restrict__ w, const double* __restrict__ exc,
                                                        const double* __restrict__ rho, int64_t ng,
                                                        double* __restrict__ part) {
  double v = 0.0;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x)
    v += w[g] * exc[g] * rho[g];
  __shared__ double sd[256];
  sd[threadIdx.x] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) sd[threadIdx.x] += sd[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sd[0];
}
}  // namespace

void Numint::vxc_matrix(int den, const double* vxc_eff, double* out) {
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "numint: den_type must be RHO, SIGMA or TAU");
  const int nc = den_ncomp(den);
  HH_REQUIRE(ncomp_ >= nc, "numint: AO component dimension insufficient");
  HH_REQUIRE(vxc_eff && out, "numint: null argument");
  const int64_t gbmax = std::min<int64_t>(ng_, std::max<int64_t>(256, (1LL << 27) / (nao_ * nc)));
  double* S = static_cast<double*>(akb_.reserve(sizeof(double) * gbmax * nao_ * nc));
  for (int64_t g0 = 0; g0 < ng_; g0 += gbmax) {
    const int64_t gb = std::min(gbmax, ng_ - g0);
    HH_DEN_DISPATCH(den, vxc_s_kernel, dim3(nblk(gb * nao_)), dim3(256), 0, stream_, ao_.as<double>(), ng_, nao_, g0,
                    gb, vxc_eff, w_.as<double>(), S);
    back_contract(ao_.as<double>(), ng_, nao_, g0, gb, nc, S, nao_, out, 1.0, g0 > 0, PREC_F64, stream_, ws_);
  }
}

double Numint::xc_energy(const double* exc, const double* rho) {
  HH_REQUIRE(exc && rho, "numint: null argument");
  constexpr int NB = 256;
  const unsigned nb = (unsigned)std::min<int64_t>(NB, ceil_div(ng_, 256));
  double* part = static_cast<double*>(aob2_.reserve(sizeof(double) * NB));
  hipLaunchKernelGGL(xc_energy_kernel, dim3(nb), dim3(256), 0, stream_, w_.as<double>(), exc, rho, ng_, part);
  HH_HIP(hipGetLastError());
  std::vector<double> h(nb);
  HH_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, stream_));
  HH_HIP(hipStreamSynchronize(stream_));
  double e = 0.0;
  for (double v : h) e += v;
  return e;
}

void Numint::rks_response_bra(int den, const double* fxc_eff, const double* mo1_bra, int64_t nprop,
                              const double* mocc, int64_t nocc, double* out, Prec prec, bool accumulate) {
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "numint: den_type must be RHO, SIGMA or TAU");
  const int nc = den_ncomp(den);
  HH_REQUIRE(ncomp_ >= nc, "numint: AO component dimension insufficient");
  const bool bf = prec == PREC_BF16X3;
  const int64_t ncol = nocc * nprop;
  const int64_t gbmax = batch_for(ncol, bf);
  const size_t es = bf ? sizeof(float) : sizeof(double);
  double* aob = static_cast<double*>(aob_.reserve(sizeof(double) * gbmax * nocc * nc));
  // BF16X3: the kets product fused with the pointwise pass when a tile of whole perturbations
  // fits 128 columns (nocc <= 128): fna perturbations, fnfr 16-column fragments per tile
  const bool fused = bf && fused_kets && nocc <= 128;
  const int fna = fused ? (int)std::max<int64_t>(1, std::min<int64_t>(8, 96 / nocc)) : 0;
  const int fnfr = fused ? (int)ceil_div(fna * nocc, (int64_t)16) : 0;
  if (fused) HH_REQUIRE(ceil_div(nprop, (int64_t)fna) < 65536, "numint: too many perturbations for one launch");
  void* akb = fused ? nullptr : akb_.reserve(es * gbmax * ncol * nc);
  if (bf) {
    ensure_bf16();
    split_kets(mo1_bra, ncol, kh_, kl_);
  }
  for (int64_t g0 = 0; g0 < ng_; g0 += gbmax) {
    const int64_t gb = std::min(gbmax, ng_ - g0);
    half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, mocc, nocc, aob, PREC_F64, stream_, ws_);
    if (bf && !fused)
      kets_bf16(nc, g0, gb, kh_.as<unsigned short>(), kl_.as<unsigned short>(), ncol, static_cast<float*>(akb));
    else if (!bf)
      half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, mo1_bra, ncol, static_cast<double*>(akb), prec, stream_,
                     ws_);
    // the x4 closed-shell factor of hess_rks.rs:603
    if (bf) {
      unsigned short *h, *l;
      s_tiled(nc, gb, ncol, sh_, sl_, &h, &l);
      const int64_t cs = ceil_div(gb, 32) * 32 * ncol;
      if (fused) {
        const dim3 grid((unsigned)ceil_div(gb, 32), (unsigned)ceil_div(nprop, (int64_t)fna));
        launch_kets_rho1_s(den, fnfr, grid, stream_, aot_hi_.as<unsigned short>(), aot_lo_.as<unsigned short>(),
                           naop_ * ngp_, ngp_, kh_.as<unsigned short>(), kl_.as<unsigned short>(), ncol,
                           (int)(naop_ / 32), aob, fxc_eff, w_.as<double>(), h, l, cs, gb, nocc, nprop, fna, ng_, g0);
      } else {
        HH_DEN_DISPATCH(den, rho1_wv_s_tiled_kernel, dim3(nblk(gb * nprop)), dim3(256), 0, stream_, aob,
                        static_cast<const float*>(akb), h, l, cs, fxc_eff, w_.as<double>(), gb, nocc, nprop, ng_, g0);
      }
      back_bf16(nc, g0, gb, h, l, ncol, out, 4.0, accumulate || g0 > 0);
    } else {
      HH_DEN_DISPATCH_T(den, false, rho1_wv_s_kernel, dim3(nblk(gb * nprop)), dim3(256), 0, stream_, aob, akb,
                        fxc_eff, w_.as<double>(), gb, nocc, nprop, ng_, g0);
      back_contract(ao_.as<double>(), ng_, nao_, g0, gb, nc, static_cast<double*>(akb), ncol, out, 4.0,
                    accumulate || g0 > 0, prec, stream_, ws_);
    }
  }
}

void Numint::uks_fxc_pot_bra_trans(int den, const double* fxc_eff, const double* rho1, int64_t nprop,
                                   const double* const* bra, const int64_t* nocc, double* const* out, Prec prec,
                                   double alpha, bool accumulate) {
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "numint: den_type must be RHO, SIGMA or TAU");
  const int nc = den_ncomp(den);
  HH_REQUIRE(ncomp_ >= nc, "numint: AO component dimension insufficient");
  const bool bf = prec == PREC_BF16X3;
  const int64_t ncol_a = nocc[0] * nprop, ncol_b = nocc[1] * nprop;
  const int64_t gbmax = batch_for(ncol_a + ncol_b, bf);
  const size_t es = bf ? sizeof(float) : sizeof(double);
  double* aob_a = static_cast<double*>(aob_.reserve(sizeof(double) * gbmax * nocc[0] * nc));
  double* aob_b = static_cast<double*>(aob2_.reserve(sizeof(double) * gbmax * nocc[1] * nc));
  void* Sa = bf ? nullptr : akb_.reserve(es * gbmax * ncol_a * nc);
  void* Sb = bf ? nullptr : akb2_.reserve(es * gbmax * ncol_b * nc);
  if (bf) ensure_bf16();
  for (int64_t g0 = 0; g0 < ng_; g0 += gbmax) {
    const int64_t gb = std::min(gbmax, ng_ - g0);
    half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, bra[0], nocc[0], aob_a, PREC_F64, stream_, ws_);
    half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, bra[1], nocc[1], aob_b, PREC_F64, stream_, ws_);
    if (bf) {
      unsigned short *ha, *la, *hb, *lb;
      s_tiled(nc, gb, ncol_a, sh_, sl_, &ha, &la);
      s_tiled(nc, gb, ncol_b, sh2_, sl2_, &hb, &lb);
      const int64_t gbp = ceil_div(gb, 32) * 32;
      HH_DEN_DISPATCH(den, uks_wv_s_tiled_kernel, dim3(nblk(gb * nprop)), dim3(256), 0, stream_, aob_a, aob_b, rho1, ha,
                      la, gbp * ncol_a, hb, lb, gbp * ncol_b, fxc_eff, w_.as<double>(), gb, nocc[0], nocc[1], nprop,
                      ng_, g0);
      back_bf16(nc, g0, gb, ha, la, ncol_a, out[0], alpha, accumulate || g0 > 0);
      back_bf16(nc, g0, gb, hb, lb, ncol_b, out[1], alpha, accumulate || g0 > 0);
    } else {
      HH_DEN_DISPATCH_T(den, false, uks_wv_s_kernel, dim3(nblk(gb * nprop)), dim3(256), 0, stream_, aob_a, aob_b, rho1,
                        Sa, Sb, fxc_eff, w_.as<double>(), gb, nocc[0], nocc[1], nprop, ng_, g0);
      back_contract(ao_.as<double>(), ng_, nao_, g0, gb, nc, static_cast<double*>(Sa), ncol_a, out[0], alpha,
                    accumulate || g0 > 0, prec, stream_, ws_);
      back_contract(ao_.as<double>(), ng_, nao_, g0, gb, nc, static_cast<double*>(Sb), ncol_b, out[1], alpha,
                    accumulate || g0 > 0, prec, stream_, ws_);
    }
  }
}

void Numint::uks_response_bra(int den, const double* fxc_eff, const double* const* mo1_bra, int64_t nprop,
                              const double* const* mocc, const int64_t* nocc, double* const* out, Prec prec,
                              bool accumulate) {
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "numint: den_type must be RHO, SIGMA or TAU");
  const int nc = den_ncomp(den);
  HH_REQUIRE(ncomp_ >= nc, "numint: AO component dimension insufficient");
  const bool bf = prec == PREC_BF16X3;
  const int64_t ncol_a = nocc[0] * nprop, ncol_b = nocc[1] * nprop;
  const int64_t gbmax = batch_for(ncol_a + ncol_b, bf);
  const size_t es = bf ? sizeof(float) : sizeof(double);
  double* aob_a = static_cast<double*>(aob_.reserve(sizeof(double) * gbmax * nocc[0] * nc));
  double* aob_b = static_cast<double*>(aob2_.reserve(sizeof(double) * gbmax * nocc[1] * nc));
  void* akb_a = akb_.reserve(es * gbmax * ncol_a * nc);
  void* akb_b = akb2_.reserve(es * gbmax * ncol_b * nc);
  if (bf) {
    ensure_bf16();
    split_kets(mo1_bra[0], ncol_a, kh_, kl_);
    split_kets(mo1_bra[1], ncol_b, kh2_, kl2_);
  }
  for (int64_t g0 = 0; g0 < ng_; g0 += gbmax) {
    const int64_t gb = std::min(gbmax, ng_ - g0);
    half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, mocc[0], nocc[0], aob_a, PREC_F64, stream_, ws_);
    half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, mocc[1], nocc[1], aob_b, PREC_F64, stream_, ws_);
    if (bf) {
      kets_bf16(nc, g0, gb, kh_.as<unsigned short>(), kl_.as<unsigned short>(), ncol_a, static_cast<float*>(akb_a));
      kets_bf16(nc, g0, gb, kh2_.as<unsigned short>(), kl2_.as<unsigned short>(), ncol_b, static_cast<float*>(akb_b));
    } else {
      half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, mo1_bra[0], ncol_a, static_cast<double*>(akb_a), prec,
                     stream_, ws_);
      half_transform(ao_.as<double>(), ng_, nao_, g0, gb, nc, mo1_bra[1], ncol_b, static_cast<double*>(akb_b), prec,
                     stream_, ws_);
    }
    // the x2 factor of hess_uks.rs:436-439
    if (bf) {
      unsigned short *ha, *la, *hb, *lb;
      s_tiled(nc, gb, ncol_a, sh_, sl_, &ha, &la);
      s_tiled(nc, gb, ncol_b, sh2_, sl2_, &hb, &lb);
      const int64_t gbp = ceil_div(gb, 32) * 32;
      HH_DEN_DISPATCH(den, uks_rho1_wv_s_tiled_kernel, dim3(nblk(gb * nprop)), dim3(256), 0, stream_, aob_a,
                      static_cast<const float*>(akb_a), aob_b, static_cast<const float*>(akb_b), ha, la,
                      gbp * ncol_a, hb, lb, gbp * ncol_b, fxc_eff, w_.as<double>(), gb, nocc[0], nocc[1], nprop,
                      ng_, g0);
      back_bf16(nc, g0, gb, ha, la, ncol_a, out[0], 2.0, accumulate || g0 > 0);
      back_bf16(nc, g0, gb, hb, lb, ncol_b, out[1], 2.0, accumulate || g0 > 0);
    } else {
      HH_DEN_DISPATCH_T(den, false, uks_rho1_wv_s_kernel, dim3(nblk(gb * nprop)), dim3(256), 0, stream_, aob_a, akb_a,
                        aob_b, akb_b, fxc_eff, w_.as<double>(), gb, nocc[0], nocc[1], nprop, ng_, g0);
      back_contract(ao_.as<double>(), ng_, nao_, g0, gb, nc, static_cast<double*>(akb_a), ncol_a, out[0], 2.0,
                    accumulate || g0 > 0, prec, stream_, ws_);
      back_contract(ao_.as<double>(), ng_, nao_, g0, gb, nc, static_cast<double*>(akb_b), ncol_b, out[1], 2.0,
                    accumulate || g0 > 0, prec, stream_, ws_);
    }
  }
}

void xceff_transform_pol(int den, int order, const double* rho, int64_t ld_rho, int64_t ld_spin, const double* xc_val,
                         int64_t ld_xc, int64_t ng, double* out, hipStream_t s) {
  HH_REQUIRE(order == 1 || order == 2, "xceff: order must be 1 (vxc) or 2 (fxc)");
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "xceff: den_type must be RHO, SIGMA or TAU");
  if (ng == 0) return;
  if (den == DEN_RHO) {
    hipLaunchKernelGGL(xceff_lda_pol_kernel, dim3(nblk(ng)), dim3(256), 0, s, order, xc_val, ld_xc, ng, out);
  } else if (den == DEN_SIGMA) {
    hipLaunchKernelGGL(xceff_pol_kernel<DEN_SIGMA>, dim3(nblk(ng)), dim3(256), 0, s, order, rho, ld_rho, ld_spin, xc_val,
                       ld_xc, ng, out);
  } else {
    hipLaunchKernelGGL(xceff_pol_kernel<DEN_TAU>, dim3(nblk(ng)), dim3(256), 0, s, order, rho, ld_rho, ld_spin, xc_val,
                       ld_xc, ng, out);
  }
  HH_HIP(hipGetLastError());
}

std::vector<int64_t> libxc_to_xcfun_map(int den, int deriv) {
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "xceff: den_type must be RHO, SIGMA or TAU");
  HH_REQUIRE(deriv >= 0 && deriv <= 2, "xceff: the map is provided for derivative orders 0-2");
  std::vector<int> kinds = den == DEN_RHO ? std::vector<int>{2}
                           : den == DEN_SIGMA ? std::vector<int>{2, 3} : std::vector<int>{2, 3, 2};
  std::vector<std::pair<int, int>> flat;  // (kind, component)
  for (int k = 0; k < (int)kinds.size(); ++k)
    for (int c = 0; c < kinds[k]; ++c) flat.push_back({k, c});
  const int nf = (int)flat.size();
  std::vector<int64_t> out{0};
  if (deriv >= 1)
    for (int i = 0; i < nf; ++i) out.push_back(1 + i);
  if (deriv >= 2) {
    // LibXC position of each (kind1, comp1, kind2, comp2) second-derivative entry
    std::vector<int64_t> pos(nf * nf, -1);
    auto flat_index = [&](int k, int c) {
      int i = 0;
      for (int kk = 0; kk < k; ++kk) i += kinds[kk];
      return i + c;
    };
    int64_t off = 1 + nf;
    for (int k1 = 0; k1 < (int)kinds.size(); ++k1)
      for (int k2 = k1; k2 < (int)kinds.size(); ++k2)
        for (int c1 = 0; c1 < kinds[k1]; ++c1)
          for (int c2 = (k1 == k2 ? c1 : 0); c2 < kinds[k2]; ++c2)
            pos[flat_index(k1, c1) * nf + flat_index(k2, c2)] = off++;
    for (int i = 0; i < nf; ++i)
      for (int j = i; j < nf; ++j) out.push_back(pos[i * nf + j]);
  }
  return out;
}

void xceff_transform(int den, int order, const double* rho, int64_t ld_rho, const double* xc_val, int64_t ld_xc,
                     int64_t ng, double* out, hipStream_t s) {
  HH_REQUIRE(order == 1 || order == 2, "xceff: order must be 1 (vxc) or 2 (fxc)");
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "xceff: den_type must be RHO, SIGMA or TAU");
  if (ng == 0) return;
  if (den == DEN_RHO) {
    hipLaunchKernelGGL(xceff_lda_kernel, dim3(nblk(ng)), dim3(256), 0, s, order, xc_val, ld_xc, ng, out);
  } else if (den == DEN_SIGMA) {
    hipLaunchKernelGGL(xceff_kernel<DEN_SIGMA>, dim3(nblk(ng)), dim3(256), 0, s, order, rho, ld_rho, xc_val, ld_xc, ng,
                       out);
  } else {
    hipLaunchKernelGGL(xceff_kernel<DEN_TAU>, dim3(nblk(ng)), dim3(256), 0, s, order, rho, ld_rho, xc_val, ld_xc, ng,
                       out);
  }
  HH_HIP(hipGetLastError());
}

}  // namespace hh
