// Strided batched split-K GEMM on gfx950 matrix cores.  See gemm.hpp for the contract.
//
// Work decomposition: a workgroup of 256 lanes (4 wave64) owns a BM x BN tile of C and walks
// its K range in BK steps.  Waves form a 2 x 2 grid; each wave owns (BM/2) x (BN/2) as
// 16 x 16 MFMA blocks.  Tiles are staged global -> registers -> LDS, with the next K step's
// global loads issued before the current step's MFMAs (register double buffering).
//
// MFMA 16x16 operand/result layout (wave64, lane l):
//   4-deep  (f64/f32): A frag = A[l%16][l/16],            B frag = B[l/16][l%16]
//   32-deep (bf16)   : A frag = A[l%16][8*(l/16) .. +7],  B frag = B[8*(l/16) .. +7][l%16]
//   result f32/bf16  : acc[r] = D[4*(l/16) + r][l%16],  r = 0..3
//   result f64       : acc[r] = D[(l/16) + 4*r][l%16]   (rows interleaved; measured on gfx950)
// A row fragment and a B column fragment have the same lane layout, so the kernel issues
// mfma(B-frag, A-frag) and accumulates C^T blocks: lane l then holds C[m = l%16][n = ...], and
// the column-major C (m contiguous) is stored by 16 consecutive lanes at consecutive addresses.
#include "gemm.hpp"

namespace hh {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

struct GemmParams {
  int64_t M, N, K;
  const void* A;
  int64_t lda;
  const void* B;
  int64_t ldb;
  void* C;
  int64_t ldc;
  double alpha, beta;
  int64_t nb0;
  int64_t sA0, sA1, sB0, sB1, sC0, sC1;
  int64_t tiles_m, tiles_n;
  int64_t kchunk;  // K elements per split (multiple of BK)
  int ksplit;
  void* partial;   // [ksplit][nbatch][N][M] accumulator-typed partial sums (ksplit > 1)
  int64_t K1;      // second operand pair for k >= K1 (K1 >= K: none)
  const void* A2;
  const void* B2;
  int64_t lda2, ldb2, sA2_0, sA2_1, sB2_0, sB2_1;
  void* C_lo;      // non-null: write C as a bf16 hi (C) / lo (C_lo) pair of u16 arrays
  const int64_t* krange;  // non-null: K range [krange[2 b1], krange[2 b1 + 1]) per second batch index
};

template <int PREC> struct PrecTraits;
template <> struct PrecTraits<PREC_F64> {
  typedef double acc_t;
  typedef v4d vacc_t;
};
template <> struct PrecTraits<PREC_F32> {
  typedef float acc_t;
  typedef v4f vacc_t;
};
template <> struct PrecTraits<PREC_BF16X3> {
  typedef float acc_t;
  typedef v4f vacc_t;
};

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

template <int PREC, typename TIn, typename TOut, int BM, int BN, int BK, bool AK, bool BKM>
__global__ __launch_bounds__(256) void gemm_kernel(GemmParams p) {
  typedef typename PrecTraits<PREC>::acc_t acc_t;
  typedef typename PrecTraits<PREC>::vacc_t vacc_t;
  typedef typename std::conditional<PREC == PREC_F64, double, float>::type ct;  // staging type
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int NA = BM * BK / 256, NB = BN * BK / 256;
  static_assert(BM * BK % 256 == 0 && BN * BK % 256 == 0, "tile/thread mismatch");
  constexpr bool BF = (PREC == PREC_BF16X3);
  // LDS layouts: f64/f32 stage [k][m] (m contiguous); bf16 stages [m][k] (k contiguous) hi + lo.
  constexpr int PADF = 4;  // elements
  constexpr int PADB = 8;  // bf16 elements (keeps 16-byte row alignment)
  constexpr int LDA_F = BM + PADF, LDB_F = BN + PADF;
  constexpr int LDK_B = BK + PADB;
  constexpr size_t SMEM_F = sizeof(ct) * (size_t)BK * (LDA_F + LDB_F);
  constexpr size_t SMEM_B = sizeof(__bf16) * 2 * (size_t)LDK_B * (BM + BN);
  constexpr size_t SMEM = BF ? SMEM_B : SMEM_F;
  __shared__ __attribute__((aligned(16))) unsigned char smem[SMEM];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  // tile / batch / split coordinates
  const int64_t tile = blockIdx.x;
  const int64_t tm = tile % p.tiles_m, tn = tile / p.tiles_m;
  const int64_t bidx = blockIdx.y;
  const int64_t b0 = bidx % p.nb0, b1 = bidx / p.nb0;
  const int kz = blockIdx.z;
  const int64_t m0 = tm * BM, n0 = tn * BN;
  const int64_t kbeg = (p.krange ? max((int64_t)0, p.krange[2 * b1]) : 0) + (int64_t)kz * p.kchunk;
  const int64_t kend = min(p.krange ? min(p.K, p.krange[2 * b1 + 1]) : p.K, kbeg + p.kchunk);

  const TIn* __restrict__ A = static_cast<const TIn*>(p.A) + b0 * p.sA0 + b1 * p.sA1;
  const TIn* __restrict__ B = static_cast<const TIn*>(p.B) + b0 * p.sB0 + b1 * p.sB1;
  const bool dual = p.K1 < p.K;
  const TIn* __restrict__ A2 = dual ? static_cast<const TIn*>(p.A2) + b0 * p.sA2_0 + b1 * p.sA2_1 : nullptr;
  const TIn* __restrict__ B2 = dual ? static_cast<const TIn*>(p.B2) + b0 * p.sB2_0 + b1 * p.sB2_1 : nullptr;

  vacc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = vacc_t{0, 0, 0, 0};

  ct ra[NA], rb[NB];

  auto load_tiles = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      const int idx = tid + 256 * e;
      const int mm = AK ? idx / BK : idx % BM;
      const int kk = AK ? idx % BK : idx / BM;
      const int64_t gm = m0 + mm, gk = k0 + kk;
      ct v = 0;
      if (gm < p.M && gk < kend) {
        if (!dual || gk < p.K1) v = (ct)(AK ? A[gm * p.lda + gk] : A[gm + gk * p.lda]);
        else {
          const int64_t k2 = gk - p.K1;
          v = (ct)(AK ? A2[gm * p.lda2 + k2] : A2[gm + k2 * p.lda2]);
        }
      }
      ra[e] = v;
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int idx = tid + 256 * e;
      const int nn = BKM ? idx / BK : idx % BN;
      const int kk = BKM ? idx % BK : idx / BN;
      const int64_t gn = n0 + nn, gk = k0 + kk;
      ct v = 0;
      if (gn < p.N && gk < kend) {
        if (!dual || gk < p.K1) v = (ct)(BKM ? B[gk + gn * p.ldb] : B[gk * p.ldb + gn]);
        else {
          const int64_t k2 = gk - p.K1;
          v = (ct)(BKM ? B2[k2 + gn * p.ldb2] : B2[k2 * p.ldb2 + gn]);
        }
      }
      rb[e] = v;
    }
  };

  auto store_tiles = [&]() {
    if constexpr (!BF) {
      ct* As = reinterpret_cast<ct*>(smem);
      ct* Bs = As + BK * LDA_F;
#pragma unroll
      for (int e = 0; e < NA; ++e) {
        const int idx = tid + 256 * e;
        const int mm = AK ? idx / BK : idx % BM;
        const int kk = AK ? idx % BK : idx / BM;
        As[kk * LDA_F + mm] = ra[e];
      }
#pragma unroll
      for (int e = 0; e < NB; ++e) {
        const int idx = tid + 256 * e;
        const int nn = BKM ? idx / BK : idx % BN;
        const int kk = BKM ? idx % BK : idx / BN;
        Bs[kk * LDB_F + nn] = rb[e];
      }
    } else {
      __bf16* Ah = reinterpret_cast<__bf16*>(smem);
      __bf16* Al = Ah + BM * LDK_B;
      __bf16* Bh = Al + BM * LDK_B;
      __bf16* Bl = Bh + BN * LDK_B;
#pragma unroll
      for (int e = 0; e < NA; ++e) {
        const int idx = tid + 256 * e;
        const int mm = AK ? idx / BK : idx % BM;
        const int kk = AK ? idx % BK : idx / BM;
        __bf16 h, l;
        split_bf16((float)ra[e], h, l);
        Ah[mm * LDK_B + kk] = h;
        Al[mm * LDK_B + kk] = l;
      }
#pragma unroll
      for (int e = 0; e < NB; ++e) {
        const int idx = tid + 256 * e;
        const int nn = BKM ? idx / BK : idx % BN;
        const int kk = BKM ? idx % BK : idx / BN;
        __bf16 h, l;
        split_bf16((float)rb[e], h, l);
        Bh[nn * LDK_B + kk] = h;
        Bl[nn * LDK_B + kk] = l;
      }
    }
  };

  auto compute_tiles = [&]() {
    if constexpr (!BF) {
      const ct* As = reinterpret_cast<const ct*>(smem);
      const ct* Bs = As + BK * LDA_F;
#pragma unroll
      for (int ks = 0; ks < BK / 4; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
        ct af[FM], bfr[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i] = As[kr * LDA_F + wm * WM + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < FN; ++j) bfr[j] = Bs[kr * LDB_F + wn * WN + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            if constexpr (PREC == PREC_F64)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bfr[j], af[i], acc[i][j], 0, 0, 0);
            else
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bfr[j], af[i], acc[i][j], 0, 0, 0);
          }
      }
    } else {
      const __bf16* Ah = reinterpret_cast<const __bf16*>(smem);
      const __bf16* Al = Ah + BM * LDK_B;
      const __bf16* Bh = Al + BM * LDK_B;
      const __bf16* Bl = Bh + BN * LDK_B;
#pragma unroll
      for (int ks = 0; ks < BK / 32; ++ks) {
        const int kc = ks * 32 + 8 * (lane >> 4);
        v8bf ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          const int r = wm * WM + i * 16 + (lane & 15);
          ah[i] = *reinterpret_cast<const v8bf*>(Ah + r * LDK_B + kc);
          al[i] = *reinterpret_cast<const v8bf*>(Al + r * LDK_B + kc);
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const int c = wn * WN + j * 16 + (lane & 15);
          bh[j] = *reinterpret_cast<const v8bf*>(Bh + c * LDK_B + kc);
          bl[j] = *reinterpret_cast<const v8bf*>(Bl + c * LDK_B + kc);
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al[i], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah[i], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah[i], acc[i][j], 0, 0, 0);
          }
      }
    }
  };

  if (kbeg < kend) {
    load_tiles(kbeg);
    for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
      __syncthreads();  // previous compute done with LDS
      store_tiles();
      __syncthreads();
      if (k0 + BK < kend) load_tiles(k0 + BK);
      compute_tiles();
    }
  }

  // epilogue: acc[i][j][r] = C[mrow0 + 16 i][ncol0 + 16 j + RSTRIDE r]  (transposed blocks)
  const int64_t mrow0 = m0 + wm * WM + (lane & 15);
  constexpr int RSTRIDE = (PREC == PREC_F64) ? 4 : 1;  // column step between acc[r] entries
  const int64_t ncol0 = n0 + wn * WN + ((PREC == PREC_F64) ? (lane >> 4) : 4 * (lane >> 4));
  if (p.ksplit > 1) {
    const int64_t nbatch = (int64_t)gridDim.y;
    acc_t* P = static_cast<acc_t*>(p.partial) + ((int64_t)kz * nbatch + bidx) * p.M * p.N;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t gm = mrow0 + i * 16, gn = ncol0 + j * 16 + r * RSTRIDE;
          if (gm < p.M && gn < p.N) P[gm + gn * p.M] = acc[i][j][r];
        }
  } else {
    TOut* C = static_cast<TOut*>(p.C) + b0 * p.sC0 + b1 * p.sC1;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t gm = mrow0 + i * 16, gn = ncol0 + j * 16 + r * RSTRIDE;
          if (gm < p.M && gn < p.N) {
            double v = p.alpha * (double)acc[i][j][r];
            const int64_t off = gm + gn * p.ldc;
            if (p.C_lo) {
              // split output: hi = bf16(v), lo = bf16(v - hi) into two u16 arrays (beta ignored)
              const float f = (float)v;
              const __bf16 h = (__bf16)f;
              const __bf16 l = (__bf16)(f - (float)h);
              static_cast<unsigned short*>(p.C)[off + (b0 * p.sC0 + b1 * p.sC1)] = __builtin_bit_cast(unsigned short, h);
              static_cast<unsigned short*>(p.C_lo)[off + (b0 * p.sC0 + b1 * p.sC1)] = __builtin_bit_cast(unsigned short, l);
            } else {
              TOut* c = C + off;
              if (p.beta != 0.0) v += p.beta * (double)(*c);
              *c = (TOut)v;
            }
          }
        }
  }
}

// C = alpha * sum_z partial[z] + beta * C, summed in a fixed order in FP64.
template <typename acc_t, typename TOut>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const acc_t* __restrict__ partial, int ksplit,
                                                            int64_t nbatch, int64_t M, int64_t N, int64_t nb0,
                                                            TOut* C, int64_t ldc, int64_t sC0, int64_t sC1,
                                                            double alpha, double beta) {
  const int64_t MN = M * N;
  const int64_t total = MN * nbatch;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t / MN, e = t % MN;
    const int64_t m = e % M, n = e / M;
    double s = 0.0;
    for (int z = 0; z < ksplit; ++z) s += (double)partial[((int64_t)z * nbatch + b) * MN + e];
    const int64_t b0 = b % nb0, b1 = b / nb0;
    TOut* c = C + b0 * sC0 + b1 * sC1 + m + n * ldc;
    double v = alpha * s;
    if (beta != 0.0) v += beta * (double)(*c);
    *c = (TOut)v;
  }
}

namespace {

int num_cus() {
  static int n = 0;
  if (n == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
      n = 256;
  }
  return n;
}

template <int PREC, typename TIn, typename TOut, int BM, int BN, int BK>
void launch_cfg(const GemmArgs& g, hipStream_t stream, DevBuf& ws) {
  typedef typename PrecTraits<PREC>::acc_t acc_t;
  GemmParams p{};
  p.M = g.M; p.N = g.N; p.K = g.K;
  p.A = g.A; p.lda = g.lda; p.B = g.B; p.ldb = g.ldb; p.C = g.C; p.ldc = g.ldc;
  p.alpha = g.alpha; p.beta = g.beta;
  p.nb0 = g.nb0;
  p.sA0 = g.sA0; p.sA1 = g.sA1; p.sB0 = g.sB0; p.sB1 = g.sB1; p.sC0 = g.sC0; p.sC1 = g.sC1;
  p.K1 = (g.K1 >= 0 && g.K1 < g.K) ? g.K1 : g.K;
  p.A2 = g.A2; p.B2 = g.B2; p.lda2 = g.lda2; p.ldb2 = g.ldb2;
  p.sA2_0 = g.sA2_0; p.sA2_1 = g.sA2_1; p.sB2_0 = g.sB2_0; p.sB2_1 = g.sB2_1;
  if (p.K1 < p.K) HH_REQUIRE(g.A2 && g.B2, "gemm: second operand pair missing");
  p.C_lo = g.C_lo;
  if (g.C_lo) HH_REQUIRE(g.ksplit == 1, "gemm: split bf16 output needs ksplit = 1");
  p.krange = g.krange;
  if (g.krange) HH_REQUIRE(p.K1 >= p.K && g.krange_max >= 0 && g.krange_max <= g.K, "gemm: grouped K ranges need "
                                                                                  "one operand pair and 0 <= krange_max <= K");
  const int64_t klen = g.krange ? std::max<int64_t>(g.krange_max, 1) : g.K;  // K length one split covers
  p.tiles_m = ceil_div(g.M, BM);
  p.tiles_n = ceil_div(g.N, BN);
  const int64_t nbatch = g.nb0 * g.nb1;
  const int64_t tiles = p.tiles_m * p.tiles_n * nbatch;
  int ksplit = g.ksplit;
  if (ksplit <= 0) {
    // enough workgroups to cover the chip ~4x, but keep >= 8 K-steps per split
    const int64_t target = 4LL * num_cus();
    ksplit = 1;
    while (tiles * ksplit < target && ceil_div(klen, (int64_t)BK * ksplit * 2) >= 8) ksplit *= 2;
  }
  int64_t kchunk = ceil_div(ceil_div(klen, ksplit), BK) * BK;
  if (kchunk <= 0) kchunk = BK;
  ksplit = (int)ceil_div(std::max<int64_t>(klen, 1), kchunk);
  p.kchunk = kchunk;
  p.ksplit = ksplit;
  if (ksplit > 1) p.partial = ws.reserve(sizeof(acc_t) * (size_t)ksplit * nbatch * g.M * g.N);
  HH_REQUIRE(p.tiles_m * p.tiles_n < (1LL << 31) && nbatch < 65536 && ksplit < 65536, "gemm grid too large");
  dim3 grid((unsigned)(p.tiles_m * p.tiles_n), (unsigned)nbatch, (unsigned)ksplit);
  if (g.ev_begin) HH_HIP(hipEventRecord(g.ev_begin, stream));
  if (g.a_kmajor && g.b_kmajor)
    hipLaunchKernelGGL((gemm_kernel<PREC, TIn, TOut, BM, BN, BK, true, true>), grid, dim3(256), 0, stream, p);
  else if (g.a_kmajor && !g.b_kmajor)
    hipLaunchKernelGGL((gemm_kernel<PREC, TIn, TOut, BM, BN, BK, true, false>), grid, dim3(256), 0, stream, p);
  else if (!g.a_kmajor && g.b_kmajor)
    hipLaunchKernelGGL((gemm_kernel<PREC, TIn, TOut, BM, BN, BK, false, true>), grid, dim3(256), 0, stream, p);
  else
    hipLaunchKernelGGL((gemm_kernel<PREC, TIn, TOut, BM, BN, BK, false, false>), grid, dim3(256), 0, stream, p);
  HH_HIP(hipGetLastError());
  if (g.ev_main_end) HH_HIP(hipEventRecord(g.ev_main_end, stream));
  if (ksplit > 1) {
    const int64_t total = nbatch * g.M * g.N;
    const unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(total, 256), 8192);
    hipLaunchKernelGGL((splitk_reduce_kernel<acc_t, TOut>), dim3(blocks), dim3(256), 0, stream,
                       static_cast<const acc_t*>(p.partial), ksplit, nbatch, g.M, g.N, g.nb0,
                       static_cast<TOut*>(g.C), g.ldc, g.sC0, g.sC1, g.alpha, g.beta);
    HH_HIP(hipGetLastError());
  }
  if (g.ev_end) HH_HIP(hipEventRecord(g.ev_end, stream));
}

template <int PREC, typename TIn, typename TOut>
void dispatch_tiles(const GemmArgs& g, hipStream_t stream, DevBuf& ws) {
  constexpr int BK = (PREC == PREC_BF16X3) ? 32 : 16;
  // small problems (or thin M/N) use 64x64 tiles, large ones 128x128
  if (g.M >= 128 && g.N >= 128 && g.M * g.N * g.nb0 * g.nb1 >= (int64_t)128 * 128 * 256)
    launch_cfg<PREC, TIn, TOut, 128, 128, BK>(g, stream, ws);
  else
    launch_cfg<PREC, TIn, TOut, 64, 64, BK>(g, stream, ws);
}

}  // namespace

template <typename TIn, typename TOut>
void gemm(const GemmArgs& g, Prec prec, hipStream_t stream, DevBuf& ws) {
  HH_REQUIRE(g.M >= 0 && g.N >= 0 && g.K >= 0, "gemm: negative size");
  HH_REQUIRE(g.nb0 >= 1 && g.nb1 >= 1, "gemm: empty batch");
  if (g.M == 0 || g.N == 0) return;
  HH_REQUIRE(g.C && ((g.A && g.B) || g.K == 0), "gemm: null operand");
  switch (prec) {
    case PREC_F64:
      if constexpr (std::is_same<TIn, double>::value) {
        dispatch_tiles<PREC_F64, TIn, TOut>(g, stream, ws);
        return;
      } else {
        throw Error(HH_ERR_INVALID, "gemm: FP64 mode needs FP64 operands");
      }
    case PREC_F32: dispatch_tiles<PREC_F32, TIn, TOut>(g, stream, ws); return;
    case PREC_BF16X3: dispatch_tiles<PREC_BF16X3, TIn, TOut>(g, stream, ws); return;
  }
  throw Error(HH_ERR_INVALID, "gemm: unknown precision mode");
}

template void gemm<double, double>(const GemmArgs&, Prec, hipStream_t, DevBuf&);
template void gemm<double, float>(const GemmArgs&, Prec, hipStream_t, DevBuf&);
template void gemm<float, float>(const GemmArgs&, Prec, hipStream_t, DevBuf&);
template void gemm<float, double>(const GemmArgs&, Prec, hipStream_t, DevBuf&);

}  // namespace hh
