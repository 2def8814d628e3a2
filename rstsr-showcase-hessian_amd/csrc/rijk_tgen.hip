// Generation of the RI-K hot operand T in BF16x3 storage (see rijk_tgen.hpp).
//
// T_a = G_a H_a with G_a = [bra_a | C_o] (nao x 2 nocc) and H_a = [O ; W_a] (2 nocc x nc): a GEMM
// with a short inner dimension whose cost is writing T (4 bytes per element as the bf16 pair),
// so the kernel is organised around the stores:
//   - a workgroup (4 waves) owns 4 * 16 * NF consecutive columns (P, i), i fastest, and a group
//     of perturbations;
//     per perturbation each lane builds its H fragments once (split into bf16 hi / lo) and
//     keeps them in registers while the wave sweeps all rows of G_a;
//   - G_a is packed per chunk as [a][row][K] bf16 hi / lo; the workgroup stages each row group
//     in LDS once (double-buffered, one LDS-only barrier per group) and the four waves read
//     their A fragments from there;
//   - three MFMA 16x16x32 bf16 products per fragment pair (lo*hi + hi*lo + hi*hi, FP32
//     accumulation), the split-bf16 scheme of the hot contraction;
//   - results are split into hi / lo and staged through LDS in groups of 16 RT rows, so each lane
//     stores 16 bytes and every column's 32 RT-byte segment goes out contiguously; the next
//     group's G rows are loaded before the current group's stores, and no barrier waits for
//     global stores.
#include "rijk_tgen.hpp"

namespace hh {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef unsigned short u16;

#ifndef TG_WAVES
#define TG_WAVES __attribute__((amdgpu_waves_per_eu(2)))
#endif
#ifndef TG_RT
#define TG_RT 2  // MFMA row tiles (16 rows each) per staged row group
#endif

__device__ __forceinline__ u16 bf_bits(__bf16 h) { return __builtin_bit_cast(u16, h); }

// (x0, x1) -> packed bf16 pairs hi = bf16(x), lo = bf16(x - hi), element 0 in the low half
__device__ __forceinline__ void split2(float x0, float x1, unsigned& hi, unsigned& lo) {
  const v2f x = {x0, x1};
  const v2bf h = __builtin_convertvector(x, v2bf);
  const v2bf l = __builtin_convertvector(x - __builtin_convertvector(h, v2f), v2bf);
  hi = __builtin_bit_cast(unsigned, h);
  lo = __builtin_bit_cast(unsigned, l);
}

// Orders this wave's LDS writes before its later LDS reads (and the reads before the next
// writes).  Each wave stages through its own LDS area, and LDS instructions of one wave execute
// in order, so no workgroup barrier is needed: a workgroup-scope fence would also wait for the
// wave's outstanding global stores and expose their latency once per row group.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// gh/gl[(al * mrows + m) * kp + kk] = split([bra[:, :, a0+al] | Co | 0][m, kk]), zero rows m >= nao
__global__ void tgen_pack_g_kernel(const float* __restrict__ bra, const float* __restrict__ Co, int64_t nao,
                                   int64_t nocc, int64_t a0, int64_t na, int64_t mrows, int kp, u16* __restrict__ gh,
                                   u16* __restrict__ gl) {
  const int64_t total = na * mrows * kp;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t kk = e % kp, m = (e / kp) % mrows, al = e / (kp * mrows);
    float x = 0.f;
    if (m < nao) {
      if (kk < nocc) x = bra[m + nao * (kk + nocc * (a0 + al))];
      else if (kk < 2 * nocc) x = Co[m + nao * (kk - nocc)];
    }
    const __bf16 h = (__bf16)x;
    gh[e] = bf_bits(h);
    gl[e] = bf_bits((__bf16)(x - (float)h));
  }
}

struct TgenParams {
  const u16* gh;  // packed G hi [na][mrows][KP]
  const u16* gl;  // packed G lo
  const float* O;
  const float* W;
  int64_t nocc, naux, ldw, nc, a0, na, ag, mrows, kbp, nn;
  u16* th;  // T hi, K-block tiled: [(P kbp + mu / 32) nn + i + nocc al] 32 + mu % 32
  u16* tl;
};

// Workgroup barrier ordering LDS only: global stores stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int KS, int NF, int RT, bool VEC>
__global__ __launch_bounds__(256) TG_WAVES void tgen_bf16x3_kernel(TgenParams p) {
  static_assert(RT == 2, "one staged row group is one K block of 32 rows");
  constexpr int KP = 32 * KS, NCW = 16 * NF, NCB = 4 * NCW;
  constexpr int ROWS = 16 * RT, LD = ROWS + 8;  // staged T column stride: 16-byte aligned rows
  constexpr int GLD = KP + 8;                   // staged G row stride: conflict-free 16-byte reads
  constexpr int GCH = ROWS * KP / 8;            // 16-byte chunks of one G array per row group
  constexpr int GPT = (2 * GCH + 255) / 256;    // chunks per thread (hi and lo arrays)
  __shared__ __attribute__((aligned(16))) u16 gbuf[2][2][ROWS * GLD];  // [buffer][hi, lo]
  __shared__ __attribute__((aligned(16))) u16 stage[4][2][NCW * LD];   // [wave][hi, lo]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t c0 = (int64_t)blockIdx.x * NCB + wave * NCW;
  const int64_t al_beg = (int64_t)blockIdx.y * p.ag;
  const int64_t al_end = min(p.na, al_beg + p.ag);
  const int64_t ng = (p.mrows + ROWS - 1) / ROWS;
  const int64_t nsteps = (al_end - al_beg) * ng;  // (perturbation, row group) steps, uniform
  u16* sh = stage[wave][0];
  u16* sl = stage[wave][1];

  // G rows of one step: global -> registers -> LDS buffer.  Two register sets: the rows of step
  // st + 2 are requested while step st computes, so each load has two steps to arrive.
  uint4 greg0[GPT], greg1[GPT];
  auto gload = [&](int64_t st, uint4 (&greg)[GPT]) {
    const int64_t al = al_beg + st / ng, mg = (st % ng) * ROWS;
    const int64_t nrows = min((int64_t)ROWS, p.mrows - mg);
    const u16* src0 = p.gh + (al * p.mrows + mg) * KP;
    const u16* src1 = p.gl + (al * p.mrows + mg) * KP;
#pragma unroll
    for (int e = 0; e < GPT; ++e) {
      const int q = tid + 256 * e;
      uint4 v = {0, 0, 0, 0};
      if (q < 2 * GCH) {
        const int arr = q / GCH, qq = q % GCH;
        if (qq / (KP / 8) < nrows) v = *reinterpret_cast<const uint4*>((arr ? src1 : src0) + qq * 8);
      }
      greg[e] = v;
    }
  };
  auto gstore = [&](int buf, const uint4 (&greg)[GPT]) {
#pragma unroll
    for (int e = 0; e < GPT; ++e) {
      const int q = tid + 256 * e;
      if (q < 2 * GCH) {
        const int arr = q / GCH, qq = q % GCH;
        const int row = qq / (KP / 8), kk = (qq % (KP / 8)) * 8;
        *reinterpret_cast<uint4*>(&gbuf[buf][arr][row * GLD + kk]) = greg[e];
      }
    }
  };

  if (nsteps > 0) {
    gload(0, greg0);
    gstore(0, greg0);
  }
  if (nsteps > 1) gload(1, greg1);
  lds_barrier();
  // H fragments: column c' = c0 + 16 nf + lane % 16, i.e. (P, i) = (c' / nocc, c' % nocc), the
  // column P + naux i of O and W; K entries 32 ks + 8 (lane / 16) .. +7
  uint4 bh[NF][KS], bl[NF][KS];
  // one (perturbation, row group) step; `ahead` receives step st + 2, `next` holds step st + 1
  auto step = [&](int64_t st, uint4 (&ahead)[GPT], const uint4 (&next)[GPT]) {
    const int64_t al = al_beg + st / ng, mg = (st % ng) * ROWS;
    const int buf = (int)(st & 1);
    if (mg == 0) {
      const int64_t a = p.a0 + al;
#pragma unroll
      for (int nf = 0; nf < NF; ++nf) {
        const int64_t cq = c0 + nf * 16 + (lane & 15);
        const bool live = cq < p.nc;
        const int64_t P = cq / p.nocc;
        const int64_t c = P + p.naux * (cq - P * p.nocc);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const int64_t k0 = ks * 32 + 8 * (lane >> 4);
          float x[8];
          // VEC (nocc % 8 == 0): the 8 entries come from one source, 32-byte aligned
          if (VEC && live && k0 < 2 * p.nocc) {
            const float* src = k0 < p.nocc ? p.O + k0 + p.nocc * c : p.W + (k0 - p.nocc) + p.nocc * a + p.ldw * c;
            const float4 u = *reinterpret_cast<const float4*>(src);
            const float4 v = *reinterpret_cast<const float4*>(src + 4);
            x[0] = u.x; x[1] = u.y; x[2] = u.z; x[3] = u.w;
            x[4] = v.x; x[5] = v.y; x[6] = v.z; x[7] = v.w;
          } else if (VEC) {
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] = 0.f;
          } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
              const int64_t kk = k0 + e;
              float v = 0.f;
              if (live) {
                if (kk < p.nocc) v = p.O[kk + p.nocc * c];
                else if (kk < 2 * p.nocc) v = p.W[(kk - p.nocc) + p.nocc * a + p.ldw * c];
              }
              x[e] = v;
            }
          }
          split2(x[0], x[1], bh[nf][ks].x, bl[nf][ks].x);
          split2(x[2], x[3], bh[nf][ks].y, bl[nf][ks].y);
          split2(x[4], x[5], bh[nf][ks].z, bl[nf][ks].z);
          split2(x[6], x[7], bh[nf][ks].w, bl[nf][ks].w);
        }
      }
    }
    // G rows two steps ahead go out before this step's stores: waiting for loads later leaves
    // the stores in flight (the vector memory counter retires in issue order)
    if (st + 2 < nsteps) gload(st + 2, ahead);

    const int nmt = (int)min((int64_t)RT, (p.mrows - mg) / 16);
    v4f acc[RT][NF];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
#pragma unroll
      for (int nf = 0; nf < NF; ++nf) acc[t][nf] = v4f{0.f, 0.f, 0.f, 0.f};
      if (t < nmt) {
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const int off = (16 * t + (lane & 15)) * GLD + ks * 32 + 8 * (lane >> 4);
          const v8bf ah = *reinterpret_cast<const v8bf*>(&gbuf[buf][0][off]);
          const v8bf alo = *reinterpret_cast<const v8bf*>(&gbuf[buf][1][off]);
#pragma unroll
          for (int nf = 0; nf < NF; ++nf) {
            const v8bf bhv = __builtin_bit_cast(v8bf, bh[nf][ks]);
            const v8bf blv = __builtin_bit_cast(v8bf, bl[nf][ks]);
            acc[t][nf] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bhv, acc[t][nf], 0, 0, 0);
            acc[t][nf] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, blv, acc[t][nf], 0, 0, 0);
            acc[t][nf] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bhv, acc[t][nf], 0, 0, 0);
          }
        }
      }
    }
    // stage: the lane holds D[row 16 t + 4 (lane / 16) + r][col 16 nf + lane % 16]
#pragma unroll
    for (int t = 0; t < RT; ++t) {
      if (t < nmt) {
#pragma unroll
        for (int nf = 0; nf < NF; ++nf) {
          const int col = nf * 16 + (lane & 15), row = 16 * t + 4 * (lane >> 4);
          uint2 hv, lv;
          split2(acc[t][nf][0], acc[t][nf][1], hv.x, lv.x);
          split2(acc[t][nf][2], acc[t][nf][3], hv.y, lv.y);
          *reinterpret_cast<uint2*>(sh + col * LD + row) = hv;
          *reinterpret_cast<uint2*>(sl + col * LD + row) = lv;
        }
      }
    }
    wave_sync();
    // write out: column col, 16-byte chunk ch (8 rows); 4 lanes cover a column's 32 rows, and
    // consecutive columns of one P are consecutive rows n = i + nocc al of the tiled T
    {
      constexpr int CPC = 2 * RT;
      const int cpc = 2 * nmt;
      const int64_t kb0 = mg / ROWS;
#pragma unroll
      for (int it = 0; it < NCW * CPC / 64; ++it) {
        const int idx = it * 64 + lane;
        const int col = idx / CPC, ch = idx % CPC;
        const int64_t cq = c0 + col;
        if (ch < cpc && cq < p.nc) {
          const int64_t P = cq / p.nocc;
          const int64_t goff = ((P * p.kbp + kb0) * p.nn + (cq - P * p.nocc) + p.nocc * al) * 32 + ch * 8;
          *reinterpret_cast<uint4*>(p.th + goff) = *reinterpret_cast<const uint4*>(sh + col * LD + ch * 8);
          *reinterpret_cast<uint4*>(p.tl + goff) = *reinterpret_cast<const uint4*>(sl + col * LD + ch * 8);
        }
      }
    }
    if (st + 1 < nsteps) gstore(buf ^ 1, next);
    lds_barrier();
  };
  for (int64_t st = 0; st < nsteps; st += 2) {
    step(st, greg0, greg1);
    if (st + 1 < nsteps) step(st + 1, greg1, greg0);
  }
}

template <int KS, int NF, bool VEC>
void launch_tgen(TgenParams p, hipStream_t s) {
  constexpr int NCB = 64 * NF;
  const int64_t gx = ceil_div(p.nc, NCB);
  // perturbations per workgroup: about 2048 workgroups in all, so each column block's O slice
  // is reused from L2 across the group
  p.ag = std::max<int64_t>(1, std::min<int64_t>(p.na, gx * p.na / 2048));
  const int64_t gy = ceil_div(p.na, p.ag);
  HH_REQUIRE(gx < (1LL << 31) && gy < 65536, "tgen: grid too large");
  hipLaunchKernelGGL((tgen_bf16x3_kernel<KS, NF, TG_RT, VEC>), dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s, p);
}

// column tiles per wave chosen so the H fragments, the prefetched G rows and the accumulators
// fit 256 VGPRs (two waves per SIMD, no spills); VF for nocc % 8 == 0, SF otherwise
template <int KS, int VF, int SF>
void launch_tgen_ks(const TgenParams& p, hipStream_t s) {
  if (p.nocc % 8 == 0) launch_tgen<KS, VF, true>(p, s);
  else launch_tgen<KS, SF, false>(p, s);
}

}  // namespace

bool tgen_supported(int64_t nocc) { return nocc >= 1 && 2 * nocc <= 256; }

void tgen_bf16x3(const float* bra, const float* Co, const float* O, const float* W, int64_t nao, int64_t nocc,
                 int64_t nprop, int64_t nc, int64_t a0, int64_t na, int64_t kpad, unsigned short* th,
                 unsigned short* tl, DevBuf& ws, hipStream_t s) {
  HH_REQUIRE(tgen_supported(nocc), "tgen: 2 nocc must be at most 256");
  // every row up to kpad is generated (rows >= nao come out zero), so T needs no clearing
  const int64_t mrows = kpad;
  HH_REQUIRE(kpad >= nao && kpad % 32 == 0, "tgen: kpad must cover nao and be a multiple of 32");
  HH_REQUIRE(nc % nocc == 0, "tgen: nc must be naux * nocc");
  HH_REQUIRE(na >= 1 && a0 >= 0 && a0 + na <= nprop && nc >= 1, "tgen: perturbation range");
  const int ks = (int)ceil_div(2 * nocc, 32);
  const int kp = 32 * ks;
  const int64_t glen = na * mrows * kp;
  u16* gh = static_cast<u16*>(ws.reserve(2 * sizeof(u16) * glen));
  u16* gl = gh + glen;
  {
    const int64_t nb = std::min<int64_t>(ceil_div(glen, 256), 8192);
    hipLaunchKernelGGL(tgen_pack_g_kernel, dim3((unsigned)nb), dim3(256), 0, s, bra, Co, nao, nocc, a0, na, mrows, kp,
                       gh, gl);
    HH_HIP(hipGetLastError());
  }
  TgenParams p;
  p.gh = gh;
  p.gl = gl;
  p.O = O;
  p.W = W;
  p.nocc = nocc;
  p.naux = nc / nocc;
  p.ldw = nocc * nprop;
  p.nc = nc;
  p.a0 = a0;
  p.na = na;
  p.mrows = mrows;
  p.kbp = kpad / 32;
  p.nn = nocc * na;
  p.th = th;
  p.tl = tl;
  switch (ks) {
    case 1: launch_tgen_ks<1, 4, 4>(p, s); break;
    case 2: launch_tgen_ks<2, 4, 2>(p, s); break;
    case 3: launch_tgen_ks<3, 3, 2>(p, s); break;
    case 4: launch_tgen_ks<4, 2, 2>(p, s); break;
    case 5: launch_tgen_ks<5, 2, 1>(p, s); break;
    case 6: launch_tgen_ks<6, 1, 1>(p, s); break;
    case 7: launch_tgen_ks<7, 1, 1>(p, s); break;
    default: launch_tgen_ks<8, 1, 1>(p, s); break;
  }
  HH_HIP(hipGetLastError());
}

}  // namespace hh
