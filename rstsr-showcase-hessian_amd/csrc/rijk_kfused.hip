// Fused BF16x3 RI-K contraction (see rijk_kfused.hpp).
//
// Work decomposition of term 1, out[mu, (i, a)] = sum_{b, P} sum_{nu in block b} B_P[mu, nu] Q[nu, P, (i, a)]:
//   - a workgroup = 4 waves (one per SIMD, up to 512 registers each) owns all 32 nb rows mu and a
//     tile of 128 columns c = i + nocc a (32 per wave), over a K slice: one nu block b and a
//     range of P; the slices' FP32 partial sums are reduced in FP64 in a fixed order afterwards;
//   - per K step (b, P): the B_P tile [32 nb rows][32 nu] (bf16 hi and lo, 64 bytes per row each)
//     is copied into LDS by direct global-to-LDS loads issued one step ahead (double-buffered);
//     each wave computes its 32 x 32 block of Q = bra_a O_P on the matrix cores
//     (v_mfma_f32_32x32x16_bf16, three products hi*hi + hi*lo + lo*hi): the A operand is the
//     bra block of the perturbations its columns touch (registers, loaded once per slice), the B
//     operand O_P[j, i_c] masked to the columns of that perturbation; the FP32 result is split
//     into bf16 hi / lo in registers and is the B operand of 2 x nb x 3 contraction MFMAs.
//   - Register layouts (32 x 32 x 16): an A operand lane l holds row l % 32, k = 8 (l / 32) + e;
//     a B operand lane holds column l % 32, the same k; accumulator register v of lane l holds
//     row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32.  Q's accumulator registers 8 kappa .. 8 kappa + 7
//     are therefore the B operand of K half kappa once the 32 nu of a block are stored in the
//     order pos = 8 c + e <-> nu = 16 (c / 2) + 8 (e / 4) + 4 (c % 2) + e % 4 (c = 2 kappa + l / 32):
//     kfused_pack_b writes B_P in that order, so no data moves between lanes.
//   - LDS rows are 64 bytes; chunk c of row r sits at slot c ^ ((r >> 2) & 3), so the 16 rows
//     of each quarter-wave's 16-byte reads fall on distinct bank groups.  Global-to-LDS loads
//     write 64 consecutive 16-byte slots per wave instruction, so each lane fetches the global
//     chunk that belongs at its slot.
//   - XCD-aware order: consecutive workgroups of one XCD take the same K slice (all column
//     tiles), so each B_P tile is fetched from HBM about once per XCD and reused from its L2.
#include "rijk_kfused.hpp"

#include "gemm_nt.hpp"

namespace hh {

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef unsigned short u16;
typedef __attribute__((address_space(3))) void* lds_ptr;

__device__ __forceinline__ u16 bf_bits(__bf16 h) { return __builtin_bit_cast(u16, h); }
__device__ __forceinline__ void split1(float x, u16& hi, u16& lo) {
  const __bf16 h = (__bf16)x;
  hi = bf_bits(h);
  lo = bf_bits((__bf16)(x - (float)h));
}

// position pos (0..31) of a 32-nu block -> nu within the block
__host__ __device__ __forceinline__ int nu_of_pos(int pos) {
  const int c = pos >> 3, e = pos & 7;
  return 16 * (c >> 1) + 8 * (e >> 2) + 4 * (c & 1) + (e & 3);
}

__global__ void pack_b_kernel(const double* __restrict__ cderi, int64_t nao, int64_t naux, int64_t nb, int64_t mp,
                              int64_t naux_e, u16* __restrict__ hi, u16* __restrict__ lo) {
  const int64_t ntp = nao * (nao + 1) / 2;
  const int64_t total = (nb * naux_e + 1) * mp * 32;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int pos = (int)(e & 31);
    const int64_t mu = (e >> 5) % mp, P = (e >> 5) / mp % naux_e, b = (e >> 5) / mp / naux_e;
    const int64_t nu = 32 * b + nu_of_pos(pos);
    float x = 0.f;
    if (mu < nao && nu < nao && P < naux && b < nb) {
      const int64_t l = mu < nu ? mu : nu, g = mu < nu ? nu : mu;
      x = (float)cderi[g * (g + 1) / 2 + l + P * ntp];
    }
    split1(x, hi[e], lo[e]);
  }
}

__global__ void pack_o_kernel(const double* __restrict__ O, int64_t nocc, int64_t naux, int64_t naux_e, int64_t jp,
                              u16* __restrict__ hi, u16* __restrict__ lo) {
  const int64_t total = (naux_e + 2) * nocc * jp;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e % jp, i = (e / jp) % nocc, P = e / jp / nocc;
    const float x = (j < nocc && P < naux) ? (float)O[j + nocc * (P + naux * i)] : 0.f;
    split1(x, hi[e], lo[e]);
  }
}

__global__ void pack_bra_kernel(const double* __restrict__ bra, int64_t nao, int64_t nocc, int64_t nprop, int64_t nb,
                                int64_t jp, u16* __restrict__ hi, u16* __restrict__ lo) {
  const int64_t total = nprop * nb * 32 * jp;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e % jp, nu = (e / jp) % (32 * nb), a = e / jp / (32 * nb);
    const float x = (j < nocc && nu < nao) ? (float)bra[nu + nao * (j + nocc * a)] : 0.f;
    split1(x, hi[e], lo[e]);
  }
}

// one 16-byte output chunk (8 consecutive kk) per thread, threads mu-fastest
__global__ void pack_x_kernel(const float* __restrict__ X, int64_t nao, int64_t naux, int64_t nocc, int64_t naux_pad,
                              u16* __restrict__ hi, u16* __restrict__ lo) {
  const int64_t nch = naux_pad * nocc / 8;
  const int64_t total = nch * nao;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t mu = e % nao, ch = e / nao;
    const int64_t kk0 = 8 * ch, j = kk0 / naux_pad, P0 = kk0 % naux_pad;
    u16 h[8], l[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int64_t P = P0 + q;
      const float x = P < naux ? X[mu + nao * (P + naux * j)] : 0.f;
      split1(x, h[q], l[q]);
    }
    const int64_t o = ((kk0 / 32) * nao + mu) * 32 + kk0 % 32;
    *reinterpret_cast<uint4*>(hi + o) = *reinterpret_cast<const uint4*>(h);
    *reinterpret_cast<uint4*>(lo + o) = *reinterpret_cast<const uint4*>(l);
  }
}

// zero the entries naux <= P < naux_pad of the term-2 operand W' (the W product writes P < naux):
// the last K block of each j, rows i + nocc a
__global__ void zero_w_pad_kernel(int64_t naux, int64_t nocc, int64_t R, int64_t naux_pad, u16* __restrict__ hi,
                                  u16* __restrict__ lo) {
  const int64_t p0 = naux % 32, npad = 32 - p0;
  const int64_t total = nocc * R * npad;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = e % npad, r = (e / npad) % R, j = e / (npad * R);
    const int64_t kb = (naux_pad * j + naux - p0) / 32;
    const int64_t o = (kb * R + r) * 32 + p0 + q;
    hi[o] = 0;
    lo[o] = 0;
  }
}

struct KfParams {
  const u16* bh;  // [nb][naux_e][mp][32] (+ one tile)
  const u16* bl;
  const u16* oh;  // [naux_e + 2][nocc][jp]
  const u16* ol;
  const u16* rh;  // [nprop][nb][32][jp]
  const u16* rl;
  float* partial;  // [nslice][ncol][mp]
  int64_t naux_e, nocc, ncol, mp, jp, pchunk, ntile, nwg;
  int nb, nh, psplit;
};

// XCD x receives workgroups x, x + 8, ...; they are given consecutive linear ids (the first
// nwg % 8 XCDs one more), so one XCD sweeps consecutive (slice, tile) pairs.
__device__ __forceinline__ int64_t xcd_remap(int64_t w, int64_t nwg) {
  const int64_t q = nwg / 8, r = nwg % 8, x = w % 8;
  return x * q + (x < r ? x : r) + w / 8;
}

__device__ __forceinline__ uint4 ld16(const u16* p) { return *reinterpret_cast<const uint4*>(p); }
__device__ __forceinline__ v8bf as_bf(uint4 v) { return __builtin_bit_cast(v8bf, v); }
__device__ __forceinline__ uint4 masked(uint4 v, bool on) {
  return on ? v : uint4{0u, 0u, 0u, 0u};
}

// NAW: perturbations one workgroup's 128 columns can touch (nocc >= 8, 17, 33, 49, 65 for NH 1..5)
template <int NH>
constexpr int naw_of() { return NH == 1 ? 17 : NH == 2 ? 9 : NH == 3 ? 5 : NH == 4 ? 4 : 3; }

template <int NMB, int NH, int NA>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void kfused_kernel(KfParams p) {
  constexpr int TILE = NMB * 32 * 32;  // u16 per array per stage
  constexpr int NAW = naw_of<NH>();
  constexpr int NFRAG = NAW * NH * 2;  // bra fragments of the workgroup: (perturbation, j block, hi / lo)
  // the two stages are separate LDS objects and every access names its stage statically, so the
  // compiler sees that reads of one stage do not depend on the copies in flight into the other
  __shared__ __attribute__((aligned(16))) u16 lds0[2][TILE];
  __shared__ __attribute__((aligned(16))) u16 lds1[2][TILE];
  // bra blocks (rows nu of block b) of the workgroup's perturbations, stored as register images
  // (64 lanes x 16 bytes per fragment): conflict-free reads, one per MFMA operand
  __shared__ __attribute__((aligned(16))) uint4 sbra[NFRAG * 64];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, h = lane >> 5, l32 = lane & 31;
  const int64_t lin = xcd_remap(blockIdx.x, p.nwg);
  const int64_t tn = lin % p.ntile, sl = lin / p.ntile;
  const int b = (int)(sl / p.psplit);
  const int64_t P0 = (sl % p.psplit) * p.pchunk;
  const int64_t P1 = min(p.naux_e, P0 + p.pchunk);  // P1 - P0 is even
  const int64_t cb = tn * 128 + 32 * w;  // first column of this wave
  const int64_t col = cb + l32;
  const bool valid = col < p.ncol;
  const int a_l = valid ? (int)(col / p.nocc) : -1;
  const int i_l = valid ? (int)(col - (int64_t)a_l * p.nocc) : 0;
  const int64_t a_wg = tn * 128 / p.nocc;  // first perturbation of the workgroup
  const int a_first = (int)(cb / p.nocc);
  const int t_off = (int)(a_first - a_wg);  // this wave's first perturbation in sbra

  // stage bra: fragment f = (tw NH + qq) 2 + arr; zero past the chunk's last perturbation
  for (int e = tid; e < NFRAG * 64; e += 256) {
    const int f = e >> 6, ln = e & 63, arr = f & 1, qq = (f >> 1) % NH, tw = (f >> 1) / NH;
    const int64_t a = a_wg + tw;
    uint4 v = {0u, 0u, 0u, 0u};
    if (a * p.nocc < p.ncol)
      v = ld16((arr ? p.rl : p.rh) + ((a * p.nb + b) * 32 + (ln & 31)) * p.jp + 16 * qq + 8 * (ln >> 5));
    sbra[e] = v;
  }

  // O_P[j, i_l] for j = 16 q + 8 h .. + 7 (the B operand of Q's product)
  // (lanes past the last column read column 0 and are masked)
  const u16* o_hi = p.oh + (int64_t)i_l * p.jp + 8 * h;
  const u16* o_lo = p.ol + (int64_t)i_l * p.jp + 8 * h;
  const int64_t o_stride = p.nocc * p.jp;
  auto load_o1 = [&](int64_t P, uint4 (&oh)[NH], uint4 (&ol)[NH], int q) {
    oh[q] = ld16(o_hi + P * o_stride + 16 * q);
    ol[q] = ld16(o_lo + P * o_stride + 16 * q);
  };
  auto load_o = [&](int64_t P, uint4 (&oh)[NH], uint4 (&ol)[NH]) {
#pragma unroll
    for (int q = 0; q < NH; ++q) {
      oh[q] = ld16(o_hi + P * o_stride + 16 * q);
      ol[q] = ld16(o_lo + P * o_stride + 16 * q);
    }
  };
  // B_P tile (32 NMB rows) -> LDS stage: groups of 64 slots (1 KB) per wave instruction, NMB per
  // wave.  Slot s of row r holds global chunk s ^ ((r >> 2) & 3); within a group that depends on
  // the lane only.
  constexpr int GPA = 2 * NMB;  // groups per array (32 NMB rows x 4 slots / 64)
  const int lane_off = (lane >> 2) * 32 + (((lane & 3) ^ ((lane >> 4) & 3)) * 8);
  constexpr int64_t TILE_LEN = (int64_t)TILE;
  // one of the NMB copy instructions of a wave (group w + 4 k)
  auto dma_one = [&](int64_t P, u16 (&dst)[2][TILE], int k) {
    const int64_t base = ((int64_t)b * p.naux_e + P) * TILE_LEN + lane_off;
    const int g = w + 4 * k;
    const int arr = g >= GPA ? 1 : 0;
    const int gg = g - arr * GPA;
    const u16* src = (arr ? p.bl : p.bh) + base + gg * 512;
    __builtin_amdgcn_global_load_lds(src, (lds_ptr)&dst[arr][gg * 512], 16, 0, 0);
  };
  auto dma_b = [&](int64_t P, u16 (&dst)[2][TILE]) {
    const int64_t base = ((int64_t)b * p.naux_e + P) * TILE_LEN + lane_off;
#pragma unroll
    for (int k = 0; k < NMB; ++k) {
      const int g = w + 4 * k;  // 4 NMB groups over the 4 waves
      const int arr = g >= GPA ? 1 : 0;
      const int gg = g - arr * GPA;
      const u16* src = (arr ? p.bl : p.bh) + base + gg * 512;
      __builtin_amdgcn_global_load_lds(src, (lds_ptr)&dst[arr][gg * 512], 16, 0, 0);
    }
  };

  v16f acc[NMB];
#pragma unroll
  for (int m = 0; m < NMB; ++m)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[m][v] = 0.f;

  // Q[nu, c] = sum_j bra[nu, j, a_c] O_P[j, i_c] over the (at most NA) perturbations of this
  // wave's columns, one (perturbation t, j block qq) pair at a time (three MFMAs).  Branch-free:
  // the O operand is zero in every lane whose column belongs to another perturbation (and in all
  // lanes for t past the wave's last one).
  constexpr int NPAIR = NA * NH;
  auto qload = [&](int pr, v8bf& ah, v8bf& al) {
    const int t = pr / NH, qq = pr % NH;
    const int tw = min(t_off + t, NAW - 1);  // past the staged range: only masked columns use it
    const int f = (tw * NH + qq) * 2;
    ah = as_bf(sbra[f * 64 + lane]);
    al = as_bf(sbra[(f + 1) * 64 + lane]);
  };
  auto qmma = [&](int pr, const v8bf& ah, const v8bf& al, const uint4 (&oh)[NH], const uint4 (&ol)[NH], v16f& q) {
    const int t = pr / NH, qq = pr % NH;
    const bool on = valid && a_l == a_first + t && t_off + t < NAW;
    const v8bf bh = as_bf(masked(oh[qq], on)), bl = as_bf(masked(ol[qq], on));
    q = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, q, 0, 0, 0);
    q = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, q, 0, 0, 0);
    q = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, q, 0, 0, 0);
  };
  // Q (accumulator layout) -> bf16 hi / lo B operands of the two K halves
  auto qsplit = [&](const v16f& q, v8bf (&qh)[2], v8bf (&ql)[2]) {
#pragma unroll
    for (int kap = 0; kap < 2; ++kap)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float x = q[8 * kap + e];
        const __bf16 hv = (__bf16)x;
        qh[kap][e] = hv;
        ql[kap][e] = (__bf16)(x - (float)hv);
      }
  };
  auto sync_stage = [&]() {
    // the next stage's global-to-LDS copies (and O loads) have landed; every wave is done with
    // this stage before the copy after the next barrier overwrites it
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  };

  // Software pipeline over P: step P contracts stage P with Q(P) (built during step P - 1) and
  // builds Q(P + 1) by pairs interleaved with the second K half of the contraction; B(P + 1) goes
  // to the other LDS stage, one copy instruction per row block of the first K half.  Steps alternate between the two stages.  In program order
  // (scheduling barriers keep it): per row block the fragments of block m + 2 are requested while
  // block m's three products run.
  // O(P + 1) is in registers when step P starts (loaded one step earlier, before the previous
  // barrier's wait): two register sets, swapped between the two steps of a loop trip.
  uint4 oah[NH], oal[NH], obh[NH], obl[NH];
  v8bf qh[2], ql[2], rh[2], rl[2];
  load_o(P0, oah, oal);
  load_o(P0 + 1, obh, obl);
  dma_b(P0, lds0);
  sync_stage();
  {
    v16f q;
#pragma unroll
    for (int v = 0; v < 16; ++v) q[v] = 0.f;
#pragma unroll
    for (int pr = 0; pr < NPAIR; ++pr) {
      v8bf ah, al;
      qload(pr, ah, al);
      qmma(pr, ah, al, oah, oal, q);
    }
    qsplit(q, qh, ql);
  }
  // step P: O(P + 1) in (uh, ul), O(P + 2) requested into (lh, ll); index P + 2 <= P1 + 1 and tile
  // P + 1 <= P1 exist (padding past every range)
  // Q(P + 1) is split into its bf16 operand registers (the second set of two, alternating between
  // steps) right after the row block QS, so the conversion overlaps the remaining products
  constexpr int QS = NPAIR + 1 < NMB ? NPAIR + 1 : -1;
  auto step = [&](int64_t P, const u16 (&cur)[2][TILE], u16 (&nxt)[2][TILE], const uint4 (&uh)[NH],
                  const uint4 (&ul)[NH], uint4 (&lh)[NH], uint4 (&ll)[NH], const v8bf (&ch)[2],
                  const v8bf (&cl)[2], v8bf (&nh)[2], v8bf (&nl)[2]) {
    const u16* sh = cur[0];
    const u16* sl2 = cur[1];
    v16f q;
#pragma unroll
    for (int v = 0; v < 16; ++v) q[v] = 0.f;
    v8bf bh0, bl0;  // bra fragments of the next Q pair
    qload(0, bh0, bl0);
#pragma unroll
    for (int kap = 0; kap < 2; ++kap) {
      auto frag = [&](int m, v8bf& ah, v8bf& al) {
        const int row = 32 * m + l32;
        const int off = row * 32 + (((2 * kap + h) ^ ((row >> 2) & 3)) * 8);
        al = *reinterpret_cast<const v8bf*>(sl2 + off);
        ah = *reinterpret_cast<const v8bf*>(sh + off);
      };
      v8bf ah0, al0, ah1, al1;
      frag(0, ah0, al0);
      if (NMB > 1) frag(1, ah1, al1);
#pragma unroll
      for (int m = 0; m < NMB; ++m) {
        // rows past 32 nb read stale LDS and are never stored
        acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al0, ch[kap], acc[m], 0, 0, 0);
        acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah0, cl[kap], acc[m], 0, 0, 0);
        acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah0, ch[kap], acc[m], 0, 0, 0);
        if (kap == 1 && m == QS) qsplit(q, nh, nl);
        if (kap == 1 && m < NPAIR) {
          qmma(m, bh0, bl0, uh, ul, q);
          if (m + 1 < NPAIR) qload(m + 1, bh0, bl0);
        }
        // the copies of B(P + 1) are spread over the first K half, one per row block: issued
        // back to back at the start of the step they delayed the first products
        if (kap == 0) dma_one(P + 1, nxt, m);
        if (kap == 0 && m < NH) load_o1(P + 2, lh, ll, m);  // O(P + 2), used two steps on
        ah0 = ah1;
        al0 = al1;
        if (m + 2 < NMB) frag(m + 2, ah1, al1);
        // keep the schedule as written: hoisting every fragment read would exceed the registers
        __builtin_amdgcn_sched_barrier(0);
      }
    }
#pragma unroll
    for (int q = NMB; q < NH; ++q) load_o1(P + 2, lh, ll, q);
#pragma unroll
    for (int pr = NMB; pr < NPAIR; ++pr) {
      qmma(pr, bh0, bl0, uh, ul, q);
      if (pr + 1 < NPAIR) qload(pr + 1, bh0, bl0);
    }
    if (QS < 0) qsplit(q, nh, nl);  // Q(P + 1); unused after the last step
    sync_stage();
  };
  for (int64_t P = P0; P < P1; P += 2) {
    step(P, lds0, lds1, obh, obl, oah, oal, qh, ql, rh, rl);
    step(P + 1, lds1, lds0, oah, oal, obh, obl, rh, rl, qh, ql);
  }

  if (valid) {
    float* dst = p.partial + (sl * p.ncol + col) * p.mp;
#pragma unroll
    for (int m = 0; m < NMB; ++m) {
      if (m < p.nb) {
#pragma unroll
        for (int g = 0; g < 4; ++g)
          *reinterpret_cast<float4*>(dst + 32 * m + 8 * g + 4 * h) =
              make_float4(acc[m][4 * g], acc[m][4 * g + 1], acc[m][4 * g + 2], acc[m][4 * g + 3]);
      }
    }
  }
}

__global__ void kfused_reduce_kernel(const float* __restrict__ partial, int nslice, int64_t ncol, int64_t mp,
                                     int64_t nao, double alpha, double beta, double* __restrict__ out) {
  const int64_t total = nao * ncol;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t mu = e % nao, c = e / nao;
    double s = 0.0;
    for (int z = 0; z < nslice; ++z) s += (double)partial[(z * ncol + c) * mp + mu];
    double* o = out + e;
    *o = beta != 0.0 ? alpha * s + beta * (*o) : alpha * s;
  }
}

unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 65536)); }

int na_span(int64_t nocc) { return 2 + (int)(30 / nocc); }

template <int NMB, int NH, int NA>
void launch(const KfParams& p, hipStream_t s) {
  static_assert(NMB * 8 * 1024 + naw_of<NH>() * NH * 2 * 1024 <= 160 * 1024, "LDS");
  hipLaunchKernelGGL((kfused_kernel<NMB, NH, NA>), dim3((unsigned)p.nwg), dim3(256), 0, s, p);
}

template <int NMB>
void launch_nh(const KfParams& p, hipStream_t s) {
  switch (p.nh) {
    case 1: launch<NMB, 1, 5>(p, s); break;
    case 2: launch<NMB, 2, 3>(p, s); break;
    case 3: launch<NMB, 3, 2>(p, s); break;
    case 4: launch<NMB, 4, 2>(p, s); break;
    default: launch<NMB, 5, 2>(p, s); break;
  }
}

}  // namespace

bool kfused_supported(int64_t nao, int64_t nocc) {
  if (nao < 1 || nocc < 8 || nocc > 80) return false;
  if (kfused_nb(nao) > 13) return false;
  return na_span(nocc) <= (nocc <= 16 ? 5 : nocc <= 32 ? 3 : 2);
}

int kfused_nmb(int64_t nb) { return nb <= 4 ? 4 : nb <= 8 ? 8 : nb <= 10 ? 10 : nb <= 12 ? 12 : 13; }

int64_t kfused_b_len(int64_t nao, int64_t naux) {
  const int64_t nb = kfused_nb(nao);
  return (nb * kfused_naux_e(naux) + 1) * 32 * kfused_nmb(nb) * 32;
}

int64_t kfused_o_len(int64_t nocc, int64_t naux) { return (kfused_naux_e(naux) + 2) * nocc * kfused_jp(nocc); }

void kfused_pack_b(const double* cderi, int64_t nao, int64_t naux, unsigned short* hi, unsigned short* lo,
                   hipStream_t s) {
  const int64_t nb = kfused_nb(nao), mp = 32 * kfused_nmb(nb);
  hipLaunchKernelGGL(pack_b_kernel, dim3(blocks_for(kfused_b_len(nao, naux))), dim3(256), 0, s, cderi, nao, naux, nb, mp,
                     kfused_naux_e(naux), hi, lo);
  HH_HIP(hipGetLastError());
}

void kfused_pack_o(const double* O, int64_t nocc, int64_t naux, unsigned short* hi, unsigned short* lo,
                   hipStream_t s) {
  const int64_t jp = kfused_jp(nocc);
  hipLaunchKernelGGL(pack_o_kernel, dim3(blocks_for(kfused_o_len(nocc, naux))), dim3(256), 0, s, O, nocc, naux,
                     kfused_naux_e(naux), jp, hi, lo);
  HH_HIP(hipGetLastError());
}

void kfused_pack_bra(const double* bra, int64_t nao, int64_t nocc, int64_t nprop, unsigned short* hi,
                     unsigned short* lo, hipStream_t s) {
  const int64_t nb = kfused_nb(nao), jp = kfused_jp(nocc);
  hipLaunchKernelGGL(pack_bra_kernel, dim3(blocks_for(nprop * nb * 32 * jp)), dim3(256), 0, s, bra, nao, nocc, nprop,
                     nb, jp, hi, lo);
  HH_HIP(hipGetLastError());
}

void kfused_pack_x(const float* X, int64_t nao, int64_t naux, int64_t nocc, unsigned short* hi, unsigned short* lo,
                   hipStream_t s) {
  const int64_t np = kfused_naux_pad(naux);
  hipLaunchKernelGGL(pack_x_kernel, dim3(blocks_for(np * nocc / 8 * nao)), dim3(256), 0, s, X, nao, naux, nocc, np, hi,
                     lo);
  HH_HIP(hipGetLastError());
}

void kfused_zero_w_pad(int64_t naux, int64_t nocc, int64_t nprop, unsigned short* hi, unsigned short* lo,
                       hipStream_t s) {
  if (naux % 32 == 0) return;
  const int64_t R = nocc * nprop, total = nocc * R * (32 - naux % 32);
  const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(total, 256), 65536));
  hipLaunchKernelGGL(zero_w_pad_kernel, dim3(blocks), dim3(256), 0, s, naux, nocc, R, kfused_naux_pad(naux), hi, lo);
  HH_HIP(hipGetLastError());
}

void kfused_term1(const KfusedArgs& a, DevBuf& ws, hipStream_t s) {
  HH_REQUIRE(kfused_supported(a.nao, a.nocc), "kfused: unsupported nao / nocc");
  HH_REQUIRE(a.bh && a.bl && a.oh && a.ol && a.rh && a.rl && a.out && a.nprop > 0 && a.naux > 0, "kfused: arguments");
  KfParams p{};
  p.bh = a.bh; p.bl = a.bl; p.oh = a.oh; p.ol = a.ol; p.rh = a.rh; p.rl = a.rl;
  p.naux_e = kfused_naux_e(a.naux);
  p.nocc = a.nocc;
  p.ncol = a.nocc * a.nprop;
  p.nb = (int)kfused_nb(a.nao);
  p.mp = 32 * (int64_t)p.nb;
  p.jp = kfused_jp(a.nocc);
  p.nh = (int)(p.jp / 16);
  p.ntile = ceil_div(p.ncol, 128);
  // P splits: the count in [1, 16] (at least 8 P per slice) whose workgroup count fills the
  // largest fraction of whole rounds of one workgroup per CU; slices hold an even number of P
  // (the padded range naux_e), as the kernel takes two steps per loop trip
  const int64_t cus = cu_count();
  const int64_t base = p.ntile * p.nb;
  double best = -1.0;
  int ps = 1;
  for (int k = 1; k <= 16; ++k) {
    if (k > 1 && p.naux_e / k < 8) break;
    const int64_t nwg = base * k;
    const double eff = (double)nwg / (double)(ceil_div(nwg, cus) * cus);
    if (eff > best + 1e-9) {
      best = eff;
      ps = k;
    }
  }
  p.pchunk = 2 * ceil_div(p.naux_e, 2 * (int64_t)ps);
  ps = (int)ceil_div(p.naux_e, p.pchunk);  // rounding the chunk to even can leave fewer slices
  p.psplit = ps;
  p.nwg = base * ps;
  HH_REQUIRE(p.nwg < (1LL << 31), "kfused: grid too large");
  const int64_t nslice = (int64_t)p.nb * ps;
  p.partial = static_cast<float*>(ws.reserve(sizeof(float) * (size_t)nslice * p.ncol * p.mp));
  if (a.ev_begin) HH_HIP(hipEventRecord(a.ev_begin, s));
  // the contraction covers NMB row blocks (rows past 32 nb are computed and dropped; the packed
  // B tiles hold 32 NMB rows)
  switch (kfused_nmb(p.nb)) {
    case 4: launch_nh<4>(p, s); break;
    case 8: launch_nh<8>(p, s); break;
    case 10: launch_nh<10>(p, s); break;
    case 12: launch_nh<12>(p, s); break;
    default: launch_nh<13>(p, s); break;
  }
  HH_HIP(hipGetLastError());
  if (a.ev_main_end) HH_HIP(hipEventRecord(a.ev_main_end, s));
  hipLaunchKernelGGL(kfused_reduce_kernel, dim3(blocks_for(a.nao * p.ncol)), dim3(256), 0, s, p.partial, (int)nslice,
                     p.ncol, p.mp, a.nao, a.alpha, a.beta, a.out);
  HH_HIP(hipGetLastError());
}

}  // namespace hh
