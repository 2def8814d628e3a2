// DFT skeleton-Hessian setup passes over a fixed grid (reference: make_hessian_setup,
// src/numint_matmul/hess_rks.rs:71-565, and make_hessian_setup_uks, hess_uks.rs:56-384).
//
// Per grid batch (the reference's make_hessian_setup_batched, hess_rks.rs:611-750, sums the
// additive intermediates over grid chunks the same way):
//   1. ao_dm0_s = ao[:, :, c] dm0_s (c < 4) on the matrix cores;
//   2. one pointwise pass per (grid point, atom): drho_s (d rho / d R_A) in registers, then the
//      kernel-weighted W_s2 = w sum_{x,s1} fxc[x, s1, y, s2] drho_s1 for both the de_fxc product
//      and the vmat_fxc operand (the reference forms the same contraction twice);
//   3. one pass per AO over the batch's points for the AO-diagonal de_vxc terms;
//   4. grid-contracted products (K = grid points) for de_fxc, the de_vxc_off AO-pair tensor,
//      vmat_ip and the fxc part of vmat_deriv1, accumulated across batches by GEMM beta = 1.
// The atom-blocked reductions with dm0 and the bra + ket assembly run once at the end.
//
// AO components: 0 value, 1-3 x y z, 4-9 xx xy xz yy yz zz, 10-19 xxx xxy xxz xyy xyz xzz yyy yyz
// yzz zzz (index tuples sorted, lexicographic order).
#include <algorithm>
#include <vector>

#include "becke.hpp"
#include "elementwise.hpp"
#include "gemm_nt.hpp"
#include "numint.hpp"

namespace hh {
namespace {

unsigned nblocks(int64_t n, int tpb = 256) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, tpb), 1 << 20));
}

template <int DEN> struct Den {
  static constexpr int NV = DEN == DEN_RHO ? 1 : (DEN == DEN_SIGMA ? 4 : 5);
};

// component of the second derivative along (i, j)
__host__ __device__ constexpr int c2(int i, int j) {
  const int a = i < j ? i : j, b = i < j ? j : i;
  return a == 0 ? 4 + b : (a == 1 ? 6 + b : 9);
}
// component of the third derivative along (i, j, k)
__host__ __device__ constexpr int c3(int i, int j, int k) {
  const int lo = i < j ? (i < k ? i : k) : (j < k ? j : k);
  const int hi = i > j ? (i > k ? i : k) : (j > k ? j : k);
  const int mid = i + j + k - lo - hi;
  return lo == 0 ? 6 + c2(mid, hi) : (lo == 1 ? 9 + c2(mid, hi) : 19);
}
static_assert(c2(0, 0) == 4 && c2(1, 0) == 5 && c2(2, 0) == 6 && c2(1, 1) == 7 && c2(2, 1) == 8 && c2(2, 2) == 9);
static_assert(c3(0, 0, 0) == 10 && c3(1, 0, 0) == 11 && c3(0, 2, 0) == 12 && c3(1, 1, 0) == 13 &&
              c3(2, 1, 0) == 14 && c3(2, 0, 2) == 15 && c3(1, 1, 1) == 16 && c3(2, 1, 1) == 17 &&
              c3(1, 2, 2) == 18 && c3(2, 2, 2) == 19);

// xx xy xz yy yz zz
__device__ constexpr int P6I[6] = {0, 0, 0, 1, 1, 2};
__device__ constexpr int P6J[6] = {0, 1, 2, 1, 2, 2};

// ---- density (hess_rks.rs:187-200): rho[g, v] at ld ng, of the batch's points ----
template <int DEN>
__global__ void setup_rho_kernel(const double* __restrict__ ao, int64_t ng, int64_t nao, int64_t g0, int64_t gb,
                                 const double* __restrict__ aod, double* __restrict__ rho) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= gb) return;
  const int64_t sa = ng * nao, sd = gb * nao;
  const double* a = ao + g0 + g;
  const double* d = aod + g;
  double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0, tau = 0.0;
  for (int64_t mu = 0; mu < nao; ++mu) {
    const double d0 = d[gb * mu];
    r0 += a[ng * mu] * d0;
    if constexpr (DEN != DEN_RHO) {
      const double ax = a[ng * mu + sa], ay = a[ng * mu + 2 * sa], az = a[ng * mu + 3 * sa];
      r1 += ax * d0;
      r2 += ay * d0;
      r3 += az * d0;
      if constexpr (DEN == DEN_TAU)
        tau += ax * d[gb * mu + sd] + ay * d[gb * mu + 2 * sd] + az * d[gb * mu + 3 * sd];
    }
  }
  rho[g0 + g] = r0;
  if constexpr (DEN != DEN_RHO) {
    rho[g0 + g + ng] = 2.0 * r1;
    rho[g0 + g + 2 * ng] = 2.0 * r2;
    rho[g0 + g + 3 * ng] = 2.0 * r3;
  }
  if constexpr (DEN == DEN_TAU) rho[g0 + g + 4 * ng] = 0.5 * tau;
}

// ---- drho (hess_rks.rs:220-279) and the fxc-weighted W, one thread per (g, A) ----
// drho / W element (g, v, s, t, A) at g + gb (v + NV (s + NS (t + 3 A))).
// W[g, y, s2, t, A] = w_g sum_{x, s1} fxc[g, x, s1, y, s2] drho[g, x, s1, t, A].
template <int DEN, int NS>
__global__ __launch_bounds__(256) void drho_w_kernel(const double* __restrict__ ao, int64_t ng, int64_t nao, int64_t g0, int64_t gb,
                              const double* __restrict__ aod0, const double* __restrict__ aod1,
                              const int64_t* __restrict__ slices, int64_t natm, const double* __restrict__ fxc,
                              const double* __restrict__ w, double* __restrict__ drho, double* __restrict__ W) {
  constexpr int NV = Den<DEN>::NV;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= gb * natm) return;
  const int64_t g = e % gb, A = e / gb;
  const int64_t p0 = slices[2 * A], p1 = slices[2 * A + 1];
  const int64_t sa = ng * nao, sd = gb * nao;
  double dr[NS][NV][3];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int t = 0; t < 3; ++t) dr[s][v][t] = 0.0;
  for (int64_t mu = p0; mu < p1; ++mu) {
    const double* a = ao + (g0 + g) + ng * mu;
    double a1[3], a2[6];
#pragma unroll
    for (int t = 0; t < 3; ++t) a1[t] = a[(1 + t) * sa];
    if constexpr (DEN != DEN_RHO) {
#pragma unroll
      for (int c = 0; c < 6; ++c) a2[c] = a[(4 + c) * sa];
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double* d = (s == 0 ? aod0 : aod1) + g + gb * mu;
      const double d0 = d[0];
#pragma unroll
      for (int t = 0; t < 3; ++t) dr[s][0][t] -= a1[t] * d0;
      if constexpr (DEN != DEN_RHO) {
        const double d1[3] = {d[sd], d[2 * sd], d[3 * sd]};
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const double atr = a2[c2(t, r) - 4];
            dr[s][1 + r][t] -= atr * d0 + a1[t] * d1[r];
            if constexpr (DEN == DEN_TAU) dr[s][4][t] -= atr * d1[r];
          }
      }
    }
  }
  constexpr int NF = DEN == DEN_RHO ? 1 : 4;  // components doubled (tau carries its 0.5 already)
  const int64_t gg = g0 + g;
  const double wg = w[gg];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        if (v < NF) dr[s][v][t] *= 2.0;
        drho[g + gb * (v + NV * (s + NS * (t + 3 * A)))] = dr[s][v][t];
      }
#pragma unroll
  for (int s2 = 0; s2 < NS; ++s2)
#pragma unroll
    for (int y = 0; y < NV; ++y) {
      double f[NS][NV];
#pragma unroll
      for (int s1 = 0; s1 < NS; ++s1)
#pragma unroll
        for (int x = 0; x < NV; ++x) f[s1][x] = fxc[gg + ng * (x + NV * (s1 + NS * (y + NV * s2)))];
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        double acc = 0.0;
#pragma unroll
        for (int s1 = 0; s1 < NS; ++s1)
#pragma unroll
          for (int x = 0; x < NV; ++x) acc += f[s1][x] * dr[s1][x][t];
        W[g + gb * (y + NV * (s2 + NS * (t + 3 * A)))] = wg * acc;
      }
    }
}

// ---- AO-diagonal de_vxc terms (hess_rks.rs:302-344), one workgroup per (AO, grid split) ----
// acc[(split * nao + mu) * 6 + k] += this batch's sum over the split's points.
template <int DEN>
__global__ void __launch_bounds__(256) vxc_diag_kernel(const double* __restrict__ ao, int64_t ng, int64_t nao,
                                                       int64_t g0, int64_t gb, const double* __restrict__ aod,
                                                       const double* __restrict__ vxc, const double* __restrict__ w,
                                                       double* __restrict__ acc, int64_t nsplit) {
  constexpr int NV = Den<DEN>::NV;
  const int64_t mu = blockIdx.x, sp = blockIdx.y;
  const int64_t len = (gb + nsplit - 1) / nsplit, gs = sp * len, ge = std::min(gb, gs + len);
  const int64_t sa = ng * nao, sd = gb * nao;
  double part[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t g = gs + threadIdx.x; g < ge; g += blockDim.x) {
    const int64_t gg = g0 + g;
    const double wg = w[gg];
    double wv[NV];
#pragma unroll
    for (int x = 0; x < NV; ++x) wv[x] = wg * vxc[gg + ng * x];
    const double* a = ao + gg + ng * mu;
    const double* d = aod + g + gb * mu;
    double dd[4];
    dd[0] = d[0];
    double aow = dd[0] * wv[0];
    if constexpr (DEN != DEN_RHO) {
#pragma unroll
      for (int r = 1; r < 4; ++r) {
        dd[r] = d[r * sd];
        aow += dd[r] * wv[r];
      }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int t = P6I[k], s = P6J[k];
      double v = 2.0 * a[c2(t, s) * sa] * aow;
      if constexpr (DEN != DEN_RHO) {
        double a3 = 0.0;
#pragma unroll
        for (int r = 0; r < 3; ++r) a3 += a[c3(t, s, r) * sa] * wv[1 + r];
        v += 2.0 * a3 * dd[0];
      }
      if constexpr (DEN == DEN_TAU) {
#pragma unroll
        for (int r = 0; r < 3; ++r) v += a[c3(r, t, s) * sa] * dd[1 + r] * wv[4];
      }
      part[k] += v;
    }
  }
  __shared__ double red[6][256];
#pragma unroll
  for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = part[k];
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 6) acc[(sp * nao + mu) * 6 + threadIdx.x] += red[threadIdx.x][0];
}

// ---- operands of the vxc products, one thread per (g, mu) ----
// aowv[g, mu, t] (g + ldg (mu + nao t)): RHO w v0 ao_t; else 0.5 w v0 ao_t + sum_r w v_r ao_tr
// aow[g, mu] (g + ldg mu): RHO w v0 ao_0; else 0.5 w v0 ao_0 + sum_r w v_r ao_r
// (rows gb <= g < ldg are zero: the K padding of the matrix-core products)
// TAU: H2[(g, r), (mu, t)] = ao_tr, H2w = 0.5 w v_tau H2, AT[(g, r), mu] = 0.5 w v_tau ao_r
// (row (g, r) = g + gb r, leading dimension 3 gb)
template <int DEN>
__global__ void vxc_ops_kernel(const double* __restrict__ ao, int64_t ng, int64_t nao, int64_t g0, int64_t gb,
                               int64_t ldg, const double* __restrict__ vxc, const double* __restrict__ w,
                               double* __restrict__ aowv, double* __restrict__ aow, double* __restrict__ H2,
                               double* __restrict__ H2w, double* __restrict__ AT) {
  constexpr int NV = Den<DEN>::NV;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= ldg * nao) return;
  const int64_t g = e % ldg, mu = e / ldg, gg = g0 + g;
  if (g >= gb) {  // K padding rows of aowv / aow
    aow[g + ldg * mu] = 0.0;
#pragma unroll
    for (int t = 0; t < 3; ++t) aowv[g + ldg * (mu + nao * t)] = 0.0;
    return;
  }
  const int64_t sa = ng * nao;
  const double* a = ao + gg + ng * mu;
  const double wg = w[gg];
  double wv[NV];
#pragma unroll
  for (int x = 0; x < NV; ++x) wv[x] = wg * vxc[gg + ng * x];
  if constexpr (DEN == DEN_RHO) {
    aow[g + ldg * mu] = wv[0] * a[0];
#pragma unroll
    for (int t = 0; t < 3; ++t) aowv[g + ldg * (mu + nao * t)] = wv[0] * a[(1 + t) * sa];
  } else {
    double v = 0.5 * wv[0] * a[0];
#pragma unroll
    for (int r = 0; r < 3; ++r) v += wv[1 + r] * a[(1 + r) * sa];
    aow[g + ldg * mu] = v;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      double u = 0.5 * wv[0] * a[(1 + t) * sa];
#pragma unroll
      for (int r = 0; r < 3; ++r) u += wv[1 + r] * a[c2(t, r) * sa];
      aowv[g + ldg * (mu + nao * t)] = u;
    }
  }
  if constexpr (DEN == DEN_TAU) {
    const double ht = 0.5 * wv[4];
    const int64_t ld = 3 * gb;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      AT[(g + gb * r) + ld * mu] = ht * a[(1 + r) * sa];
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        const double h = a[c2(t, r) * sa];
        H2[(g + gb * r) + ld * (mu + nao * t)] = h;
        H2w[(g + gb * r) + ld * (mu + nao * t)] = ht * h;
      }
    }
  }
}

// ---- operand of the fxc part of vmat_deriv1 (hess_rks.rs:509-536 / hess_uks.rs:191-240) ----
// with W' = W of output spin s, component 0 halved and tau quartered:
// AW[blk][g, (mu, t, A)] = sum_{c < 4 (RHO: c = 0)} W'[g, c, t, A] ao_c[g, mu]    (blk 0)
//                        = W'[g, 4, t, A] ao_r[g, mu]                             (blk r = 1..3, TAU)
// element g + ldg (mu + nao (t + 3 A)) of block blk at offset blk * ldg * nao * 3 natm; rows
// gb <= g < ldg are zero (the K padding of the matrix-core product).
template <int DEN, int NS>
__global__ void fxc_ops_kernel(const double* __restrict__ ao, int64_t ng, int64_t nao, int64_t g0, int64_t gb,
                               int64_t ldg, const double* __restrict__ W, int s, int64_t natm,
                               double* __restrict__ AW) {
  // one thread per (g, group of FX_MU AOs): the W values of a point are read once per group
  constexpr int FX_MU = 4;
  constexpr int NV = Den<DEN>::NV;
  constexpr int NC = DEN == DEN_RHO ? 1 : 4;
  const int64_t nq = (nao + FX_MU - 1) / FX_MU;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= ldg * nq) return;
  const int64_t g = e % ldg, mu0 = (e / ldg) * FX_MU;
  const int nm = (int)(nao - mu0 < FX_MU ? nao - mu0 : FX_MU);
  const int64_t blk = ldg * nao * 3 * natm;
  if (g >= gb) {  // K padding rows
    for (int64_t tA = 0; tA < 3 * natm; ++tA)
      for (int i = 0; i < nm; ++i) {
        const int64_t o = g + ldg * (mu0 + i + nao * tA);
        AW[o] = 0.0;
        if constexpr (DEN == DEN_TAU) {
#pragma unroll
          for (int r = 1; r < 4; ++r) AW[o + r * blk] = 0.0;
        }
      }
    return;
  }
  const int64_t sa = ng * nao;
  double ac[FX_MU][4];
#pragma unroll
  for (int i = 0; i < FX_MU; ++i) {
    const double* a = ao + (g0 + g) + ng * (mu0 + (i < nm ? i : 0));
#pragma unroll
    for (int c = 0; c < NC; ++c) ac[i][c] = a[c * sa];
  }
  for (int64_t tA = 0; tA < 3 * natm; ++tA) {
    const double* Wp = W + g + gb * NV * (s + NS * tA);  // component c at Wp[gb * c]
    double wc[NC];
    wc[0] = 0.5 * Wp[0];
#pragma unroll
    for (int c = 1; c < NC; ++c) wc[c] = Wp[gb * c];
    const double w4 = DEN == DEN_TAU ? 0.25 * Wp[gb * (NV - 1)] : 0.0;
#pragma unroll
    for (int i = 0; i < FX_MU; ++i) {
      if (i >= nm) break;
      double v = wc[0] * ac[i][0];
#pragma unroll
      for (int c = 1; c < NC; ++c) v += wc[c] * ac[i][c];
      const int64_t o = g + ldg * (mu0 + i + nao * tA);
      AW[o] = v;
      if constexpr (DEN == DEN_TAU) {
#pragma unroll
        for (int r = 1; r < 4; ++r) AW[o + r * blk] = w4 * ac[i][r];
      }
    }
  }
}

// AO components 0..3 of the batch with the point axis padded to ldg (zero rows):
// out[(c nao + mu) ldg + g] = ao[g0 + g, mu, c]
__global__ void stage_ao_kernel(const double* __restrict__ ao, int64_t ng, int64_t nao, int64_t g0, int64_t gb,
                                int64_t ldg, double* __restrict__ out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= ldg * nao * 4) return;
  const int64_t g = e % ldg, cm = e / ldg;
  out[e] = g < gb ? ao[g0 + g + ng * cm] : 0.0;
}

// ---- finalisation ----
// de[t, s, A, B] = acc[t + 3 A, s + 3 B] (hess_rks.rs:296-299 transpose)
__global__ void de_fxc_final_kernel(const double* __restrict__ acc, int64_t natm, double* __restrict__ de) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t n3 = 3 * natm;
  if (e >= n3 * n3) return;
  const int64_t t = e % 3, s = (e / 3) % 3, A = (e / 9) % natm, B = e / (9 * natm);
  de[e] = acc[(t + 3 * A) + n3 * (s + 3 * B)];
}

// de[t, s, A, B] = [A == B] sum_{mu in A} sum_split acc[split][mu][pair(t, s)] (hess_rks.rs:340-351)
__global__ void de_diag_final_kernel(const double* __restrict__ acc, int64_t nsplit, int64_t nao,
                                     const int64_t* __restrict__ slices, int64_t natm, double* __restrict__ de) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= 9 * natm * natm) return;
  const int64_t t = e % 3, s = (e / 3) % 3, A = (e / 9) % natm, B = e / (9 * natm);
  double v = 0.0;
  if (A == B) {
    const int lo = t < s ? (int)t : (int)s, hi = t < s ? (int)s : (int)t;
    const int k = lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
    for (int64_t sp = 0; sp < nsplit; ++sp)
      for (int64_t mu = slices[2 * A]; mu < slices[2 * A + 1]; ++mu) v += acc[(sp * nao + mu) * 6 + k];
  }
  de[e] = v;
}

// de_off[t, s, A, B] = sum_{mu in A, nu in B} (C[mu + nao t, nu + nao s] + C[nu + nao s, mu + nao t]) dm0[mu, nu]
// with C [3 nao, 3 nao] the unsymmetrised AO-pair tensor (hess_rks.rs:401-418); one workgroup each.
__global__ void __launch_bounds__(256) de_off_final_kernel(const double* __restrict__ C, int64_t nao,
                                                           const double* __restrict__ dm0,
                                                           const int64_t* __restrict__ slices, int64_t natm,
                                                           double* __restrict__ de) {
  const int64_t e = blockIdx.x;
  const int64_t t = e % 3, s = (e / 3) % 3, A = (e / 9) % natm, B = e / (9 * natm);
  const int64_t a0 = slices[2 * A], na = slices[2 * A + 1] - a0;
  const int64_t b0 = slices[2 * B], nb = slices[2 * B + 1] - b0;
  const int64_t ld = 3 * nao;
  double v = 0.0;
  for (int64_t i = threadIdx.x; i < na * nb; i += blockDim.x) {
    const int64_t mu = a0 + i % na, nu = b0 + i / na;
    v += (C[(mu + nao * t) + ld * (nu + nao * s)] + C[(nu + nao * s) + ld * (mu + nao * t)]) * dm0[mu + nao * nu];
  }
  __shared__ double red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) de[e] = red[0];
}

// vmat_ip[mu, nu, t] = Cip[mu + nao t, nu]
__global__ void vmat_ip_final_kernel(const double* __restrict__ Cip, int64_t nao, double* __restrict__ out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= 3 * nao * nao) return;
  const int64_t mu = e % nao, nu = (e / nao) % nao, t = e / (nao * nao);
  out[e] = Cip[(mu + nao * t) + 3 * nao * nu];
}

// vmat_deriv1[mu, nu, t, A] = F[mu, nu, tA] + F[nu, mu, tA] - [mu in A] ip[mu, nu, t] - [nu in A] ip[nu, mu, t]
// with F[mu, nu, tA] = CfxT[nu, mu + nao tA] (hess_rks.rs:471-565, bra + ket assembly)
__global__ void vmat_deriv1_final_kernel(const double* __restrict__ Cfx, const double* __restrict__ Cip,
                                         const int64_t* __restrict__ atom_of, int64_t nao, int64_t natm,
                                         double* __restrict__ out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= nao * nao * 3 * natm) return;
  const int64_t mu = e % nao, nu = (e / nao) % nao, tA = e / (nao * nao);
  const int64_t t = tA % 3, A = tA / 3;
  double v = Cfx[nu + nao * (mu + nao * tA)] + Cfx[mu + nao * (nu + nao * tA)];
  if (atom_of[mu] == A) v -= Cip[(mu + nao * t) + 3 * nao * nu];
  if (atom_of[nu] == A) v -= Cip[(nu + nao * t) + 3 * nao * mu];
  out[e] = v;
}

// ---- Becke grid shift (hess_rks_becke.rs:916-1266, hess_uks_becke.rs:501-930) ----
// A point of atom G moves with G: its total derivative is d/dR_A + [A == G] d/dr, and
// d/dr = -sum_A d/dR_A on the fixed-grid density.  On the batch's drho and W (all points of
// atom G) that is the row of G replaced by minus the sum of the other rows; one thread per
// (g, v, s, t).  Element (g, j, t, A) at g + gb (j + nvns (t + 3 A)), j = v + NV s.
__global__ void shift_fix_kernel(double* __restrict__ drho, double* __restrict__ W, int64_t gb, int64_t nvns,
                                 int64_t natm, int64_t G) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= gb * nvns * 3) return;
  const int64_t g = e % gb, j = (e / gb) % nvns, t = e / (gb * nvns);
  double sd = 0.0, sw = 0.0;
  for (int64_t A = 0; A < natm; ++A) {
    if (A == G) continue;
    const int64_t i = g + gb * (j + nvns * (t + 3 * A));
    sd += drho[i];
    sw += W[i];
  }
  const int64_t iG = g + gb * (j + nvns * (t + 3 * G));
  drho[iG] = -sd;
  W[iG] = -sw;
}

// Phi = W_t + vxc (x) dw: the field of vmat_fxc + vmat_becke_fxc + vmat_becke_dw
// (hess_rks_becke.rs:850-908); dw [natm, 3, ng] (A, t, g).
__global__ void phi_kernel(double* __restrict__ W, const double* __restrict__ vxc, const double* __restrict__ dw,
                           int64_t ng, int64_t g0, int64_t gb, int64_t nvns, int64_t natm) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= gb * nvns * 3 * natm) return;
  const int64_t g = e % gb, j = (e / gb) % nvns, tA = e / (gb * nvns);
  W[e] += vxc[g0 + g + ng * j] * dw[tA * ng + g0 + g];
}

// c[g] = exc[g] sum_s rho_s[g, 0], the weights of the ddw contraction (hess_rks_becke.rs:976-978);
// 32 points per workgroup, 8 threads per point split the AO sum.
__global__ void __launch_bounds__(256) cddw_kernel(const double* __restrict__ ao, int64_t ng, int64_t nao, int64_t g0,
                                                   int64_t gb, const double* __restrict__ aod0,
                                                   const double* __restrict__ aod1, int nspin,
                                                   const double* __restrict__ exc, double* __restrict__ c) {
  __shared__ double red[8][32];
  const int lane = threadIdx.x & 31, part = threadIdx.x >> 5;
  const int64_t g = (int64_t)blockIdx.x * 32 + lane;
  double r = 0.0;
  if (g < gb)
    for (int64_t mu = part; mu < nao; mu += 8) {
      const double a = ao[g0 + g + ng * mu];
      r += a * aod0[g + gb * mu];
      if (nspin == 2) r += a * aod1[g + gb * mu];
    }
  red[part][lane] = r;
  __syncthreads();
  if (part == 0 && g < gb) {
    double v = 0.0;
    for (int p = 0; p < 8; ++p) v += red[p][lane];
    c[g0 + g] = exc[g0 + g] * v;
  }
}

// out[i] = sum_G in[G n + i]
__global__ void sum_blocks_kernel(const double* __restrict__ in, int64_t nblk, int64_t n, double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = 0.0;
  for (int64_t b = 0; b < nblk; ++b) v += in[b * n + i];
  out[i] = v;
}

// Per-grid-atom kernel of the vxc grid-shift parts (hess_rks_becke.rs:736-767, spin-summed as
// hess_uks_becke.rs:357-394): pv[G][u][t][s] = sum_spin dao_diag_G[u][pair(t, s)]
//   + sum_v (C_G[u + nao t, v + nao s] + C_G[v + nao s, u + nao t]) dm0[u, v].
// diag [G][split][u][6]; C_G at G * 9 nao^2.
__global__ void becke_pv_kernel(const double* __restrict__ diag0, const double* __restrict__ diag1,
                                const double* __restrict__ C0, const double* __restrict__ C1,
                                const double* __restrict__ D0, const double* __restrict__ D1, int nspin,
                                int64_t nsplit, int64_t nao, int64_t natm, double* __restrict__ pv) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= 9 * nao * natm) return;
  const int64_t ts = e % 9, u = (e / 9) % nao, G = e / (9 * nao);
  const int t = (int)(ts / 3), s = (int)(ts % 3);
  const int lo = t < s ? t : s, hi = t < s ? s : t;
  const int k = lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
  const int64_t ld = 3 * nao;
  double v = 0.0;
  for (int sp = 0; sp < nspin; ++sp) {
    const double* dg = sp ? diag1 : diag0;
    for (int64_t p = 0; p < nsplit; ++p) v += dg[((G * nsplit + p) * nao + u) * 6 + k];
    const double* C = (sp ? C1 : C0) + G * 9 * nao * nao;
    const double* D = sp ? D1 : D0;
    for (int64_t w = 0; w < nao; ++w)
      v += (C[(u + nao * t) + ld * (w + nao * s)] + C[(w + nao * s) + ld * (u + nao * t)]) * D[u + nao * w];
  }
  pv[e] = v;
}

// PA[G][A][t][s] = sum_{u in A} pv[G][u][t][s]
__global__ void becke_pa_kernel(const double* __restrict__ pv, const int64_t* __restrict__ slices, int64_t nao,
                                int64_t natm, double* __restrict__ PA) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= 9 * natm * natm) return;
  const int64_t ts = e % 9, A = (e / 9) % natm, G = e / (9 * natm);
  double v = 0.0;
  for (int64_t u = slices[2 * A]; u < slices[2 * A + 1]; ++u) v += pv[(G * nao + u) * 9 + ts];
  PA[e] = v;
}

// de_xc_skeleton[t, s, A, B] (hess_rks_becke.rs:1227-1237), a = t + 3 A, b = s + 3 B:
//   X[a, b] + X[b, a] - Xt[a, b]       de_fxc + de_becke_atom_1/2/3 + de_becke_full_1
//                                      (X = Phi^T drho_t, Xt = W_t^T drho_t)
//   + ddc[a n3 + b]                    de_becke_full_2
//   + Y[t, s, A, B] + Y[s, t, B, A]    de_becke_vxc_diag / off, Y[t, s, A, B] =
//                                      [A == B] / 2 sum_A' PA[B][A'][t][s] - PA[B][A][t][s]
//   + sum_spin de_vxc_diag + de_vxc_off
__global__ void becke_de_final_kernel(const double* __restrict__ X, const double* __restrict__ Xt,
                                      const double* __restrict__ ddc, const double* __restrict__ PA,
                                      const double* __restrict__ f0, const double* __restrict__ f1,
                                      const double* __restrict__ f2, const double* __restrict__ f3, int nfix,
                                      int64_t natm, double* __restrict__ de) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t n3 = 3 * natm;
  if (e >= n3 * n3) return;
  const int64_t t = e % 3, s = (e / 3) % 3, A = (e / 9) % natm, B = e / (9 * natm);
  const int64_t a = t + 3 * A, b = s + 3 * B;
  double v = X[a + n3 * b] + X[b + n3 * a] - Xt[a + n3 * b] + ddc[a * n3 + b];
  double y1 = -PA[(B * natm + A) * 9 + 3 * t + s], y2 = -PA[(A * natm + B) * 9 + 3 * s + t];
  if (A == B)
    for (int64_t Ap = 0; Ap < natm; ++Ap) {
      y1 += 0.5 * PA[(B * natm + Ap) * 9 + 3 * t + s];
      y2 += 0.5 * PA[(A * natm + Ap) * 9 + 3 * s + t];
    }
  v += y1 + y2;
  const double* f[4] = {f0, f1, f2, f3};
  for (int i = 0; i < nfix; ++i) v += f[i][e];
  de[e] = v;
}

// vmat_deriv1_grid[mu, nu, t, A] (hess_rks_becke.rs:1238): the bra + ket assembly of the Phi
// field (CfxT[nu, mu + nao tA]), the basis part -[mu in A] ip - [nu in A] ip^T, and vmat_becke_vxc = ip_A + ip_A^T of the
// points of atom A (hess_rks_becke.rs:893-895).
__global__ void vmat_grid_final_kernel(const double* __restrict__ Cfx, const double* __restrict__ Cip,
                                       const double* __restrict__ CipG, const int64_t* __restrict__ atom_of,
                                       int64_t nao, int64_t natm, double* __restrict__ out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= nao * nao * 3 * natm) return;
  const int64_t mu = e % nao, nu = (e / nao) % nao, tA = e / (nao * nao);
  const int64_t t = tA % 3, A = tA / 3;
  const int64_t ld = 3 * nao;
  double v = Cfx[nu + nao * (mu + nao * tA)] + Cfx[mu + nao * (nu + nao * tA)];
  if (atom_of[mu] == A) v -= Cip[(mu + nao * t) + ld * nu];
  if (atom_of[nu] == A) v -= Cip[(nu + nao * t) + ld * mu];
  const double* ipA = CipG + A * 3 * nao * nao;
  v += ipA[(mu + nao * t) + ld * nu] + ipA[(nu + nao * t) + ld * mu];
  out[e] = v;
}

// ---- products with a small output: C[a, b] += sum_k X[k + ld a] Y[k + ld b], a < na, b < nb <= 96 ----
// (the n3 x n3 contractions over grid points and rho components).  One workgroup per K chunk holds
// the whole output in registers, a thread owning a 6 x 6 sub-block; the chunk partials are summed
// in a fixed order by skinny_reduce_kernel.
constexpr int SK_T = 32;  // K tile
__global__ void __launch_bounds__(256) skinny_kernel(const double* __restrict__ X, const double* __restrict__ Y,
                                                     int64_t ld, int64_t K, int na, int nb, int64_t kchunk,
                                                     double* __restrict__ part) {
  __shared__ double xs[SK_T][96 + 1], ys[SK_T][96 + 1];
  const int tid = threadIdx.x;
  const int ra = (tid % 16) * 6, rb = (tid / 16) * 6;
  const int64_t k0 = blockIdx.x * kchunk, k1 = min(K, k0 + kchunk);
  double acc[6][6];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) acc[i][j] = 0.0;
  for (int64_t kt = k0; kt < k1; kt += SK_T) {
    for (int e = tid; e < SK_T * 96; e += 256) {
      const int kk = e % SK_T, c = e / SK_T;
      const int64_t k = kt + kk;
      xs[kk][c] = (k < k1 && c < na) ? X[k + ld * c] : 0.0;
      ys[kk][c] = (k < k1 && c < nb) ? Y[k + ld * c] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < SK_T; ++kk) {
      double x[6], y[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        x[i] = xs[kk][ra + i];
        y[i] = ys[kk][rb + i];
      }
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) acc[i][j] += x[i] * y[j];
    }
    __syncthreads();
  }
  double* P = part + blockIdx.x * (int64_t)na * nb;
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j)
      if (ra + i < na && rb + j < nb) P[(ra + i) + na * (rb + j)] = acc[i][j];
}

// 64 outputs per workgroup, the chunks split over 4 thread rows and summed in LDS (fixed order)
__global__ void __launch_bounds__(256) skinny_reduce_kernel(const double* __restrict__ part, int nchunk, int na,
                                                            int nb, double alpha, double* __restrict__ C,
                                                            int64_t ldc) {
  __shared__ double red[4][64];
  const int el = threadIdx.x & 63, row = threadIdx.x >> 6;
  const int e = blockIdx.x * 64 + el;
  const int nn = na * nb;
  double v0 = 0.0, v1 = 0.0;
  if (e < nn) {
    int c = row;
    for (; c + 4 < nchunk; c += 8) {
      v0 += part[(int64_t)c * nn + e];
      v1 += part[(int64_t)(c + 4) * nn + e];
    }
    if (c < nchunk) v0 += part[(int64_t)c * nn + e];
  }
  red[row][el] = v0 + v1;
  __syncthreads();
  if (row == 0 && e < nn) C[(e % na) + ldc * (e / na)] += alpha * (red[0][el] + red[1][el] + red[2][el] + red[3][el]);
}

void skinny_tn(int na, int nb, int64_t K, const double* X, const double* Y, int64_t ld, double* C, int64_t ldc,
               double alpha, hipStream_t s, DevBuf& ws) {
  HH_REQUIRE(na <= 96 && nb <= 96, "skinny product: output dimensions above 96");
  if (K <= 0) return;
  int64_t kchunk = std::max<int64_t>(SK_T * 4, ceil_div(ceil_div(K, 256), SK_T) * SK_T);
  const int nchunk = (int)ceil_div(K, kchunk);
  double* part = static_cast<double*>(ws.reserve(sizeof(double) * (size_t)nchunk * na * nb));
  hipLaunchKernelGGL(skinny_kernel, dim3((unsigned)nchunk), dim3(256), 0, s, X, Y, ld, K, na, nb, kchunk, part);
  HH_HIP(hipGetLastError());
  hipLaunchKernelGGL(skinny_reduce_kernel, dim3((unsigned)ceil_div(na * nb, 64)), dim3(256), 0, s, part, nchunk, na,
                     nb, alpha, C, ldc);
  HH_HIP(hipGetLastError());
}

// C[m, n] += alpha sum_k A[k, m] B[k, n] (+ A2/B2 over the second K range), FP64
void gemm_tn(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
             int64_t ldc, double alpha, hipStream_t s, DevBuf& ws, int64_t K1 = -1, const double* A2 = nullptr,
             int64_t lda2 = 0, const double* B2 = nullptr, int64_t ldb2 = 0) {
  GemmArgs g;
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.a_kmajor = true;
  g.B = B; g.ldb = ldb; g.b_kmajor = true;
  g.C = C; g.ldc = ldc;
  g.alpha = alpha; g.beta = 1.0;
  if (K1 >= 0) {
    g.K1 = K1; g.A2 = A2; g.lda2 = lda2; g.B2 = B2; g.ldb2 = ldb2;
  }
  gemm<double, double>(g, PREC_F64, s, ws);
}

// C[m, n] += alpha sum_k A[k + ld m] B[k + ld n] on the matrix cores (NT FP64); K = ld is the
// padded point count (a multiple of 32, zero rows beyond the batch)
void nt_acc(int64_t M, int64_t N, int64_t ld, const double* A, const double* B, double* C, int64_t ldc, double alpha,
            hipStream_t s, DevBuf& ws) {
  NtArgs a;
  a.M = M; a.N = N; a.K = ld;
  a.A = A; a.lda = ld; a.B = B; a.ldb = ld;
  a.C = C; a.ldc = ldc; a.alpha = alpha; a.beta = 1.0;
  a.ksplit = -1;
  gemm_nt(a, NT_F64, s, ws);
}

// fxc part of vmat_deriv1, transposed: CT[nu, mu + nao tA] += sum_g ao_b[g, nu] AW_b[g, (mu, tA)]
// over the AW blocks b (1, TAU 4) with the staged AO components b
void vmat_fx(int nblk, const double* stage, const double* AW, int64_t ldg, int64_t nao, int64_t M_fx, double* CT,
             hipStream_t s, DevBuf& ws) {
  for (int b = 0; b < nblk; ++b) nt_acc(nao, M_fx, ldg, stage + b * nao * ldg, AW + b * ldg * M_fx, CT, nao, 1.0, s, ws);
}

// out[g, mu, c] = sum_nu ao[g0 + g, nu, c] dm0[nu, mu], c < nc
void ao_dm0(const double* ao, int64_t ng, int64_t nao, int64_t g0, int64_t gb, int nc, const double* dm0, double* out,
            hipStream_t s, DevBuf& ws) {
  GemmArgs g;
  g.M = gb; g.N = nao; g.K = nao;
  g.A = ao + g0; g.lda = ng; g.a_kmajor = false; g.sA0 = ng * nao;
  g.B = dm0; g.ldb = nao; g.b_kmajor = true;
  g.C = out; g.ldc = gb; g.sC0 = gb * nao;
  g.nb0 = nc;
  gemm<double, double>(g, PREC_F64, s, ws);
}

#define HH_SETUP_DISPATCH(den, KERNEL, ...)                                   \
  do {                                                                        \
    if ((den) == DEN_RHO) hipLaunchKernelGGL(KERNEL<DEN_RHO>, __VA_ARGS__);   \
    else if ((den) == DEN_SIGMA) hipLaunchKernelGGL(KERNEL<DEN_SIGMA>, __VA_ARGS__); \
    else hipLaunchKernelGGL(KERNEL<DEN_TAU>, __VA_ARGS__);                    \
    HH_HIP(hipGetLastError());                                                \
  } while (0)

#define HH_SETUP_DISPATCH_NS(den, ns, KERNEL, ...)                                     \
  do {                                                                                 \
    if ((ns) == 1) {                                                                   \
      if ((den) == DEN_RHO) hipLaunchKernelGGL((KERNEL<DEN_RHO, 1>), __VA_ARGS__);     \
      else if ((den) == DEN_SIGMA) hipLaunchKernelGGL((KERNEL<DEN_SIGMA, 1>), __VA_ARGS__); \
      else hipLaunchKernelGGL((KERNEL<DEN_TAU, 1>), __VA_ARGS__);                      \
    } else {                                                                           \
      if ((den) == DEN_RHO) hipLaunchKernelGGL((KERNEL<DEN_RHO, 2>), __VA_ARGS__);     \
      else if ((den) == DEN_SIGMA) hipLaunchKernelGGL((KERNEL<DEN_SIGMA, 2>), __VA_ARGS__); \
      else hipLaunchKernelGGL((KERNEL<DEN_TAU, 2>), __VA_ARGS__);                      \
    }                                                                                  \
    HH_HIP(hipGetLastError());                                                         \
  } while (0)
}  // namespace

int setup_ao_ncomp(int den) { return den == DEN_RHO ? 10 : 20; }

void Numint::setup_rho(int den, int nspin, const double* const* dm0, double* rho) {
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "numint: den_type must be RHO, SIGMA or TAU");
  HH_REQUIRE(nspin == 1 || nspin == 2, "numint: nspin must be 1 or 2");
  const int nc = den_ncomp(den), nv = den_nvar(den);
  HH_REQUIRE(ncomp_ >= nc, "numint: AO component dimension insufficient");
  const int64_t gbmax = batch > 0 ? std::min(batch, ng_)
                                  : std::min<int64_t>(ng_, std::max<int64_t>(256, (1LL << 30) / (8 * nao_ * nc)));
  double* aod = static_cast<double*>(aob_.reserve(sizeof(double) * gbmax * nao_ * nc));
  for (int s = 0; s < nspin; ++s)
    for (int64_t g0 = 0; g0 < ng_; g0 += gbmax) {
      const int64_t gb = std::min(gbmax, ng_ - g0);
      ao_dm0(ao_.as<double>(), ng_, nao_, g0, gb, nc, dm0[s], aod, stream_, ws_);
      HH_SETUP_DISPATCH(den, setup_rho_kernel, dim3(nblocks(gb)), dim3(256), 0, stream_, ao_.as<double>(), ng_, nao_,
                        g0, gb, aod, rho + ng_ * nv * s);
    }
}

void Numint::hessian_setup(int den, int nspin, const double* const* dm0, int64_t natm, const int64_t* aoslice,
                           const double* vxc, const double* fxc, const SetupOut& out) {
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "numint: den_type must be RHO, SIGMA or TAU");
  HH_REQUIRE(nspin == 1 || nspin == 2, "numint: nspin must be 1 or 2");
  HH_REQUIRE(ncomp_ >= setup_ao_ncomp(den),
             "numint: the Hessian setup needs AO derivatives to order 2 (RHO, 10 components) or 3 (SIGMA/TAU, 20)");
  HH_REQUIRE(natm > 0 && aoslice && vxc && fxc && out.de_fxc, "numint: null argument or no atoms");
  std::vector<int64_t> sl(aoslice, aoslice + 2 * natm), atom_of(nao_, -1);
  for (int64_t A = 0; A < natm; ++A) {
    HH_REQUIRE(sl[2 * A] >= 0 && sl[2 * A] <= sl[2 * A + 1] && sl[2 * A + 1] <= nao_, "numint: AO slice out of range");
    for (int64_t mu = sl[2 * A]; mu < sl[2 * A + 1]; ++mu) {
      HH_REQUIRE(atom_of[mu] < 0, "numint: AO slices overlap");
      atom_of[mu] = A;
    }
  }
  for (int s = 0; s < nspin; ++s)
    HH_REQUIRE(dm0[s] && out.de_vxc_diag[s] && out.de_vxc_off[s] && out.vmat_ip[s] && out.vmat_deriv1[s],
               "numint: null per-spin argument");
  const int nc = den_ncomp(den), nv = den_nvar(den);
  const bool tau = den == DEN_TAU;
  const int nblk_aw = tau ? 4 : 1;
  const int64_t n3 = 3 * natm, M_fx = nao_ * n3;

  // per-point scratch (doubles): ao_dm0 per spin, drho + W, vxc operands, staged AO, fxc operand blocks
  const int64_t per_point = nspin * nao_ * nc + 2 * nv * nspin * n3 + 8 * nao_ + (tau ? 21 * nao_ : 0) +
                            nblk_aw * M_fx;
  int64_t gbmax = batch > 0 ? batch : (int64_t)(1LL << 30) / (8 * per_point);
  gbmax = std::max<int64_t>(64, (std::min(ng_, gbmax) + 63) / 64 * 64);

  // persistent accumulators
  const int64_t nsplit = std::max<int64_t>(1, std::min<int64_t>(64, ceil_div(2048, nao_)));
  const int64_t n_acc_fxc = n3 * n3, n_diag = nsplit * nao_ * 6, n_off = 9 * nao_ * nao_, n_ip = 3 * nao_ * nao_,
                n_fx = M_fx * nao_;
  const int64_t n_acc = n_acc_fxc + nspin * (n_diag + n_off + n_ip + n_fx);
  double* acc = static_cast<double*>(rho_.reserve(sizeof(double) * n_acc));
  fill<double>(acc, n_acc, 0.0, stream_);
  double* acc_fxc = acc;
  auto acc_diag = [&](int s) { return acc + n_acc_fxc + s * (n_diag + n_off + n_ip + n_fx); };
  auto acc_off = [&](int s) { return acc_diag(s) + n_diag; };
  auto acc_ip = [&](int s) { return acc_off(s) + n_off; };
  auto acc_fx = [&](int s) { return acc_ip(s) + n_ip; };

  // slices and atom map on the device
  int64_t* dsl = static_cast<int64_t*>(akb2_.reserve(sizeof(int64_t) * (2 * natm + nao_)));
  HH_HIP(hipMemcpyAsync(dsl, sl.data(), sizeof(int64_t) * 2 * natm, hipMemcpyHostToDevice, stream_));
  HH_HIP(hipMemcpyAsync(dsl + 2 * natm, atom_of.data(), sizeof(int64_t) * nao_, hipMemcpyHostToDevice, stream_));
  const int64_t* datom = dsl + 2 * natm;

  // batch scratch
  const int64_t n_aod = gbmax * nao_ * nc, n_dr = gbmax * nv * nspin * n3;
  const int64_t n_scr = nspin * n_aod + 2 * n_dr + 8 * gbmax * nao_ + (tau ? 21 * gbmax * nao_ : 0) +
                        nblk_aw * gbmax * M_fx;
  double* scr = static_cast<double*>(akb_.reserve(sizeof(double) * n_scr));
  double* aod[2] = {scr, scr + n_aod};
  double* drho = scr + nspin * n_aod;
  double* W = drho + n_dr;
  double* aowv = W + n_dr;                 // [ldg, nao, 3]
  double* aow = aowv + 3 * gbmax * nao_;   // [ldg, nao]
  double* H2 = aow + gbmax * nao_;         // TAU: [3 gb, 3 nao]
  double* H2w = H2 + (tau ? 9 * gbmax * nao_ : 0);
  double* AT = H2w + (tau ? 9 * gbmax * nao_ : 0);  // TAU: [3 gb, nao]
  double* stage = AT + (tau ? 3 * gbmax * nao_ : 0);  // [4][nao, ldg]
  double* AW = stage + 4 * gbmax * nao_;            // [nblk_aw][ldg, M_fx]

  const double* ao = ao_.as<double>();
  const double* w = w_.as<double>();
  const double fo = den == DEN_RHO ? 1.0 : 2.0;  // hess_rks.rs:364 / 377
  for (int64_t g0 = 0; g0 < ng_; g0 += gbmax) {
    const int64_t gb = std::min(gbmax, ng_ - g0), ldg = ceil_div(gb, 32) * 32;
    for (int s = 0; s < nspin; ++s) ao_dm0(ao, ng_, nao_, g0, gb, nc, dm0[s], aod[s], stream_, ws_);
    hipLaunchKernelGGL(stage_ao_kernel, dim3(nblocks(ldg * nao_ * 4)), dim3(256), 0, stream_, ao, ng_, nao_, g0, gb,
                       ldg, stage);
    HH_HIP(hipGetLastError());
    HH_SETUP_DISPATCH_NS(den, nspin, drho_w_kernel, dim3(nblocks(gb * natm)), dim3(256), 0, stream_, ao, ng_, nao_,
                         g0, gb, aod[0], aod[nspin - 1], dsl, natm, fxc, w, drho, W);
    // de_fxc: sum over (g, y, s) of W[(g,y,s), (t,A)] drho[(g,y,s), (s',B)]
    const int64_t Kf = gb * nv * nspin;
    if (n3 <= 96) skinny_tn((int)n3, (int)n3, Kf, W, drho, Kf, acc_fxc, n3, 1.0, stream_, ws2_);
    else gemm_tn(n3, n3, Kf, W, Kf, drho, Kf, acc_fxc, n3, 1.0, stream_, ws_);
    for (int s = 0; s < nspin; ++s) {
      const double* vxc_s = vxc + ng_ * nv * s;
      HH_SETUP_DISPATCH(den, vxc_diag_kernel, dim3((unsigned)nao_, (unsigned)nsplit), dim3(256), 0, stream_, ao, ng_,
                        nao_, g0, gb, aod[s], vxc_s, w, acc_diag(s), nsplit);
      HH_SETUP_DISPATCH(den, vxc_ops_kernel, dim3(nblocks(ldg * nao_)), dim3(256), 0, stream_, ao, ng_, nao_, g0, gb,
                        ldg, vxc_s, w, aowv, aow, H2, H2w, AT);
      const double* ao0 = ao + g0;
      const double* ao1 = ao + g0 + ng_ * nao_;  // components 1..3 as [gb, 3 nao], ld ng
      // de_vxc_off AO-pair tensor C[(mu, t), (nu, s)] (hess_rks.rs:359-400)
      nt_acc(3 * nao_, 3 * nao_, ldg, aowv, stage + nao_ * ldg, acc_off(s), 3 * nao_, fo, stream_, ws_);
      if (tau) gemm_tn(3 * nao_, 3 * nao_, 3 * gb, H2w, 3 * gb, H2, 3 * gb, acc_off(s), 3 * nao_, 1.0, stream_, ws_);
      // vmat_ip Cip[(mu, t), nu] (hess_rks.rs:421-469)
      if (den == DEN_RHO)
        gemm_tn(3 * nao_, nao_, gb, ao1, ng_, aow, ldg, acc_ip(s), 3 * nao_, 1.0, stream_, ws_);
      else
        gemm_tn(3 * nao_, nao_, 2 * gb, ao1, ng_, aow, ldg, acc_ip(s), 3 * nao_, 1.0, stream_, ws_, gb, aowv, ldg, ao0,
                ng_);
      if (tau) gemm_tn(3 * nao_, nao_, 3 * gb, H2, 3 * gb, AT, 3 * gb, acc_ip(s), 3 * nao_, 1.0, stream_, ws_);
      // fxc part of vmat_deriv1: CfxT[nu, (mu, t, A)]
      HH_SETUP_DISPATCH_NS(den, nspin, fxc_ops_kernel, dim3(nblocks(ldg * ceil_div(nao_, 4))), dim3(256), 0, stream_, ao,
                           ng_, nao_, g0, gb, ldg, W, s, natm, AW);
      vmat_fx(nblk_aw, stage, AW, ldg, nao_, M_fx, acc_fx(s), stream_, ws_);
    }
  }
  hipLaunchKernelGGL(de_fxc_final_kernel, dim3(nblocks(n3 * n3)), dim3(256), 0, stream_, acc_fxc, natm, out.de_fxc);
  HH_HIP(hipGetLastError());
  for (int s = 0; s < nspin; ++s) {
    hipLaunchKernelGGL(de_diag_final_kernel, dim3(nblocks(9 * natm * natm)), dim3(256), 0, stream_, acc_diag(s),
                       nsplit, nao_, dsl, natm, out.de_vxc_diag[s]);
    HH_HIP(hipGetLastError());
    hipLaunchKernelGGL(de_off_final_kernel, dim3((unsigned)(9 * natm * natm)), dim3(256), 0, stream_, acc_off(s), nao_,
                       dm0[s], dsl, natm, out.de_vxc_off[s]);
    HH_HIP(hipGetLastError());
    hipLaunchKernelGGL(vmat_ip_final_kernel, dim3(nblocks(3 * nao_ * nao_)), dim3(256), 0, stream_, acc_ip(s), nao_,
                       out.vmat_ip[s]);
    HH_HIP(hipGetLastError());
    hipLaunchKernelGGL(vmat_deriv1_final_kernel, dim3(nblocks(nao_ * nao_ * n3)), dim3(256), 0, stream_, acc_fx(s),
                       acc_ip(s), datom, nao_, natm, out.vmat_deriv1[s]);
    HH_HIP(hipGetLastError());
  }
}

void Numint::hessian_setup_becke(int den, int nspin, const double* const* dm0, int64_t natm, const int64_t* aoslice,
                                 const double* vxc, const double* fxc, const BeckeIn& bin, const SetupOut& out,
                                 const BeckeOut& bout) {
  HH_REQUIRE(den >= DEN_RHO && den <= DEN_TAU, "numint: den_type must be RHO, SIGMA or TAU");
  HH_REQUIRE(nspin == 1 || nspin == 2, "numint: nspin must be 1 or 2");
  HH_REQUIRE(ncomp_ >= setup_ao_ncomp(den),
             "numint: the Hessian setup needs AO derivatives to order 2 (RHO, 10 components) or 3 (SIGMA/TAU, 20)");
  HH_REQUIRE(natm > 0 && aoslice && vxc && fxc && bout.de_xc_skeleton, "numint: null argument or no atoms");
  HH_REQUIRE(bin.becke && bin.coords && bin.wquad && bin.split && bin.exc, "numint: null Becke argument");
  HH_REQUIRE(bin.becke->natm() == natm, "numint: the grid shift needs the full atom list (Becke tables of natm atoms)");
  HH_REQUIRE(bin.becke->stream() == stream_, "numint: the Becke tables must use the grid handle's stream");
  HH_REQUIRE(bin.split[0] == 0 && bin.split[natm] == ng_, "numint: atom grid boundaries must run from 0 to ngrids");
  for (int64_t A = 0; A < natm; ++A) HH_REQUIRE(bin.split[A] <= bin.split[A + 1], "numint: atom grid boundaries must be non-decreasing");
  std::vector<int64_t> sl(aoslice, aoslice + 2 * natm), atom_of(nao_, -1);
  for (int64_t A = 0; A < natm; ++A) {
    HH_REQUIRE(sl[2 * A] >= 0 && sl[2 * A] <= sl[2 * A + 1] && sl[2 * A + 1] <= nao_, "numint: AO slice out of range");
    for (int64_t mu = sl[2 * A]; mu < sl[2 * A + 1]; ++mu) {
      HH_REQUIRE(atom_of[mu] < 0, "numint: AO slices overlap");
      atom_of[mu] = A;
    }
  }
  for (int64_t mu = 0; mu < nao_; ++mu) HH_REQUIRE(atom_of[mu] >= 0, "numint: the grid shift needs every AO in an atom slice");
  bool fixed_vmat = true;
  for (int s = 0; s < nspin; ++s) {
    HH_REQUIRE(dm0[s] && bout.vmat_deriv1_grid[s], "numint: null per-spin argument");
    fixed_vmat = fixed_vmat && out.vmat_deriv1[s];
  }
  const int nc = den_ncomp(den), nv = den_nvar(den);
  const bool tau = den == DEN_TAU;
  const int nblk_aw = tau ? 4 : 1;
  const int64_t n3 = 3 * natm, M_fx = nao_ * n3, nvns = (int64_t)nv * nspin;

  const int64_t per_point = nspin * nao_ * nc + 2 * nv * nspin * n3 + 8 * nao_ + (tau ? 21 * nao_ : 0) +
                            nblk_aw * M_fx;
  int64_t gbmax = batch > 0 ? batch : (int64_t)(1LL << 30) / (8 * per_point);
  gbmax = std::max<int64_t>(64, (std::min(ng_, gbmax) + 63) / 64 * 64);

  // accumulators: [Xf | Xt | X] then per spin [diag_G | C_G | Cip_G | Fgrid | Ffixed]
  const int64_t nsplit = std::max<int64_t>(1, std::min<int64_t>(64, ceil_div(2048, nao_)));
  const int64_t n_d = natm * nsplit * nao_ * 6, n_o = natm * 9 * nao_ * nao_, n_i = natm * 3 * nao_ * nao_,
                n_f = M_fx * nao_, n_ff = fixed_vmat ? M_fx * nao_ : 0;
  const int64_t n_spin = n_d + n_o + n_i + n_f + n_ff;
  const int64_t n_acc = 3 * n3 * n3 + nspin * n_spin;
  double* acc = static_cast<double*>(rho_.reserve(sizeof(double) * n_acc));
  fill<double>(acc, n_acc, 0.0, stream_);
  double* acc_xf = acc;
  double* acc_xt = acc + n3 * n3;
  double* acc_x = acc + 2 * n3 * n3;
  auto acc_d = [&](int s) { return acc + 3 * n3 * n3 + s * n_spin; };
  auto acc_o = [&](int s) { return acc_d(s) + n_d; };
  auto acc_i = [&](int s) { return acc_o(s) + n_o; };
  auto acc_f = [&](int s) { return acc_i(s) + n_i; };
  auto acc_ff = [&](int s) { return acc_f(s) + n_f; };

  // whole-grid buffers: dw, ddw weights c, ddc, C / Cip sums, pv, PA, grid-fixed de scratch
  const int64_t n_fixde = 9 * natm * natm;
  const int64_t n_aux = n3 * ng_ + ng_ + n3 * n3 + nspin * 12 * nao_ * nao_ + 9 * natm * nao_ + 9 * natm * natm +
                        2 * nspin * n_fixde;
  double* aux = static_cast<double*>(aob_.reserve(sizeof(double) * n_aux));
  double* dw = aux;
  double* cw = dw + n3 * ng_;
  double* ddc = cw + ng_;
  double* csum = ddc + n3 * n3;                 // per spin [C sum 9 nao^2 | Cip sum 3 nao^2]
  double* pv = csum + nspin * 12 * nao_ * nao_;
  double* PA = pv + 9 * natm * nao_;
  double* fixde = PA + 9 * natm * natm;         // per spin [diag | off]
  auto c_sum = [&](int s) { return csum + s * 12 * nao_ * nao_; };
  auto ip_sum = [&](int s) { return c_sum(s) + 9 * nao_ * nao_; };

  int64_t* dsl = static_cast<int64_t*>(akb2_.reserve(sizeof(int64_t) * (2 * natm + nao_)));
  HH_HIP(hipMemcpyAsync(dsl, sl.data(), sizeof(int64_t) * 2 * natm, hipMemcpyHostToDevice, stream_));
  HH_HIP(hipMemcpyAsync(dsl + 2 * natm, atom_of.data(), sizeof(int64_t) * nao_, hipMemcpyHostToDevice, stream_));
  const int64_t* datom = dsl + 2 * natm;

  const int64_t n_aod = gbmax * nao_ * nc, n_dr = gbmax * nvns * n3;
  const int64_t n_scr = nspin * n_aod + 2 * n_dr + 8 * gbmax * nao_ + (tau ? 21 * gbmax * nao_ : 0) +
                        nblk_aw * gbmax * M_fx;
  double* scr = static_cast<double*>(akb_.reserve(sizeof(double) * n_scr));
  double* aod[2] = {scr, scr + n_aod};
  double* drho = scr + nspin * n_aod;
  double* W = drho + n_dr;
  double* aowv = W + n_dr;
  double* aow = aowv + 3 * gbmax * nao_;
  double* H2 = aow + gbmax * nao_;
  double* H2w = H2 + (tau ? 9 * gbmax * nao_ : 0);
  double* AT = H2w + (tau ? 9 * gbmax * nao_ : 0);
  double* stage = AT + (tau ? 3 * gbmax * nao_ : 0);
  double* AW = stage + 4 * gbmax * nao_;

  // Becke weight derivatives dw [natm, 3, ng] of the atom-grouped grid
  {
    Becke::Out bo;
    bo.dw = dw;
    bin.becke->partition(bin.coords, ng_, true, bin.split, bin.wquad, bin.hardness, 1, bo);
  }

  const double* ao = ao_.as<double>();
  const double* w = w_.as<double>();
  const double fo = den == DEN_RHO ? 1.0 : 2.0;
  for (int64_t G = 0; G < natm; ++G)
    for (int64_t g0 = bin.split[G]; g0 < bin.split[G + 1]; g0 += gbmax) {
      const int64_t gb = std::min(gbmax, bin.split[G + 1] - g0), ldg = ceil_div(gb, 32) * 32;
      for (int s = 0; s < nspin; ++s) ao_dm0(ao, ng_, nao_, g0, gb, nc, dm0[s], aod[s], stream_, ws_);
      hipLaunchKernelGGL(stage_ao_kernel, dim3(nblocks(ldg * nao_ * 4)), dim3(256), 0, stream_, ao, ng_, nao_, g0, gb,
                         ldg, stage);
      HH_HIP(hipGetLastError());
      HH_SETUP_DISPATCH_NS(den, nspin, drho_w_kernel, dim3(nblocks(gb * natm)), dim3(256), 0, stream_, ao, ng_, nao_,
                           g0, gb, aod[0], aod[nspin - 1], dsl, natm, fxc, w, drho, W);
      const int64_t Kf = gb * nvns;
      auto fxc_vmat = [&](int s, double* C) {
        HH_SETUP_DISPATCH_NS(den, nspin, fxc_ops_kernel, dim3(nblocks(ldg * ceil_div(nao_, 4))), dim3(256), 0, stream_,
                             ao, ng_, nao_, g0, gb, ldg, W, s, natm, AW);
        vmat_fx(nblk_aw, stage, AW, ldg, nao_, M_fx, C, stream_, ws_);
      };
      // grid-fixed de_fxc and vmat_deriv1 (fxc part)
      auto xprod = [&](double* C) {
        if (n3 <= 96) skinny_tn((int)n3, (int)n3, Kf, W, drho, Kf, C, n3, 1.0, stream_, ws2_);
        else gemm_tn(n3, n3, Kf, W, Kf, drho, Kf, C, n3, 1.0, stream_, ws_);
      };
      if (out.de_fxc) xprod(acc_xf);
      if (fixed_vmat)
        for (int s = 0; s < nspin; ++s) fxc_vmat(s, acc_ff(s));
      // the point carried by G: total drho and W, their product, then the Phi field
      hipLaunchKernelGGL(shift_fix_kernel, dim3(nblocks(gb * nvns * 3)), dim3(256), 0, stream_, drho, W, gb, nvns,
                         natm, G);
      HH_HIP(hipGetLastError());
      xprod(acc_xt);
      hipLaunchKernelGGL(phi_kernel, dim3(nblocks(gb * nvns * n3)), dim3(256), 0, stream_, W, vxc, dw, ng_, g0, gb,
                         nvns, natm);
      HH_HIP(hipGetLastError());
      xprod(acc_x);
      hipLaunchKernelGGL(cddw_kernel, dim3((unsigned)ceil_div(gb, 32)), dim3(256), 0, stream_, ao, ng_, nao_, g0, gb, aod[0],
                         aod[nspin - 1], nspin, bin.exc, cw);
      HH_HIP(hipGetLastError());
      for (int s = 0; s < nspin; ++s) {
        const double* vxc_s = vxc + ng_ * nv * s;
        HH_SETUP_DISPATCH(den, vxc_diag_kernel, dim3((unsigned)nao_, (unsigned)nsplit), dim3(256), 0, stream_, ao, ng_,
                          nao_, g0, gb, aod[s], vxc_s, w, acc_d(s) + G * nsplit * nao_ * 6, nsplit);
        HH_SETUP_DISPATCH(den, vxc_ops_kernel, dim3(nblocks(ldg * nao_)), dim3(256), 0, stream_, ao, ng_, nao_, g0,
                          gb, ldg, vxc_s, w, aowv, aow, H2, H2w, AT);
        const double* ao0 = ao + g0;
        const double* ao1 = ao + g0 + ng_ * nao_;
        double* CG = acc_o(s) + G * 9 * nao_ * nao_;
        double* IG = acc_i(s) + G * 3 * nao_ * nao_;
        nt_acc(3 * nao_, 3 * nao_, ldg, aowv, stage + nao_ * ldg, CG, 3 * nao_, fo, stream_, ws_);
        if (tau) gemm_tn(3 * nao_, 3 * nao_, 3 * gb, H2w, 3 * gb, H2, 3 * gb, CG, 3 * nao_, 1.0, stream_, ws_);
        if (den == DEN_RHO)
          gemm_tn(3 * nao_, nao_, gb, ao1, ng_, aow, ldg, IG, 3 * nao_, 1.0, stream_, ws_);
        else
          gemm_tn(3 * nao_, nao_, 2 * gb, ao1, ng_, aow, ldg, IG, 3 * nao_, 1.0, stream_, ws_, gb, aowv, ldg, ao0,
                  ng_);
        if (tau) gemm_tn(3 * nao_, nao_, 3 * gb, H2, 3 * gb, AT, 3 * gb, IG, 3 * nao_, 1.0, stream_, ws_);
        fxc_vmat(s, acc_f(s));
      }
    }

  // de_becke_full_2: ddw contracted with exc rho[0]
  {
    Becke::Out bo;
    bo.contract_ddw = cw;
    bo.nset_ddw = 1;
    bo.ddc = ddc;
    bin.becke->partition(bin.coords, ng_, true, bin.split, bin.wquad, bin.hardness, 2, bo);
  }

  // grid-fixed keys
  if (out.de_fxc) {
    hipLaunchKernelGGL(de_fxc_final_kernel, dim3(nblocks(n3 * n3)), dim3(256), 0, stream_, acc_xf, natm, out.de_fxc);
    HH_HIP(hipGetLastError());
  }
  const double* fx[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int s = 0; s < nspin; ++s) {
    hipLaunchKernelGGL(sum_blocks_kernel, dim3(nblocks(9 * nao_ * nao_)), dim3(256), 0, stream_, acc_o(s), natm,
                       9 * nao_ * nao_, c_sum(s));
    HH_HIP(hipGetLastError());
    hipLaunchKernelGGL(sum_blocks_kernel, dim3(nblocks(3 * nao_ * nao_)), dim3(256), 0, stream_, acc_i(s), natm,
                       3 * nao_ * nao_, ip_sum(s));
    HH_HIP(hipGetLastError());
    double* dd = out.de_vxc_diag[s] ? out.de_vxc_diag[s] : fixde + 2 * s * n_fixde;
    double* doff = out.de_vxc_off[s] ? out.de_vxc_off[s] : fixde + (2 * s + 1) * n_fixde;
    hipLaunchKernelGGL(de_diag_final_kernel, dim3(nblocks(9 * natm * natm)), dim3(256), 0, stream_, acc_d(s),
                       natm * nsplit, nao_, dsl, natm, dd);
    HH_HIP(hipGetLastError());
    hipLaunchKernelGGL(de_off_final_kernel, dim3((unsigned)(9 * natm * natm)), dim3(256), 0, stream_, c_sum(s), nao_,
                       dm0[s], dsl, natm, doff);
    HH_HIP(hipGetLastError());
    fx[2 * s] = dd;
    fx[2 * s + 1] = doff;
    if (out.vmat_ip[s]) {
      hipLaunchKernelGGL(vmat_ip_final_kernel, dim3(nblocks(3 * nao_ * nao_)), dim3(256), 0, stream_, ip_sum(s), nao_,
                         out.vmat_ip[s]);
      HH_HIP(hipGetLastError());
    }
    if (fixed_vmat) {
      hipLaunchKernelGGL(vmat_deriv1_final_kernel, dim3(nblocks(nao_ * nao_ * n3)), dim3(256), 0, stream_, acc_ff(s),
                         ip_sum(s), datom, nao_, natm, out.vmat_deriv1[s]);
      HH_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(vmat_grid_final_kernel, dim3(nblocks(nao_ * nao_ * n3)), dim3(256), 0, stream_, acc_f(s),
                       ip_sum(s), acc_i(s), datom, nao_, natm, bout.vmat_deriv1_grid[s]);
    HH_HIP(hipGetLastError());
  }
  // grid-shift vxc parts and the assembly
  hipLaunchKernelGGL(becke_pv_kernel, dim3(nblocks(9 * nao_ * natm)), dim3(256), 0, stream_, acc_d(0),
                     acc_d(nspin - 1), acc_o(0), acc_o(nspin - 1), dm0[0], dm0[nspin - 1], nspin, nsplit, nao_, natm,
                     pv);
  HH_HIP(hipGetLastError());
  hipLaunchKernelGGL(becke_pa_kernel, dim3(nblocks(9 * natm * natm)), dim3(256), 0, stream_, pv, dsl, nao_, natm, PA);
  HH_HIP(hipGetLastError());
  hipLaunchKernelGGL(becke_de_final_kernel, dim3(nblocks(n3 * n3)), dim3(256), 0, stream_, acc_x, acc_xt, ddc, PA,
                     fx[0], fx[1], fx[2], fx[3], 2 * nspin, natm, bout.de_xc_skeleton);
  HH_HIP(hipGetLastError());
}

}  // namespace hh
