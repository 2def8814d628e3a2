// Block Krylov solver for (I + A) x = b, GPU resident (reference: src/util/krylov_block.rs).
//
// The basis is kept non-normalised, as in the reference, so the squared norms of the new trial
// vectors are the convergence signal.  Orthogonalisation of a block (the reference's serial
// column-by-column MGS, krylov_block.rs:225-253) runs in groups of up to GB columns: a group is
// first projected against the columns already kept with two classical Gram-Schmidt passes on
// the matrix cores (CGS2); inside the group, Gram-Schmidt in column order is the LDL^T
// factorisation of the group's Gram matrix G = V^T V: the Schur complement d_t of pivot t is the
// squared norm MGS would find for column t, the reference's drop rule (||v||^2 <= lindep) drops
// pivot t (its column of L is zeroed, so it projects nothing), and the kept vectors are
// V L^{-T}.  The factorisation runs in one workgroup on the group's m x m Gram matrix in LDS;
// the products over the long axis run on the whole chip.  Two such passes (the second on the
// first pass's output, whose Gram matrix is nearly diagonal) restore orthogonality to working
// precision; the second pass's Schur complements are the returned norms.  Reductions run in a
// fixed order, so results are reproducible run to run.
#include "krylov.hpp"

#include "elementwise.hpp"
#include "krylov_dev.hpp"
#include "linalg.hpp"

namespace hh {

namespace {
constexpr int GB = KR_GB;
constexpr int LDL_T = KR_LDL_T;

// LDL^T with the drop rule on the group's Gram matrix G [m, m] (every kept pivot selected):
// V Tm holds the kept Gram-Schmidt vectors in order, dout[c] their squared norms, *kout the count.
__global__ __launch_bounds__(LDL_T) void ldl_drop_kernel(const double* __restrict__ G, int m, double lindep,
                                                         double* __restrict__ Tm, double* __restrict__ dout,
                                                         int* __restrict__ kout) {
  ldl_drop_block(G, m, lindep, -__builtin_huge_val(), Tm, dout, kout, nullptr);
}

__global__ void scale_rows_kernel(double* D, int64_t rows, int64_t cols, const double* dn) {
  const int64_t total = rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x)
    D[e] /= dn[e % rows];
}

__global__ void add_diag_kernel(double* H, int64_t nd, const double* d) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nd; i += (int64_t)gridDim.x * blockDim.x)
    H[i + i * nd] += d[i];
}

// out = b - x - ax
__global__ void residual_kernel(const double* b, const double* x, const double* ax, double* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = b[i] - x[i] - ax[i];
}

unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 8192)); }

struct Scratch {
  DevBuf d, dn, dk, h, g, diag, tmp, g2, tm, w;
};
Scratch& scratch() {
  static thread_local Scratch* s = new Scratch;  // lives for the process
  return *s;
}

void d2d(double* dst, const double* src, int64_t count, hipStream_t s) {
  if (count > 0) HH_HIP(hipMemcpyAsync(dst, src, sizeof(double) * count, hipMemcpyDeviceToDevice, s));
}

// target[n, nset] += xs c,  with (xs^T ax + diag(inner)) c = xs^T b   (krylov_block.rs:204-215)
void projected_solve_add(const double* xs, const double* ax, const std::vector<double>& inner, int64_t n, int64_t nd,
                         const double* bsrc, int64_t nset, double* target, hipStream_t s, DevBuf& ws) {
  Scratch& sc = scratch();
  double* H = static_cast<double*>(sc.h.reserve(sizeof(double) * nd * nd));
  double* G = static_cast<double*>(sc.g.reserve(sizeof(double) * nd * nset));
  double* dg = static_cast<double*>(sc.diag.reserve(sizeof(double) * nd));
  HH_HIP(hipMemcpyAsync(dg, inner.data(), sizeof(double) * nd, hipMemcpyHostToDevice, s));
  GemmArgs g;
  g.M = nd; g.N = nd; g.K = n;
  g.A = xs; g.lda = n; g.a_kmajor = true;
  g.B = ax; g.ldb = n; g.b_kmajor = true;
  g.C = H; g.ldc = nd;
  gemm<double, double>(g, PREC_F64, s, ws);
  hipLaunchKernelGGL(add_diag_kernel, dim3(nblk(nd)), dim3(256), 0, s, H, nd, dg);
  HH_HIP(hipGetLastError());
  GemmArgs r;
  r.M = nd; r.N = nset; r.K = n;
  r.A = xs; r.lda = n; r.a_kmajor = true;
  r.B = bsrc; r.ldb = n; r.b_kmajor = true;
  r.C = G; r.ldc = nd;
  gemm<double, double>(r, PREC_F64, s, ws);
  lu_solve(H, nd, G, nset, s, ws);
  GemmArgs u;
  u.M = n; u.N = nset; u.K = nd;
  u.A = xs; u.lda = n; u.a_kmajor = false;
  u.B = G; u.ldb = nd; u.b_kmajor = true;
  u.C = target; u.ldc = n;
  u.beta = 1.0;
  gemm<double, double>(u, PREC_F64, s, ws);
}

}  // namespace

int64_t orth_block(double* V, int64_t n, int64_t m, double lindep, double* R, int64_t r_off,
                   std::vector<double>& norms, hipStream_t s, DevBuf& ws) {
  Scratch& sc = scratch();
  const int64_t gmax = std::min<int64_t>(std::max<int64_t>(m, 1), GB);
  double* D = static_cast<double*>(sc.d.reserve(sizeof(double) * std::max<int64_t>(m, 1) * gmax));
  double* dn = static_cast<double*>(sc.dn.reserve(sizeof(double) * (std::max<int64_t>(m, 1) + GB)));
  int* dk = static_cast<int*>(sc.dk.reserve(64));
  double* Gm = static_cast<double*>(sc.g2.reserve(sizeof(double) * gmax * gmax));
  double* Tm = static_cast<double*>(sc.tm.reserve(sizeof(double) * gmax * gmax));
  double* W = static_cast<double*>(sc.w.reserve(sizeof(double) * n * gmax));
  double* d1 = dn + std::max<int64_t>(m, 1);  // first-pass pivots (not returned)
  double* Rk = R + r_off * n;
  int64_t kept = 0;
  std::vector<double> hn(GB);
  // G = X^T X (cb x cb), then Tm / pivots / count from the LDL kernel, then Y = X Tm
  auto pass = [&](const double* X, int cb, double* Y, double* dout) {
    GemmArgs g;
    g.M = cb; g.N = cb; g.K = n;
    g.A = X; g.lda = n; g.a_kmajor = true;
    g.B = X; g.ldb = n; g.b_kmajor = true;
    g.C = Gm; g.ldc = cb;
    gemm<double, double>(g, PREC_F64, s, ws);
    hipLaunchKernelGGL(ldl_drop_kernel, dim3(1), dim3(LDL_T), 0, s, Gm, cb, lindep, Tm, dout, dk);
    HH_HIP(hipGetLastError());
    GemmArgs u;
    u.M = n; u.N = cb; u.K = cb;
    u.A = X; u.lda = n; u.a_kmajor = false;
    u.B = Tm; u.ldb = cb; u.b_kmajor = true;
    u.C = Y; u.ldc = n;
    gemm<double, double>(u, PREC_F64, s, ws);
  };
  for (int64_t c0 = 0; c0 < m; c0 += GB) {
    const int cb = (int)std::min<int64_t>(GB, m - c0);
    double* Vc = V + c0 * n;
    if (kept > 0) {
      for (int p = 0; p < 2; ++p) {
        GemmArgs g;  // D = Rk^T Vc
        g.M = kept; g.N = cb; g.K = n;
        g.A = Rk; g.lda = n; g.a_kmajor = true;
        g.B = Vc; g.ldb = n; g.b_kmajor = true;
        g.C = D; g.ldc = kept;
        gemm<double, double>(g, PREC_F64, s, ws);
        hipLaunchKernelGGL(scale_rows_kernel, dim3(nblk(kept * cb)), dim3(256), 0, s, D, kept, (int64_t)cb, dn);
        HH_HIP(hipGetLastError());
        GemmArgs u;  // Vc -= Rk D
        u.M = n; u.N = cb; u.K = kept;
        u.A = Rk; u.lda = n; u.a_kmajor = false;
        u.B = D; u.ldb = kept; u.b_kmajor = true;
        u.C = Vc; u.ldc = n;
        u.alpha = -1.0; u.beta = 1.0;
        gemm<double, double>(u, PREC_F64, s, ws);
      }
    }
    // two Gram / LDL passes: Vc -> W (kept columns first, zero columns after), W -> R[:, kept..)
    // (columns past the kept ones are zero and are overwritten by the next group)
    pass(Vc, cb, W, d1);
    pass(W, cb, Rk + kept * n, dn + kept);
    int k_new = 0;
    HH_HIP(hipMemcpyAsync(&k_new, dk, sizeof(int), hipMemcpyDeviceToHost, s));
    HH_HIP(hipStreamSynchronize(s));
    if (k_new > 0) {
      HH_HIP(hipMemcpyAsync(hn.data(), dn + kept, sizeof(double) * k_new, hipMemcpyDeviceToHost, s));
      HH_HIP(hipStreamSynchronize(s));
      for (int i = 0; i < k_new; ++i) norms.push_back(hn[i]);
    }
    kept += k_new;
  }
  return kept;
}

void krylov_block(const Aop& aop, const double* b, const double* x0, int64_t n, int64_t nset, double tol,
                  int max_cycle, int max_space, double lindep, double* x, KrylovStats* stats, hipStream_t s) {
  HH_REQUIRE(n > 0 && nset > 0, "krylov_block: empty system");
  HH_REQUIRE(max_space >= 1 && max_cycle >= 0, "krylov_block: max_space must be >= 1");
  KrylovStats st;
  DevBuf ws, bufs;
  const int64_t max_basis = nset * (int64_t)(max_space + 1);
  const int64_t blk = n * nset;
  double* base = static_cast<double*>(bufs.reserve(sizeof(double) * (2 * n * max_basis + 6 * blk)));
  double* xs = base;
  double* ax = xs + n * max_basis;
  double* xacc = ax + n * max_basis;
  double* bres = xacc + blk;
  double* x1 = bres + blk;
  double* axt = x1 + blk;
  double* xnew = axt + blk;
  double* xtmp = xnew + blk;
  Scratch& sc = scratch();
  double* coeffs = static_cast<double*>(sc.tmp.reserve(sizeof(double) * max_basis * nset));
  double* dinner = static_cast<double*>(sc.diag.reserve(sizeof(double) * max_basis));
  (void)dinner;

  if (x0) d2d(xacc, x0, blk, s);
  else fill<double>(xacc, blk, 0.0, s);
  const double conv = std::max(lindep, tol * tol);
  int total = 0, restart = 0;
  int64_t last_nd = 0;
  double last_max = 0.0;
  std::vector<double> inner_all, ip;
  bool have_basis = false;

  while (total < max_cycle) {
    if (restart == 0 && !x0) {
      d2d(bres, b, blk, s);
    } else {
      aop(xacc, nset, xtmp);
      ++st.aop_calls;
      hipLaunchKernelGGL(residual_kernel, dim3(nblk(blk)), dim3(256), 0, s, b, xacc, xtmp, bres, blk);
      HH_HIP(hipGetLastError());
    }
    // orthogonalise the residual block (V is scratch: copy first)
    d2d(xnew, bres, blk, s);
    ip.clear();
    int64_t nb = orth_block(xnew, n, nset, lindep, x1, 0, ip, s, ws);
    if (nb == 0) {
      last_nd = 0;
      break;
    }
    inner_all.clear();
    int64_t nd = 0;
    bool converged = false;
    for (int inner = 0; inner < max_space; ++inner) {
      if (total >= max_cycle) break;
      ++total;
      aop(x1, nb, axt);
      ++st.aop_calls;
      d2d(xs + nd * n, x1, nb * n, s);
      d2d(ax + nd * n, axt, nb * n, s);
      inner_all.insert(inner_all.end(), ip.begin(), ip.end());
      nd += nb;
      have_basis = true;
      // coeffs = (xs^T axt) / inner ;  x1_new = axt - xs coeffs
      {
        double* dinn = static_cast<double*>(sc.diag.reserve(sizeof(double) * max_basis));
        HH_HIP(hipMemcpyAsync(dinn, inner_all.data(), sizeof(double) * nd, hipMemcpyHostToDevice, s));
        GemmArgs g;
        g.M = nd; g.N = nb; g.K = n;
        g.A = xs; g.lda = n; g.a_kmajor = true;
        g.B = axt; g.ldb = n; g.b_kmajor = true;
        g.C = coeffs; g.ldc = nd;
        gemm<double, double>(g, PREC_F64, s, ws);
        hipLaunchKernelGGL(scale_rows_kernel, dim3(nblk(nd * nb)), dim3(256), 0, s, coeffs, nd, nb, dinn);
        HH_HIP(hipGetLastError());
        d2d(xnew, axt, nb * n, s);
        GemmArgs u;
        u.M = n; u.N = nb; u.K = nd;
        u.A = xs; u.lda = n; u.a_kmajor = false;
        u.B = coeffs; u.ldb = nd; u.b_kmajor = true;
        u.C = xnew; u.ldc = n;
        u.alpha = -1.0; u.beta = 1.0;
        gemm<double, double>(u, PREC_F64, s, ws);
        // the host copy of inner_all must stay alive until the copy above ran
        HH_HIP(hipStreamSynchronize(s));
      }
      std::vector<double> nip;
      const int64_t kept = orth_block(xnew, n, nb, lindep, xtmp, 0, nip, s, ws);
      double max_ip = 0.0;
      for (double v : nip) max_ip = std::max(max_ip, v);
      // keep only directions above the convergence threshold, in order
      ip.clear();
      int64_t nk = 0;
      for (int64_t c = 0; c < kept;) {  // runs of kept columns, one copy each
        if (!(nip[c] > conv)) {
          ++c;
          continue;
        }
        int64_t e = c;
        while (e < kept && nip[e] > conv) ip.push_back(nip[e++]);
        d2d(x1 + nk * n, xtmp + c * n, (e - c) * n, s);
        nk += e - c;
        c = e;
      }
      nb = nk;
      last_max = max_ip;
      if (max_ip < conv) {
        converged = true;
        break;
      }
      if (nb == 0) break;  // every new direction was filtered (cannot happen when max_ip >= conv)
    }
    last_nd = nd;
    if (converged || total >= max_cycle) break;
    // hard restart
    projected_solve_add(xs, ax, inner_all, n, nd, bres, nset, xacc, s, ws);
    ++restart;
  }
  if (last_nd > 0 && have_basis) projected_solve_add(xs, ax, inner_all, n, last_nd, bres, nset, xacc, s, ws);
  d2d(x, xacc, blk, s);
  HH_HIP(hipStreamSynchronize(s));
  st.cycles = total;
  st.restarts = restart;
  st.max_innerprod = last_max;
  if (stats) *stats = st;
  if (last_max >= conv) {
    char buf[256];
    snprintf(buf, sizeof buf,
             "krylov_block failed to converge: max(||v||^2) = %.3e >= tol^2 = %.3e after %d cycles (%d restarts, "
             "max_cycle = %d, max_space = %d)",
             last_max, conv, total, restart, max_cycle, max_space);
    throw Error(HH_ERR_NOT_CONVERGED, buf);
  }
}

}  // namespace hh
