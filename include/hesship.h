/* hesship: MI355X (gfx950) data-parallel core of analytic SCF/DFT Hessians.
 *
 * C ABI.  Conventions shared by every entry point:
 *  - Tensors are FP64, column-major, with the same logical shapes as the Rust reference
 *    (ajz34/rstsr-showcase-hessian): e.g. bra [nao, nocc, nprop] has element (mu, i, a) at
 *    mu + nao * (i + nocc * a).  Array arguments are DEVICE pointers unless the comment says
 *    "host".
 *  - Every call is enqueued on the handle's HIP stream (or `stream`, NULL = default stream)
 *    and is asynchronous with respect to the host unless stated otherwise.
 *  - Every call returns an hh_status.  The reference panics where this ABI returns an error
 *    code; hh_last_error() gives the message (thread-local, valid until the next call).
 */
#ifndef HESSHIP_H
#define HESSHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t hh_status;
#define HH_OK 0
#define HH_ERR_INVALID 1          /* shape / argument check failed (reference: check_shape! panic) */
#define HH_ERR_HIP 2              /* HIP runtime error */
#define HH_ERR_NOT_CONVERGED 3    /* Krylov did not converge (reference: panic, krylov_block.rs:186-194) */
#define HH_ERR_NOMEM 4            /* device allocation failed */
#define HH_ERR_STATE 5            /* call made before the preparation it needs */
#define HH_ERR_INTERNAL 6

/* Arithmetic of the contraction kernels.  FP64 is the reference arithmetic. */
typedef int32_t hh_precision;
#define HH_PREC_FP64 0
#define HH_PREC_FP32 1            /* FP32 MFMA, FP32 accumulate */
#define HH_PREC_BF16X3 2          /* split-BF16 x3 products on MFMA, FP32 accumulate */

const char* hh_last_error(void);
const char* hh_version(void);

/* ---------------------------------------------------------------------------------------- */
/* Triangular packing (src/util/density_matrices.rs:48-77 pack_triu_tilde; rstsr unpack_tri) */

/* out[nao_tp, nmat] = packed upper triangle of dm[nao, nao, nmat], off-diagonals x2
 * (replaces pack_triu_tilde, density_matrices.rs:61-77). */
hh_status hh_pack_triu_tilde(const double* dm, int64_t nao, int64_t nmat, double* out, void* stream);
/* out[nao, nao, nmat] = symmetric matrix from packed upper triangle tp[nao_tp, nmat]
 * (replaces rstsr `unpack_tri(Upper, FlagSymm::Sy)` as used in hess_r.rs:139,160). */
hh_status hh_unpack_triu(const double* tp, int64_t nao, int64_t nmat, double* out, void* stream);

/* ---------------------------------------------------------------------------------------- */
/* RI-JK response (src/ri_jk/hess_r.rs, src/ri_jk/hess_u.rs)                                 */

typedef struct hh_rijk hh_rijk;

/* Create an RI-JK component from device-resident cderi [nao_tp, naux] (packed upper, column
 * major).  The handle keeps its own copy of cderi and derived layouts; the caller's buffer may
 * be freed after the call returns.  nspin = 1 (RHF, replaces RHessRIJK::new_with_cderi,
 * hess_r.rs:1775-1794) or 2 (UHF, UHessRIJK, hess_u.rs:38-146). */
hh_status hh_rijk_create(const double* cderi, int64_t nao, int64_t naux, double scale_j, double scale_k,
                         int32_t nspin, void* stream, hh_rijk** out);
hh_status hh_rijk_destroy(hh_rijk* h);

/* Store MO coefficients and occupations for the response (replaces make_response_preparation,
 * hess_r.rs:1910-1914 / hess_u.rs:263-268).  mo_coeff[s]: device [nao, nmo]; mo_occ[s]: HOST
 * [nmo]; occupied orbitals are those with occupation > 0, in their original order. */
hh_status hh_rijk_prepare_response(hh_rijk* h, const double* const* mo_coeff, const double* const* mo_occ,
                                   const int64_t* nmo);

/* Number of occupied orbitals of spin s after hh_rijk_prepare_response. */
hh_status hh_rijk_nocc(const hh_rijk* h, int32_t s, int64_t* nocc);

/* resp[s] [nao, nocc_s, nprop] = response of bra[s] [nao, nocc_s, nprop]
 * (replaces RHessRIJK::get_response_bra, hess_r.rs:1916-1953, and UHessRIJK::get_response_bra,
 * hess_u.rs:270-313).  `precision` selects the K/J contraction arithmetic. */
hh_status hh_rijk_get_response_bra(hh_rijk* h, const double* const* bra, int64_t nprop, double* const* resp,
                                   hh_precision precision);

/* Separated J/K core (replaces get_rijk_response_bra_separated, hess_r.rs:82-186).
 * j_ao [nao, nao, nprop] (may be NULL to skip J); k_bra[s] [nao, nocc_s, nprop] (k_bra may be
 * NULL to skip K).  The handle's scale_j / scale_k are NOT applied here (as in the reference). */
hh_status hh_rijk_response_bra_separated(hh_rijk* h, const double* const* bra, int64_t nprop, double* j_ao,
                                         double* const* k_bra, hh_precision precision);

/* Aux-index chunk of the K contraction (the reference's nbatch_aux, hard-coded 72 at
 * hess_r.rs:1941 "TODO tunable"); 0 = size it from the intermediate's memory. */
hh_status hh_rijk_set_aux_chunk(hh_rijk* h, int64_t naux_chunk);

/* Stage timing with HIP events on the handle's stream (the reference keeps per-stage wall
 * times in `timing`, hess_r.rs:1746).  When on, each response call waits for its own events.
 * out[0] hot K-contraction kernel ms (summed), out[1] K total ms, out[2] J total ms,
 * out[3] whole response ms, out[4] number of hot-kernel launches. */
hh_status hh_rijk_set_profiling(hh_rijk* h, int32_t on);
hh_status hh_rijk_get_timing(const hh_rijk* h, double* out5);
hh_status hh_rijk_reset_timing(hh_rijk* h);

/* ---------------------------------------------------------------------------------------- */
/* Utility: the library's strided batched GEMM (no reference counterpart; exposed so callers  */
/* and tests can drive the matrix-core kernel directly).                                      */
/* C[m,n] = alpha sum_k A[m,k] B[k,n] + beta C[m,n]; A[m,k] at a_kmajor ? m*lda+k : m+k*lda;   */
/* B[k,n] at b_kmajor ? k+n*ldb : k*ldb+n; C[m,n] at m+n*ldc; batch b0 (fastest), b1.         */
hh_status hh_util_gemm(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                       int32_t a_kmajor, const double* B, int64_t ldb, int32_t b_kmajor, double beta, double* C,
                       int64_t ldc, int64_t nb0, int64_t nb1, const int64_t* strides6, hh_precision precision,
                       void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HESSHIP_H */
