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
/* Decomposition policy of the RI metric (J2CDecompPolicy, src/ri_jk/decompose.rs:8-16). */
#define HH_J2C_CD 0               /* Cholesky j2c = L L^T, cderi = j3c L^-T */
#define HH_J2C_EIG 1              /* eigen: cderi = j3c J^-1/2 over eigenvalues >= threshold */

/* The same from the 3-centre integrals: j3c [nao_tp, naux] (packed upper, column-major) and the
 * metric j2c [naux, naux] (device), decomposed on the device in FP64 (generate_cderi_with_decomp,
 * hess_r.rs:1725-1735, with get_j2c_decomp, pure_decompose.rs:61-73; threshold > 0 is
 * Some(threshold), threshold <= 0 is None):
 *   HH_J2C_CD  (decomp_j2c_cd, pure_decompose.rs:28-55): j2c = L L^T, cderi = j3c L^-T.  If the
 *              Cholesky factorisation fails or a diagonal element of L is <= threshold, j2c is
 *              rebuilt as V max(e, threshold) V^T from its eigendecomposition and factorised again;
 *              with no threshold a failure gives HH_ERR_INVALID (the reference panics).
 *   HH_J2C_EIG (decomp_j2c_eig, pure_decompose.rs:11-20): J^-1/2 = V_k e_k^-1/2 V_k^T over the
 *              eigenvalues e_k >= threshold (no threshold: J2C_THRESH = 1e-13, decompose.rs:5),
 *              cderi = j3c J^-1/2 (solve_by_j2c with J2CDecompose::Eig, side Right).
 * The handle keeps the factor (L, or J^-1/2) for the skeleton's metric terms. */
hh_status hh_rijk_create_from_j3c_decomp(const double* j3c, const double* j2c, int64_t nao, int64_t naux,
                                         int32_t policy, double threshold, double scale_j, double scale_k,
                                         int32_t nspin, void* stream, hh_rijk** out);
/* hh_rijk_create_from_j3c_decomp with HH_J2C_CD (RHessRIJK::new_without_cderi uses
 * J2CDecompOption {Cd, Some(1e-14), Upper}, hess_r.rs:1751-1753). */
hh_status hh_rijk_create_from_j3c(const double* j3c, const double* j2c, int64_t nao, int64_t naux, double threshold,
                                  double scale_j, double scale_k, int32_t nspin, void* stream, hh_rijk** out);
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

/* Perturbations per K-contraction batch (the GPU counterpart of the reference's memory knob
 * nbatch_aux, hard-coded 72 at hess_r.rs:1941 "TODO tunable": here the whole aux range is one
 * contraction and the intermediate is bounded by batching perturbations); 0 = size it from
 * memory. */
hh_status hh_rijk_set_batch(hh_rijk* h, int64_t nprop_batch);

/* Stage timing with HIP events on the handle's stream (the reference keeps per-stage wall
 * times in `timing`, hess_r.rs:1746).  When on, each response call waits for its own events.
 * out[0] hot K-contraction kernel ms (summed), out[1] K total ms, out[2] J total ms,
 * out[3] whole response ms, out[4] number of hot-kernel launches. */
/* BF16x3 RI-K algorithm: 0 (default) the fused contraction with its operand generated in place
 * (8 <= nocc <= 80, nao <= 512), 2 the materialised operand; the same algebra as the reference's
 * K loop (hess_r.rs:144-183).  FP64 / FP32 are unaffected. */
hh_status hh_rijk_set_k_path(hh_rijk* h, int32_t path);
/* Response algorithm of get_response_bra: 0 (default) per call, the reference's algebra
 * (hess_r.rs:82-186); 1 resident (RHF and UHF, FP64 and BF16x3): the response matrix
 * M[(mu,i),(nu,j)] = 4 scale_j sum_P X_P[mu,i] X_P[nu,j] - scale_k sum_P (B_P[mu,nu] O_P[j,i] +
 * X_P[mu,j] X_P[nu,i]) (X_P = B_P C_o, O_P = C_o^T X_P; UHF: rows and columns over both spins,
 * J weight 2 scale_j between every pair of spins, K within each spin) is built on the first call
 * after prepare_response and each call is resp = M bra.  Falls back to the per-call algorithm
 * when M would need more than budget_bytes (<= 0: keep the current budget, 24 GiB by default)
 * or for FP32. */
hh_status hh_rijk_set_response_path(hh_rijk* h, int32_t path, int64_t budget_bytes);
hh_status hh_rijk_set_profiling(hh_rijk* h, int32_t on);
hh_status hh_rijk_get_timing(const hh_rijk* h, double* out5);
hh_status hh_rijk_reset_timing(hh_rijk* h);

/* Skeleton (fixed-orbital) second derivative of the RI-J/K energy and the first-derivative Fock
 * terms (get_rijk_skeleton_decomposed_separated, hess_r.rs:202-1718, as consumed by
 * RHessRIJK::make_skeleton_hess / get_deriv1_bra, hess_r.rs:1858-1909, and UHessRIJK,
 * hess_u.rs:150-261).  The 3-centre derivative integrals come from a generator called per
 * auxiliary range (generator_hess_intor_j3c_by_aux, cint_handling.rs:80-92): it fills `out`
 * (device, column-major) with int3c2e_<kind> over all AO pairs and the auxiliary functions
 * [p0, p1): [nao, nao, p1-p0, 3] for IP1 / IP2 and [nao, nao, p1-p0, 3, 3] for the second
 * derivatives (libcint's 9 components reshaped as hess_intor does, cint_handling.rs:20-44), and
 * must enqueue its work on `stream` (or finish before returning).  Non-zero return aborts. */
#define HH_INT3C_IP1 0
#define HH_INT3C_IP2 1
#define HH_INT3C_IPIP1 2
#define HH_INT3C_IPVIP1 3
#define HH_INT3C_IP1IP2 4
#define HH_INT3C_IPIP2 5
typedef int32_t (*hh_int3c_deriv_fn)(void* ctx, int32_t kind, int64_t p0, int64_t p1, double* out, void* stream);
/* mo_coeff[s] device [nao, nmo[s]], mo_occ[s] HOST [nmo[s]] (one entry per handle spin);
 * j2c device [naux, naux] (the metric the handle's cderi was decomposed with by Cholesky; NULL
 * for a handle made by hh_rijk_create_from_j3c, which keeps its factor); j2c_ip1 [naux, naux, 3],
 * j2c_ipip1 and j2c_ip1ip2 [naux, naux, 3, 3] (int2c2e_ip1 / ipip1 / ip1ip2, device);
 * aux_ranges HOST [nrange, 2] covering [0, naux) in order (the reference's aux_ranges, shell
 * aligned); aoslice / auxslice HOST [natm, 2] AO and auxiliary ranges of the atoms in the list.
 * Outputs (device): de [3, 3, natm, natm] = make_skeleton_hess (scale_j de_J - 0.5 scale_k de_K
 * RHF, scale_j de_J - scale_k sum_s de_K UHF); deriv1_bra[s] [nao, nocc_s, 3, natm] =
 * get_deriv1_bra (NULL to skip).  Synchronous. */
hh_status hh_rijk_skeleton(hh_rijk* h, const double* const* mo_coeff, const int64_t* nmo, const double* const* mo_occ,
                           const double* j2c, const double* j2c_ip1, const double* j2c_ipip1, const double* j2c_ip1ip2,
                           hh_int3c_deriv_fn int3c, void* ctx, int64_t nrange, const int64_t* aux_ranges, int64_t natm,
                           const int64_t* aoslice, const int64_t* auxslice, double* de, double* const* deriv1_bra);

/* ---------------------------------------------------------------------------------------- */
/* Gaussian-basis integrals and AO values (the libcint calls of the reference's Hessian path)  */

typedef struct hh_cint hh_cint;
/* Integral kinds: int2c2e / int3c2e (hess_r.rs:1731, pure_decompose.rs:139) and the derivative
 * integrals of the skeleton, libcint's ip = the gradient on the electron coordinate of one
 * function: ip1 (d mu nu|P), ip2 (mu nu|d P), ipip1 (dd mu nu|P), ipvip1 (d mu d nu|P), ip1ip2
 * (d mu nu|d P), ipip2 (mu nu|dd P) (hess_r.rs:537-540, 1058, 1356); 2-centre: PLAIN, IP1
 * (dP|Q), IPIP1 (ddP|Q), IP1IP2 (dP|dQ) (hess_r.rs:416, 788-789). */
#define HH_INT_PLAIN 0
#define HH_INT_IP1 1
#define HH_INT_IP2 2
#define HH_INT_IPIP1 3
#define HH_INT_IPVIP1 4
#define HH_INT_IP1IP2 5
#define HH_INT_IPIP2 6
/* one-electron operators of hh_cint_int1e */
#define HH_INT1E_OVLP 0
#define HH_INT1E_KIN 1
#define HH_INT1E_NUC 2
#define HH_INT1E_RINV 3
/* A basis in libcint's format (what the reference's CInt holds): atm [natm][6] (charge,
 * ptr_coord, ...), bas [nbas][8] (atom, l, nprim, nctr, kappa, ptr_exp, ptr_coeff, -), env
 * [nenv] (coordinates in bohr; contraction coefficients [nprim, nctr] used as given, i.e. with
 * the radial normalisation folded in as libcint expects).  l <= 5.  cart: 0 spherical (real solid
 * harmonics, libcint's order), 1 Cartesian.  Host arrays, copied. */
hh_status hh_cint_create(const int32_t* atm, int32_t natm, const int32_t* bas, int32_t nbas, const double* env,
                         int64_t nenv, int32_t cart, hh_cint** out);
hh_status hh_cint_destroy(hh_cint* h);
/* number of AO functions; ao_loc (HOST, may be NULL) receives nbas + 1 AO offsets of the shells */
hh_status hh_cint_nao(const hh_cint* h, int64_t* nao, int64_t* ao_loc);
/* eval_gto (CInt::eval_gto "deriv0".."deriv3" as NIMatmul::prepare_ao calls it, nimatmul.rs:94-99):
 * out (device) [ng, nao, ncomp] column-major, ncomp = 1, 4, 10, 20 (value, then x, y, z, then
 * xx, xy, xz, yy, yz, zz, then the ten third derivatives in the same order); coords (device)
 * [ng][3]. */
hh_status hh_cint_eval_gto(const hh_cint* h, int32_t deriv, const double* coords, int64_t ng, double* out,
                           void* stream);
/* int2c2e of one basis (hess_intor(aux, "int2c2e[_ip1|_ipip1|_ip1ip2]", "s1"), cint_handling.rs:20-44):
 * out (device) [naux, naux, comp], comp 1 / 3 / 9 / 9 (9 as [3, 3], the later function's
 * direction fastest).  Synchronous. */
hh_status hh_cint_int2c2e(const hh_cint* aux, int32_t kind, double* out, void* stream);
/* int3c2e over every AO pair of `mol` and the auxiliary shells [sh0, sh1) of `aux`
 * (generator_hess_intor_j3c_by_aux, cint_handling.rs:80-92): out (device) [nao, nao, n01, comp].
 * Synchronous. */
hh_status hh_cint_int3c2e(const hh_cint* mol, const hh_cint* aux, int32_t kind, int32_t sh0, int32_t sh1,
                          double* out, void* stream);
/* int3c2e with "s2ij" symmetry (hess_intor_cross(&[mol, mol, aux], "int3c2e", "s2ij"),
 * hess_r.rs:1731): out (device) [nao_tp, n01], the upper triangle mu <= nu packed column-major
 * (index nu (nu + 1) / 2 + mu), the j3c that hh_rijk_create_from_j3c takes.  Synchronous. */
/* One-electron integrals out [nao, nao, comp] (column-major), the libcint int1e family the
 * Hessian's core and overlap parts take (hcore.rs:28-34, 47-82, 155-177; ovlp.rs:31-32, 76):
 *   op HH_INT1E_OVLP (int1e_ovlp), HH_INT1E_KIN (int1e_kin), HH_INT1E_NUC (int1e_nuc:
 *   sum_C -Z_C / |r - R_C|), HH_INT1E_RINV (int1e_rinv with the origin at nucleus `atom`, no
 *   charge: with_rinv_at_nucleus);
 *   kind HH_INT_PLAIN (comp 1), HH_INT_IP1 (comp 3: ip on the first function, e.g. int1e_ipovlp,
 *   int1e_ipkin, int1e_ipnuc, int1e_iprinv), HH_INT_IPIP1 (comp 9: ipip on the first, int1e_ipip*),
 *   HH_INT_IPVIP1 (comp 9 = 3 t + s, t on the first function and s on the second: int1e_ipovlpip,
 *   int1e_ipkinip, int1e_ipnucip, int1e_iprinvip).  Nuclear charges are atm[.][0] of the
 *   molecule's libcint arrays.  Synchronous. */
hh_status hh_cint_int1e(const hh_cint* mol, int32_t op, int32_t kind, int32_t atom, double* out, void* stream);

/* Restricted Hartree-Fock with RI-JK (density-fitted J and K over the auxiliary basis `aux`, the
 * metric of hh_rijk_create_from_j3c_decomp with HH_J2C_CD and threshold 1e-14) on the device:
 * the mo_coeff / mo_occ / mo_energy the reference's RHessSCF takes (rscf.rs:30-95; its tests read
 * them from PySCF, tests/hessian_rhf/mod.rs:39-41).  nocc doubly occupied orbitals; converged when
 * |dE| < conv_tol and the orthogonal-basis commutator max|X^T (FDS - SDF) X| < sqrt(conv_tol)
 * (core-Hamiltonian guess, DIIS over 8 Fock matrices).  Outputs on the device: mo_coeff [nao, nao],
 * mo_energy [nao] ascending; host: *energy (total, with nuclear repulsion), *cycles (either may be
 * NULL).  HH_ERR_NOT_CONVERGED after max_cycle cycles; HH_ERR_INVALID for a linearly dependent
 * basis.  Synchronous. */
hh_status hh_scf_rhf(const hh_cint* mol, const hh_cint* aux, int32_t nocc, double conv_tol, int32_t max_cycle,
                     double* mo_coeff, double* mo_energy, double* energy, int32_t* cycles, void* stream);

/* One-electron terms of the analytic Hessian (src/hessian/nuc_repl.rs, hcore.rs, ovlp.rs) on the
 * device, all atoms.  Hessians out [3, 3, natm, natm] (column-major; block (., ., B, A) as the
 * reference fills it); first derivatives out [nao, nao, 3, natm].  The basis shells must be
 * grouped by atom (aoslice_by_atom).  Synchronous. */
/* get_nuc_repl_hess (nuc_repl.rs:14-58) */
hh_status hh_hess_nuc_repl(const hh_cint* mol, double* out, void* stream);
/* get_hess_hcore (hcore.rs:105-125) with generator_hcore_deriv2 (hcore.rs:15-92): dm0 [nao, nao] */
hh_status hh_hess_hcore(const hh_cint* mol, const double* dm0, double* out, void* stream);
/* get_hess_ovlp (ovlp.rs:24-52): dme0 [nao, nao] the energy-weighted density */
hh_status hh_hess_ovlp(const hh_cint* mol, const double* dme0, double* out, void* stream);
/* generator_hcore_deriv1 (hcore.rs:140-161) for every atom: out[:, :, :, A] = its [nao, nao, 3] */
hh_status hh_hess_hcore_deriv1(const hh_cint* mol, double* out, void* stream);
/* generator_ovlp_deriv1 (ovlp.rs:67-82) for every atom */
hh_status hh_hess_ovlp_deriv1(const hh_cint* mol, double* out, void* stream);
/* out [nao, nao] = sum_i w[i] C[:, i] C[:, i]^T, C [nao, nmo] on the device, w [nmo] HOST:
 * get_dm0_restricted (w = mo_occ) and get_dme0_restricted (w = mo_occ * mo_energy). */
hh_status hh_hess_weighted_density(const double* mo_coeff, int64_t nao, int64_t nmo, const double* w, double* out,
                                   void* stream);
hh_status hh_cint_int3c2e_s2ij(const hh_cint* mol, const hh_cint* aux, int32_t sh0, int32_t sh1, double* out,
                               void* stream);
/* An hh_int3c_deriv_fn for hh_rijk_skeleton over the engine: ctx = an hh_cint_pair made from the
 * AO and auxiliary bases; the auxiliary ranges it is called with must be shell aligned. */
typedef struct hh_cint_pair hh_cint_pair;
hh_status hh_cint_pair_create(const hh_cint* mol, const hh_cint* aux, hh_cint_pair** out);
hh_status hh_cint_pair_destroy(hh_cint_pair* p);
int32_t hh_cint_int3c_deriv(void* ctx, int32_t kind, int64_t p0, int64_t p1, double* out, void* stream);

/* ---------------------------------------------------------------------------------------- */
/* Numerical-integration grid contractions (src/numint_matmul)                               */

typedef struct hh_numint hh_numint;
/* XCDenType (src/xceff/flags.rs:19-24) */
#define HH_DEN_RHO 0
#define HH_DEN_SIGMA 1
#define HH_DEN_TAU 2

/* Grid handle from device AO values ao [ng, nao, ncomp] (column-major; component 0 = value,
 * 1..3 = gradient, as NIMatmul's cached AO, nimatmul.rs:105-126) and weights [ng]; both are
 * copied.  (The reference's NIMatmul::new evaluates AOs itself with libcint, nimatmul.rs:41-99;
 * AO evaluation is outside this library.) */
hh_status hh_numint_create(const double* ao, int64_t ng, int64_t nao, int64_t ncomp, const double* weights,
                           void* stream, hh_numint** out);
hh_status hh_numint_destroy(hh_numint* h);
/* Grid points per batch (the reference's nbatch, nimatmul.rs:18-36); 0 = size from memory. */
hh_status hh_numint_set_batch(hh_numint* h, int64_t nbatch);
/* BF16X3 RKS response: compute the kets product inside the pointwise pass (default 1; no
 * reference counterpart, results bit-identical either way; nocc <= 128, else unfused). */
hh_status hh_numint_set_fused(hh_numint* h, int32_t on);
/* rho1 [ng, nvar, nprop] from bra [nao, nocc] and kets [nao, nocc, nprop]
 * (make_rho_from_one_bra_mult_ket, nimatmul.rs:196-229 / pure_eval_rho.rs:216-334). */
hh_status hh_numint_rho1_one_bra_mult_ket(hh_numint* h, int32_t den_type, const double* bra, int64_t nocc,
                                          const double* kets, int64_t nprop, double* rho1, hh_precision precision);
/* out [nao, nocc, nprop] = make_rks_fxc_pot_with_eff_bra_trans(fxc_eff [ng, nvar, nvar], rho1, bra)
 * (nimatmul.rs:361-389 / pure_xcpot.rs:370-465). */
hh_status hh_numint_fxc_pot_bra_trans(hh_numint* h, int32_t den_type, const double* fxc_eff, const double* rho1,
                                      int64_t nprop, const double* bra, int64_t nocc, double* out,
                                      hh_precision precision);
/* out [nao, nocc, nprop] = get_rks_response_bra (hess_rks.rs:570-605, x4 closed-shell factor),
 * batched over the grid (get_rks_response_bra_batched, hess_rks.rs:752-797). */
hh_status hh_numint_rks_response_bra(hh_numint* h, int32_t den_type, const double* fxc_eff, const double* mo1_bra,
                                     int64_t nprop, const double* mocc, int64_t nocc, double* out,
                                     hh_precision precision);

/* UKS: out[s] [nao, nocc[s], nprop] = uks_fxc_pot_with_eff_bra_trans(fxc_eff, rho1, bra)[s]
 * (pure_xcpot.rs:921-1028): fxc_eff [ng, nvar, 2, nvar, 2], rho1 [ng, nvar, 2, nprop] (the
 * two spins' densities stacked on axis 2), bra[s] [nao, nocc[s]]; two arrays per argument. */
hh_status hh_numint_uks_fxc_pot_bra_trans(hh_numint* h, int32_t den_type, const double* fxc_eff, const double* rho1,
                                          int64_t nprop, const double* const* bra, const int64_t* nocc,
                                          double* const* out, hh_precision precision);
/* out[s] [nao, nocc[s], nprop] = get_uks_response_bra (hess_uks.rs:390-441; rho1 of both spins
 * from mocc[s] and mo1_bra[s] [nao, nocc[s], nprop], x2 factor), batched over the grid
 * (get_uks_response_bra_batched, hess_uks.rs:589-650). */
hh_status hh_numint_uks_response_bra(hh_numint* h, int32_t den_type, const double* fxc_eff,
                                     const double* const* mo1_bra, int64_t nprop, const double* const* mocc,
                                     const int64_t* nocc, double* const* out, hh_precision precision);

/* DFT skeleton-Hessian setup over a fixed grid: nspin 1 = RKS (make_hessian_setup,
 * hess_rks.rs:71-172, dm0 the closed-shell density), nspin 2 = UKS (make_hessian_setup_uks,
 * hess_uks.rs:286-384, one dm0 [nao, nao] per spin), batched over the grid as
 * make_hessian_setup_batched (hess_rks.rs:611-750).  The handle's AO tensor must carry
 * derivatives to order 2 (RHO: ncomp >= 10) or 3 (SIGMA/TAU: ncomp >= 20), components ordered
 * value, x, y, z, xx, xy, xz, yy, yz, zz, xxx, xxy, xxz, xyy, xyz, xzz, yyy, yyz, yzz, zzz.
 * hh_numint_setup_rho: rho [ng, nvar, nspin] (the density of get_rho_vxc_fxc, hess_rks.rs:174-218 /
 * get_rho_vxc_fxc_uks, hess_uks.rs:56-108) for the caller's functional evaluation (LibXC is
 * outside this library). */
hh_status hh_numint_setup_rho(hh_numint* h, int32_t den_type, int32_t nspin, const double* const* dm0,
                              double* rho);
/* The additive setup intermediates from the transformed kernels vxc_eff [ng, nvar, nspin] and
 * fxc_eff [ng, nvar, nspin, nvar, nspin] (weights not applied; hh_xceff_transform /
 * hh_xceff_transform_polarized): de_fxc [3, 3, natm, natm] and per spin de_vxc_diag,
 * de_vxc_off [3, 3, natm, natm], vmat_ip [nao, nao, 3], vmat_deriv1 [nao, nao, 3, natm] (get_de_fxc,
 * get_de_vxc_diag, get_de_vxc_off, get_vmat_ip, get_vmat_deriv1: hess_rks.rs:220-565; UKS
 * hess_uks.rs:110-284).  aoslice: HOST [natm, 2] AO ranges (p0, p1) of the atoms in the list
 * (aoslice_by_atom()[atm_list], columns 2-3).  The skeleton Hessian is
 * de_fxc + sum_s (de_vxc_diag[s] + de_vxc_off[s]) (make_skeleton_hess, hess_rks.rs:866-872,
 * hess_uks.rs:705-720); get_deriv1_ao is vmat_deriv1. */
hh_status hh_numint_hessian_setup(hh_numint* h, int32_t den_type, int32_t nspin, const double* const* dm0,
                                  int64_t natm, const int64_t* aoslice, const double* vxc_eff, const double* fxc_eff,
                                  double* de_fxc, double* const* de_vxc_diag, double* const* de_vxc_off,
                                  double* const* vmat_ip, double* const* vmat_deriv1);

/* ---------------------------------------------------------------------------------------- */
/* Becke partition and the grid-shift setup (src/numint_matmul/becke_partition.rs,           */
/* hess_rks_becke.rs, hess_uks_becke.rs)                                                     */

typedef struct hh_becke hh_becke;
/* Molecular tables (BeckeMolTables::new, becke_partition.rs:163-205): atm_coords HOST [natm, 3];
 * adjustment HOST [natm, natm] row-major, entry (A, B) = adjustment_factor[A][B] (antisymmetric;
 * Becke 1988 eqs. A2, A5).  Work is enqueued on `stream`. */
hh_status hh_becke_create(int64_t natm, const double* atm_coords, const double* adjustment, void* stream,
                          hh_becke** out);
hh_status hh_becke_destroy(hh_becke* b);
/* becke_partition_with_tables (becke_partition.rs:259-340).  grid_coords DEVICE [ngrids, 3];
 * quadrature_weights DEVICE [ngrids].  by_atom != 0: atm_indices HOST [natm + 1] cumulative grid
 * boundaries (AtmIndices::ByAtom; must start at 0, end at ngrids, be non-decreasing); else HOST
 * [ngrids] generating atoms (ByGrid; values >= natm are padding and carry no weight).  deriv 0..2.
 * Outputs (DEVICE, NULL = not produced; derivative outputs above `deriv` are ignored):
 * w [ngrids]; dw [natm, 3, ngrids]; ddw [natm, 3, natm, 3, ngrids] (C order); contractions with
 * contract_x [nset, ngrids]: c [nset_w], dc [natm, 3, nset_dw], ddc [natm, 3, natm, 3, nset_ddw].
 * Second derivatives support at most 48 atoms (per-point atom-pair tables in LDS). */
hh_status hh_becke_partition(hh_becke* b, const double* grid_coords, int64_t ngrids, int32_t by_atom,
                             const int64_t* atm_indices, const double* quadrature_weights, int32_t hardness,
                             int32_t deriv, double* w, double* dw, double* ddw, const double* contract_w,
                             int64_t nset_w, double* c, const double* contract_dw, int64_t nset_dw, double* dc,
                             const double* contract_ddw, int64_t nset_ddw, double* ddc);
/* The setup with the Becke grid shift (make_hessian_setup_becke, hess_rks_becke.rs:1106-1266;
 * make_hessian_setup_becke_uks, hess_uks_becke.rs:726-930), as consumed by
 * RHessKSNIMatmulBecke::make_skeleton_hess / get_deriv1_ao (hess_rks_becke.rs:1389-1401): the
 * handle's grid is grouped by generating atom (atm_quad_split HOST [natm + 1]) and its weights are
 * the partitioned ones; grid_coords DEVICE [ng, 3]; quadrature_weights DEVICE [ng] (before the
 * partition); exc DEVICE [ng] the energy per particle (energy density exc * sum_s rho_s[0]);
 * vxc_eff / fxc_eff as hh_numint_hessian_setup.  The full atom list is required (aoslice covers
 * every AO; the Becke tables hold the same natm atoms and use the handle's stream).
 * Outputs: de_xc_skeleton [3, 3, natm, natm] and per spin vmat_deriv1_grid [nao, nao, 3, natm]
 * (both translation invariant); optional grid-fixed keys (NULL array or entry = skipped):
 * de_fxc, de_vxc_diag, de_vxc_off, vmat_ip, vmat_deriv1 as hh_numint_hessian_setup. */
hh_status hh_numint_hessian_setup_becke(hh_numint* h, int32_t den_type, int32_t nspin, const double* const* dm0,
                                        int64_t natm, const int64_t* aoslice, const double* exc, const double* vxc_eff,
                                        const double* fxc_eff, hh_becke* becke, const double* grid_coords,
                                        const double* quadrature_weights, const int64_t* atm_quad_split,
                                        int32_t hardness, double* de_xc_skeleton, double* const* vmat_deriv1_grid,
                                        double* de_fxc, double* const* de_vxc_diag, double* const* de_vxc_off,
                                        double* const* vmat_ip, double* const* vmat_deriv1);

/* XC effective kernels from LibXC-variable derivatives, unpolarised (transform_xc_inner,
 * xc_deriv.rs:305-428): rho [ng, >= nvar] (ld_rho), xc_val [ng, ncol] (ld_xc) with columns
 * [exc | 1st derivatives | unique 2nd derivatives]; order 1 -> out [ng, nvar] (vxc_eff),
 * order 2 -> out [ng, nvar, nvar] (fxc_eff).  RHO: out = the order's raw column. */
hh_status hh_xceff_transform(int32_t den_type, int32_t order, const double* rho, int64_t ld_rho, const double* xc_val,
                             int64_t ld_xc, int64_t ng, double* out, void* stream);
/* Spin-polarised transform (transform_xc_inner with LibXCSpin::Polarized: vxc_unfold_sigma_spin1,
 * xc_deriv.rs:188-243, and the order-2 sigma pair term, xc_deriv.rs:379-424): rho [ng, nvar, 2]
 * with element (g, x, s) at g + ld_rho*x + ld_spin*s; xc_val [ng, ncol] (ld_xc) in XCFun order
 * (rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb, tau_a, tau_b; second derivatives as combinations
 * with replacement); order 1 -> out [ng, nvar, 2], order 2 -> out [ng, nvar, 2, nvar, 2]. */
hh_status hh_xceff_transform_polarized(int32_t den_type, int32_t order, const double* rho, int64_t ld_rho,
                                       int64_t ld_spin, const double* xc_val, int64_t ld_xc, int64_t ng, double* out,
                                       void* stream);
/* Columns of LibXC's polarised output [zk | 1st | 2nd derivatives] listed in XCFun order
 * (libxc_to_xcfun_mapping, xc_deriv.rs:17-72), deriv 0..2.  Host.  *len = map length; map may be
 * NULL to query it, else it needs capacity >= *len. */
hh_status hh_xceff_libxc_to_xcfun_map(int32_t den_type, int32_t deriv, int64_t* map, int64_t capacity,
                                      int64_t* len);

/* Restricted Kohn-Sham with RI-JK: hh_scf_rhf's SCF with exact exchange scaled by exx (the
 * hybrid fraction, hyb_exx_coef) and the functional list (xc_ids / xc_scales HOST arrays of
 * nfunc LibXC ids and scales, as get_rho_vxc_fxc's list, hess_rks.rs:193-216) evaluated on the
 * grid of `grid` (an hh_numint whose AO values carry first derivatives for SIGMA / TAU; it must
 * run on `stream`); den_type the list's strictest family.  E = E_core + E_J - exx E_K + E_xc. */
hh_status hh_scf_rks(const hh_cint* mol, const hh_cint* aux, hh_numint* grid, int32_t den_type, int32_t nfunc,
                     const int32_t* xc_ids, const double* xc_scales, double exx, int32_t nocc, double conv_tol,
                     int32_t max_cycle, double* mo_coeff, double* mo_energy, double* energy, int32_t* cycles,
                     void* stream);
/* Unrestricted HF (grid NULL: xc arguments ignored) or KS: nalpha / nbeta singly occupied
 * orbitals; mo_coeff[2] [nao, nao] and mo_energy[2] [nao] device arrays per spin channel (the
 * mo_coeff / mo_occ / mo_energy pairs of UHessSCF, uscf.rs:32-77); otherwise as hh_scf_rks,
 * with the functional evaluated spin-polarised. */
hh_status hh_scf_uhf(const hh_cint* mol, const hh_cint* aux, hh_numint* grid, int32_t den_type, int32_t nfunc,
                     const int32_t* xc_ids, const double* xc_scales, double exx, int32_t nalpha, int32_t nbeta,
                     double conv_tol, int32_t max_cycle, double* const* mo_coeff, double* const* mo_energy,
                     double* energy, int32_t* cycles, void* stream);
/* The exchange-correlation potential matrix out [nao, nao] = dE_xc / dD (the Kohn-Sham Fock
 * term; no reference counterpart on the Hessian path, which takes converged orbitals) from
 * vxc_eff [ng, nvar] of hh_xc_eval_eff (weights not applied).  Synchronous. */
hh_status hh_numint_vxc_matrix(hh_numint* h, int32_t den_type, const double* vxc_eff, double* out);

/* Exchange-correlation functionals evaluated on the device (the values LibXC's compute_xc gives
 * libxc_eval_inner, src/xceff/libxc_wrap.rs:32-82; functional families as determine_den_type,
 * libxc_wrap.rs:10-23).  Functional ids are LibXC's: HH_XC_LDA_X 1 (Slater), HH_XC_LDA_C_VWN 7
 * (VWN5), HH_XC_LDA_C_VWN_RPA 8, HH_XC_GGA_X_B88 106, HH_XC_GGA_C_LYP 131, HH_XC_HYB_GGA_XC_B3LYP
 * 402 (0.08 Slater + 0.72 B88 + 0.19 VWN_RPA + 0.81 LYP; the 0.2 of exact exchange is the
 * caller's K scale), HH_XC_MGGA_X_TPSS 202, HH_XC_MGGA_C_TPSS 231, HH_XC_HYB_MGGA_XC_TPSS0 396
 * (0.75 TPSS exchange + TPSS correlation; 0.25 exact exchange).  Derivative order 0..2. */
#define HH_XC_LDA_X 1
#define HH_XC_LDA_C_VWN 7
#define HH_XC_LDA_C_VWN_RPA 8
#define HH_XC_GGA_X_B88 106
#define HH_XC_GGA_C_LYP 131
#define HH_XC_HYB_GGA_XC_B3LYP 402
#define HH_XC_MGGA_X_TPSS 202
#define HH_XC_MGGA_C_TPSS 231
#define HH_XC_HYB_MGGA_XC_TPSS0 396
/* Host.  *den_type (HH_DEN_RHO / HH_DEN_SIGMA / HH_DEN_TAU, determine_den_type), *exx the fraction of exact
 * exchange (xc_func.hyb_exx_coef()), *ncol the raw column count of hh_xc_eval at (nspin, deriv).
 * Any output may be NULL.  HH_ERR_INVALID for an id this library does not evaluate. */
hh_status hh_xc_info(int32_t xc_id, int32_t nspin, int32_t deriv, int32_t* den_type, double* exx, int64_t* ncol);
/* libxc_eval_inner: out[g + ld_out*c] = column c of LibXC's output at point g, columns
 * [zk | vrho | vsigma | vtau | v2rho2 | v2rhosigma | v2rhotau | v2sigma2 | v2sigmatau | v2tau2] up to
 * `deriv`, the blocks the functional's family has (LibXC's order: unpolarised (rho, sigma, tau);
 * polarised vrho (a, b), vsigma (aa, ab, bb), vtau (a, b), v2rho2 (aa, ab, bb), v2rhosigma
 * (a x {aa, ab, bb}, b x ...), v2rhotau (a x {a, b}, b x ...), v2sigma2 (upper triangle,
 * row-major), v2sigmatau (aa x {a, b}, ab x ..., bb x ...), v2tau2 (aa, ab, bb); no Laplacian
 * columns), as the reference holds them after transposing LibXC's point-major polarised blocks
 * (libxc_eval_eff_serial, libxc_wrap.rs:91-120).  zk is the energy per particle.  rho
 * [ng, nvar, nspin] with element (g, x, s) at g + ld_rho*x + ld_spin*s (x = 0 the density, 1..3
 * its gradient, 4 tau; sigma is formed from the gradient).  Points with total density below
 * 1e-15 give zeros. */
hh_status hh_xc_eval(int32_t xc_id, int32_t nspin, int32_t deriv, const double* rho, int64_t ld_rho, int64_t ld_spin,
                     int64_t ng, double* out, int64_t ld_out, void* stream);
/* One term of get_rho_vxc_fxc's functional loop (hess_rks.rs:205-216, hess_rks_becke.rs:195-205,
 * hess_uks.rs:96-106): libxc_eval_eff of the functional at derivative order 2 (exc, vxc_eff,
 * fxc_eff: libxc_wrap.rs:88-122 with transform_xc_inner, xc_deriv.rs:305-428), scaled by `scale`
 * and added (accumulate != 0) or stored into the leading slice of the caller's arrays:
 *   exc [ng]; vxc_eff [ng, nvar, (2)]; fxc_eff [ng, nvar, (2), nvar, (2)], point index fastest,
 * nvar that of `den_type` (the strictest family of the caller's functional list).  The chain
 * rule to the density components is applied to the functional's derivatives in registers.  Any
 * output may be NULL.  rho as hh_xc_eval. */
hh_status hh_xc_eval_eff(int32_t xc_id, int32_t nspin, int32_t den_type, double scale, const double* rho,
                         int64_t ld_rho, int64_t ld_spin, int64_t ng, int32_t accumulate, double* exc, double* vxc_eff,
                         double* fxc_eff, void* stream);

/* ---------------------------------------------------------------------------------------- */
/* Block Krylov solver for (I + A) x = b (src/util/krylov_block.rs)                         */

typedef struct {
  int32_t cycles;       /* trial-block aop calls (the reference's total_cycles) */
  int32_t restarts;     /* hard restarts */
  int32_t aop_calls;    /* all aop calls, including residual evaluations */
  int32_t refinements;  /* FP64 refinement sweeps (CPHF mixed precision; 0 otherwise) */
  double max_innerprod; /* last max ||v||^2 of the new trial block */
  double residual;      /* final max column 2-norm of the FP64 residual (CPHF mixed precision) */
  int32_t fp64_fallback; /* 1 if a mixed-precision CPHF solve stagnated (a sweep gained less than
                            10x) and was replaced by the plain FP64 solve from scratch */
} hh_krylov_stats;

/* ax[n, nblock] = A x[n, nblock]; device pointers, column-major; must enqueue its work on the
 * stream passed to hh_krylov_block (or finish before returning).  Non-zero return aborts. */
typedef int32_t (*hh_aop_fn)(void* ctx, const double* x, int64_t n, int64_t nblock, double* ax);

/* x[n, nset] solving (I + A) x = b (replaces krylov_block, krylov_block.rs:39-197; same
 * convergence test, restart rule and defaults).  x0 may be NULL.  Returns HH_ERR_NOT_CONVERGED
 * where the reference panics (x then holds the last approximation).  Synchronous. */
hh_status hh_krylov_block(hh_aop_fn aop, void* ctx, const double* b, const double* x0, int64_t n, int64_t nset,
                          double tol, int32_t max_cycle, int32_t max_space, double lindep, double* x,
                          hh_krylov_stats* stats, void* stream);

/* Non-normalised Gram-Schmidt of the columns of V [n, m] in column order with the drop rule
 * ||v||^2 <= lindep (replaces orth_block, krylov_block.rs:225-253).  Kept columns go to R in
 * order (R holds m columns; those past the kept ones are left zero or unspecified), their squared
 * norms to norms_out (host, m entries), the count to *kept.  V is used as scratch.  Synchronous. */
hh_status hh_krylov_orth_block(double* V, int64_t n, int64_t m, double lindep, double* R, double* norms_out,
                               int64_t* kept, void* stream);

/* ---------------------------------------------------------------------------------------- */
/* Dimensionless CPHF driver, restricted (src/hessian/rscf.rs) and unrestricted              */
/* (src/hessian/uscf.rs) SCF                                                                  */

typedef struct hh_cphf hh_cphf;

/* mo_coeff: device [nao, nmo]; mo_occ, mo_energy: HOST [nmo] (occupied orbitals first).
 * Config as HessSCFConfig (rscf.rs:5-18): level_shift, cphf_tol, cphf_max_cycle,
 * cphf_max_space, cphf_lindep. */
hh_status hh_cphf_create(const double* mo_coeff, int64_t nao, int64_t nmo, const double* mo_occ,
                         const double* mo_energy, double level_shift, double tol, int32_t max_cycle,
                         int32_t max_space, double lindep, void* stream, hh_cphf** out);
hh_status hh_cphf_destroy(hh_cphf* h);
/* Per-sweep / per-cycle diagnostics on stderr (the reference prints per Krylov cycle,
 * krylov_block.rs:148-151). */
hh_status hh_cphf_set_verbose(hh_cphf* h, int32_t on);
/* Replay the response operator from HIP graphs (default on; no reference counterpart): within a
 * solve, an operator call with the same operands as an earlier one (the Krylov trial block of a
 * given width) is captured once and replayed as one graph launch.  Applies when the driver runs on
 * a created stream (the legacy default stream cannot be captured) and every component enqueues on
 * it; graphs are dropped whenever a library buffer is freed.
 * hh_cphf_graph_stats: graphs captured and graph launches so far. */
hh_status hh_cphf_set_graphs(hh_cphf* h, int32_t on);
hh_status hh_cphf_graph_stats(hh_cphf* h, int64_t* captures, int64_t* launches);
/* Number of finalize calls (hh_cphf_finalize / hh_ucphf_finalize) that took the response of mo1
 * from the last mixed-precision solve instead of evaluating it in FP64 again (no reference
 * counterpart; the reference's finalize, rscf.rs:328-353, always evaluates it).  The solve's
 * final residual holds that response: FP64, except that the responses of late, small corrections
 * may be taken in the reduced precision when their error bound is below 5 % of cphf_tol.  It is
 * reused when the mo1 passed to finalize is bit-identical to the solve's output and no component
 * was added since (components must not be re-prepared between the solve and finalize). */
hh_status hh_cphf_finalize_reuses(hh_cphf* h, int64_t* count);
/* Add an RI-JK component (nspin 1 for the restricted driver, 2 for the unrestricted one,
 * prepared with the same MOs) to the response list (the reference's el_list, rscf.rs:28,
 * uscf.rs:24).  The component must outlive the driver. */
hh_status hh_cphf_add_rijk(hh_cphf* h, hh_rijk* rijk);
/* Add the DFT grid response (RHessKSNIMatmul::get_response_bra, hess_rks.rs:886-909) with
 * fxc_eff [ng, nvar, nvar] (device; must outlive the driver) to the response list. */
hh_status hh_cphf_add_numint(hh_cphf* h, hh_numint* ni, int32_t den_type, const double* fxc_eff);
/* rhs [nmo, nocc, nprop] from f1mo, s1mo [nmo, nocc, nprop] (compute_dimless_cphf_rhs,
 * rscf.rs:170-178). */
hh_status hh_cphf_dimless_rhs(hh_cphf* h, const double* f1mo, const double* s1mo, int64_t nprop, double* rhs);
/* resp = C^T sum_el get_response_bra(C mo1) (response_mo, rscf.rs:210-220). */
hh_status hh_cphf_response_mo(hh_cphf* h, const double* mo1, int64_t nprop, double* resp, hh_precision precision);
/* response_dimless_cphf, rscf.rs:237-263. */
hh_status hh_cphf_response_dimless(hh_cphf* h, const double* mo1, int64_t nprop, double* resp,
                                   hh_precision precision);
/* solve_dimless_cphf, rscf.rs:278-300.  precision FP64: the reference algorithm.  FP32 /
 * BF16X3: Krylov on the reduced-precision response inside FP64 iterative refinement, stopping
 * when every column's FP64 residual 2-norm is below cphf_tol.  diag (device [nmo, nocc], may be
 * NULL): right diagonal preconditioner on the virtual rows (x = diag * y). */
hh_status hh_cphf_solve_dimless(hh_cphf* h, const double* rhs, int64_t nprop, double* mo1, hh_precision precision,
                                const double* diag, hh_krylov_stats* stats);
/* finalize_cphf, rscf.rs:328-353: mo1_out [nmo, nocc, nprop], mo_e1 [nocc, nocc, nprop]. */
hh_status hh_cphf_finalize(hh_cphf* h, const double* f1mo, const double* s1mo, const double* mo1, int64_t nprop,
                           double* mo1_out, double* mo_e1);
/* get_cphf_hess, rscf.rs:371-399: de [3, 3, natm, natm] from [.., .., 3, natm] inputs. */
hh_status hh_cphf_hess(hh_cphf* h, const double* f1mo, const double* s1mo, const double* mo1, const double* mo_e1,
                       int64_t natm, double* de);

/* Unrestricted driver: the same handle type, two spins.  mo_coeff[2] device [nao, nmo[s]];
 * mo_occ[2], mo_energy[2] HOST [nmo[s]] (occupied orbitals first in each spin).  Replaces
 * UHessSCF (uscf.rs:32-56).  The hh_cphf_* per-call entry points above are for the restricted
 * driver; the hh_ucphf_* ones below take one array per spin, [nmo[s], nocc[s], nprop] (mo_e1:
 * [nocc[s], nocc[s], nprop]), like the reference's [TsrView; 2] arguments. */
hh_status hh_ucphf_create(const double* const* mo_coeff, int64_t nao, const int64_t* nmo,
                          const double* const* mo_occ, const double* const* mo_energy, double level_shift, double tol,
                          int32_t max_cycle, int32_t max_space, double lindep, void* stream, hh_cphf** out);
/* compute_dimless_cphf_rhs, uscf.rs:165-174. */
hh_status hh_ucphf_dimless_rhs(hh_cphf* h, const double* const* f1mo, const double* const* s1mo, int64_t nprop,
                               double* const* rhs);
/* response_mo, uscf.rs:203-218. */
hh_status hh_ucphf_response_mo(hh_cphf* h, const double* const* mo1, int64_t nprop, double* const* resp,
                               hh_precision precision);
/* response_dimless_cphf, uscf.rs:229-265. */
hh_status hh_ucphf_response_dimless(hh_cphf* h, const double* const* mo1, int64_t nprop, double* const* resp,
                                    hh_precision precision);
/* solve_dimless_cphf, uscf.rs:278-349: one Krylov solve over the spin blocks stacked per column;
 * precision as hh_cphf_solve_dimless (no preconditioner argument). */
hh_status hh_ucphf_solve_dimless(hh_cphf* h, const double* const* rhs, int64_t nprop, double* const* mo1,
                                 hh_precision precision, hh_krylov_stats* stats);
/* finalize_cphf, uscf.rs:369-407. */
hh_status hh_ucphf_finalize(hh_cphf* h, const double* const* f1mo, const double* const* s1mo,
                            const double* const* mo1, int64_t nprop, double* const* mo1_out, double* const* mo_e1);
/* get_cphf_hess, uscf.rs:409-440: de [3, 3, natm, natm]. */
hh_status hh_ucphf_hess(hh_cphf* h, const double* const* f1mo, const double* const* s1mo, const double* const* mo1,
                        const double* const* mo_e1, int64_t natm, double* de);
/* Add the UKS grid response (UHessKSNIMatmul::get_response_bra, hess_uks.rs:740-770) with
 * fxc_eff [ng, nvar, 2, nvar, 2] (device; must outlive the driver) to the unrestricted driver's
 * response list; the occupied MOs are the driver's. */
hh_status hh_ucphf_add_numint(hh_cphf* h, hh_numint* ni, int32_t den_type, const double* fxc_eff);

/* ---------------------------------------------------------------------------------------- */
/* Many CPHF systems in one batch (no reference counterpart: the reference solves one molecule  */
/* per RHessSCF / UHessSCF, rscf.rs:481-505, each with its own krylov_block call).              */

typedef struct hh_cphf_batch hh_cphf_batch;

/* A batch of dimensionless CPHF solves advanced in lockstep: each step of a Krylov cycle (the
 * response operator's three products, the projection coefficients, the update, both Gram / LDL
 * passes) is one grouped launch over every system still running, and the host reads one small
 * array per cycle for all systems.  Per system the solve is the reference's krylov_block
 * (krylov_block.rs:39-197: same basis, drop rule, convergence test and projected solve) on that
 * system's solve_dimless_cphf equations (rscf.rs:278-300, uscf.rs:278-349) in FP64, with one
 * subspace (no hard restart: as the reference with cphf_max_space >= cphf_max_cycle). */
hh_status hh_cphf_batch_create(double tol, int32_t max_cycle, double lindep, void* stream, hh_cphf_batch** out);
hh_status hh_cphf_batch_destroy(hh_cphf_batch* b);
/* Add one system: the driver (restricted or unrestricted; its orbitals and orbital energies;
 * level_shift 0) and the RI-JK handle prepared with the same orbitals, whose resident response
 * matrix (built here if it is not) is the system's response operator.  rhs / mo1: one device
 * pointer per spin, [nmo[s], nocc[s], nprop], 1 <= nprop <= 128.  rhs is copied at once; mo1 is
 * written by hh_cphf_batch_solve.  The driver and the handle must outlive the batch. */
hh_status hh_cphf_batch_add(hh_cphf_batch* b, hh_cphf* drv, hh_rijk* rijk, const double* const* rhs, int64_t nprop,
                            double* const* mo1);
/* Solve every system.  stats (may be NULL): one entry per system in the order added.  Returns
 * HH_ERR_NOT_CONVERGED if a system did not converge within max_cycle (every mo1 is still
 * written).  Synchronous. */
hh_status hh_cphf_batch_solve(hh_cphf_batch* b, hh_krylov_stats* stats);
/* Systems in the batch, lockstep cycles of the last solve and the most launches one cycle made. */
hh_status hh_cphf_batch_info(hh_cphf_batch* b, int64_t* nsys, int64_t* cycles, int64_t* launches_per_cycle);
/* Wall seconds of the last solve's phases: the resident matrices' builds, set-up with the first
 * (right-hand side) block, the lockstep cycles, the projected solves with the solutions'
 * assembly. */
hh_status hh_cphf_batch_times(hh_cphf_batch* b, double* seconds4);

/* ---------------------------------------------------------------------------------------- */
/* Learned diagonal CPHF preconditioner (no reference counterpart: the reference preconditions */
/* CPHF only by the orbital-energy scaling of the dimensionless form, rscf.rs:237-263).       */

typedef struct hh_mlp hh_mlp;
#define HH_MLP_NPARAM 8641  /* 3 -> 64 -> 64 -> 64 -> 1, ReLU, FP32 */

/* Network with seeded He-uniform init (an untrained network predicts ~0, i.e. diag ~1). */
hh_status hh_mlp_create(uint64_t seed, void* stream, hh_mlp** out);
hh_status hh_mlp_destroy(hh_mlp* h);
/* y [n] = f(x [n, 3]) (float, device, column-major). */
hh_status hh_mlp_forward(hh_mlp* h, const float* x, int64_t n, float* y);
/* Weighted MSE sum w (f(x) - t)^2 / sum w (host *loss) and its gradient grad [HH_MLP_NPARAM]
 * (device). */
hh_status hh_mlp_loss_and_grad(hh_mlp* h, const float* x, const float* t, const float* w, int64_t n, float* grad,
                               double* loss);
/* `epochs` full-batch Adam steps; *loss = loss before the last step.  Synchronous. */
hh_status hh_mlp_train(hh_mlp* h, const float* x, const float* t, const float* w, int64_t n, int32_t epochs, float lr,
                       double* loss);
hh_status hh_mlp_get_params(const hh_mlp* h, float* host_out);
hh_status hh_mlp_set_params(hh_mlp* h, const float* host_in);

/* RI pair norms pair[m, i] = sum_P ((C^T B_P C_o)[m, i])^2 for all MOs m of mo_coeff [nao, nmo]
 * (device) and the handle's occupied orbitals of spin s.  Synchronous. */
hh_status hh_rijk_pair_norms(hh_rijk* h, int32_t s, const double* mo_coeff, int64_t nmo, double* pair);
/* Features x [nvir*nocc, 3] (float) of the virtual-occupied pairs (sample a + nvir*i):
 * standardised ln(e_a - e_i), ln pair[a,i], ln rms_p |r[a,i,p]| for r [nmo, nocc, nprop]. */
hh_status hh_cphf_precond_features(hh_cphf* h, const double* pair, const double* r, int64_t nprop, float* x);
/* Training target from a (residual r, solution xs) pair: t = ln clip(sum_p r xs / sum_p r^2,
 * 0.05, 20), weight w = sum_p r^2 (both float [nvir*nocc]). */
hh_status hh_cphf_precond_target(hh_cphf* h, const double* r, const double* xs, int64_t nprop, float* t, float* w);
/* diag [nmo, nocc] = exp(clip(y, -3, 3)) on virtual rows, 1 on occupied rows; pass it as the
 * `diag` argument of hh_cphf_solve_dimless. */
hh_status hh_cphf_precond_diag(hh_cphf* h, const float* y, double* diag);

/* ---------------------------------------------------------------------------------------- */
/* Dense FP64 linear algebra on the device                                                    */

/* Solve H X = G in place (H overwritten by LU factors, G by X); replaces solve_general as
 * used by projected_solve (krylov_block.rs:214).  HH_ERR_INVALID if singular.  Synchronous. */
hh_status hh_linalg_solve(double* H, int64_t nd, double* G, int64_t nrhs, void* stream);
/* The same solve for callers that cannot allocate device memory (code that must not call
 * hipMalloc, tight loops, harnesses that forbid it); no counterpart in the reference.  Allocates
 * nothing: the pivots and status word live in a static device array of 8 * 8192 + 256 bytes per
 * GPU, so nd <= 8192 (else HH_ERR_INVALID).  The array is shared by all calls on one GPU and keeps
 * the last call's contents: calls that overlap on one GPU race.  Synchronous. */
hh_status hh_linalg_solve_noalloc(double* H, int64_t nd, double* G, int64_t nrhs, void* stream);
/* Symmetric eigendecomposition (rstsr eigh as decomp_j2c_eig / decomp_j2c_cd use it,
 * pure_decompose.rs:11-20, 41-45): the lower triangle of H [nd, nd] is read; H receives the
 * orthonormal eigenvectors (columns), w [nd] (device) the eigenvalues in ascending order.
 * Parallel cyclic Jacobi on the device.  HH_ERR_INVALID if not finite.  Synchronous. */
hh_status hh_linalg_eigh(double* H, int64_t nd, double* w, void* stream);
/* Lower Cholesky in place (decomp_j2c_cd, pure_decompose.rs:107-133).  HH_ERR_INVALID if not
 * positive definite.  Synchronous. */
hh_status hh_linalg_cholesky(double* H, int64_t nd, void* stream);
/* Triangular solve with lower L [nd, nd]: side_right = 0: X[nd, n2] <- L^-1 X;
 * side_right = 1: X[n2, nd] <- X L^-T (solve_by_j2c side Right, pure_decompose.rs:179-380). */
hh_status hh_linalg_trsm_lower(const double* L, int64_t nd, double* X, int64_t n2, int64_t ldx, int32_t side_right,
                               void* stream);

/* ---------------------------------------------------------------------------------------- */
/* Harmonic analysis (src/hessian/vib.rs).  HOST arrays: the readout of a [3 natm, 3 natm]    */
/* Hessian is FP64 CPU work (the caller copies the Hessian from the device).  Synchronous.     */

#define HH_ROTOR_ATOM 0
#define HH_ROTOR_LINEAR 1
#define HH_ROTOR_REGULAR 2
#define HH_TRV_V 0                /* vibration ("V") */
#define HH_TRV_TR 1               /* translation / rotation ("TR") */
#define HH_TRV_NONE 2             /* near-zero force constant, not TR ("-") */

/* Output arrays of hh_vib_harmonic_analysis (fields of VibInfo, vib.rs:221-256); ndof = 3 natm.
 * Any pointer may be NULL (not written).  Per-mode arrays [ndof]; q, w, x [ndof, ndof] column
 * major with modes as columns.  omega is the magnitude of the frequency in cm^-1, imag = 1 for
 * an imaginary mode. */
typedef struct hh_vib_out {
  double* omega;
  int32_t* imag;
  double* q;              /* mass-weighted normal modes */
  double* w;              /* un-mass-weighted */
  double* x;              /* un-mass-weighted, normalised */
  int64_t* degeneracy;
  int32_t* trv;           /* HH_TRV_* */
  double* mu;             /* reduced mass [u] */
  double* k;              /* force constant [mDyne/A] */
  double* dq0;            /* RMS deviation v=0 [a0 u^1/2] */
  double* qtp0;           /* turning point v=0, mass-weighted [a0 u^1/2] */
  double* xtp0;           /* turning point v=0, Cartesian [a0] */
  double* theta_vib;      /* characteristic temperature [K] */
} hh_vib_out;

/* Frequencies and normal modes of hess [3 natm, 3 natm] (E_h/a0^2, not mass-weighted) at geom
 * [3, natm] (a0) with masses [natm] (u), translations / rotations projected out on request
 * (replaces harmonic_analysis, vib.rs:336-552). */
hh_status hh_vib_harmonic_analysis(const double* hess, const double* geom, const double* mass, int64_t natm,
                                   int32_t project_trans, int32_t project_rot, const hh_vib_out* out);
/* Rotational constants out3[3] of a mass-centred geometry; unit 0 GHz, 1 cm^-1
 * (replaces rotation_const, vib.rs:152-197). */
hh_status hh_vib_rotation_const(const double* mass, const double* geom_c, int64_t natm, int32_t unit,
                                double* out3);
/* HH_ROTOR_* from rotational constants in GHz (replaces get_rotor_type, vib.rs:203-211). */
hh_status hh_vib_rotor_type(const double* rc_ghz, int32_t* rotor);
/* Orthonormal translation (space bit 1) / rotation (bit 2) basis tr [3 natm, ncol] (buffer for
 * 6 columns), rank by the default tolerance (replaces get_tr_space, vib.rs:52-54). */
hh_status hh_vib_tr_space(const double* mass, const double* geom, int64_t natm, int32_t space, double* tr,
                          int64_t* ncol);

/* ---------------------------------------------------------------------------------------- */
/* Utility: the library's strided batched GEMM (no reference counterpart; exposed so callers  */
/* and tests can drive the matrix-core kernel directly).                                      */
/* C[m,n] = alpha sum_k A[m,k] B[k,n] + beta C[m,n]; A[m,k] at a_kmajor ? m*lda+k : m+k*lda;   */
/* B[k,n] at b_kmajor ? k+n*ldb : k*ldb+n; C[m,n] at m+n*ldc; batch b0 (fastest), b1.         */
hh_status hh_util_gemm(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                       int32_t a_kmajor, const double* B, int64_t ldb, int32_t b_kmajor, double beta, double* C,
                       int64_t ldc, int64_t nb0, int64_t nb1, const int64_t* strides6, hh_precision precision,
                       void* stream);

/* The matrix-core NT product C[m,n] = alpha sum_k A[m*lda+k] B[n*ldb+k] + beta C[m+n*ldc] in FP64
 * (no reference counterpart; the building block of the FP64 contractions).  K a multiple of 32,
 * lda / ldb multiples of 8.  config 0..2 selects a tile configuration (208x128, 192x128,
 * 192x192) with split rule ksplit (> 0 fixed, 0 / -1 automatic); config -1 uses the autotuner:
 * the first product of a shape class (M, N and K rounded up to powers of two) times every
 * configuration on `stream` and the choice is kept for the process (hh_util_gemm_nt_choice
 * reports it, config -1 if none).  hh_util_set_autotune(0) keeps the default configuration for
 * classes not yet tuned. */
hh_status hh_util_gemm_nt(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                          const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int32_t config,
                          int32_t ksplit, void* stream);
hh_status hh_util_set_autotune(int32_t enable);
/* Measured MFMA issue-rate ceiling in TFLOP/s (no reference counterpart; the denominator beside
 * the published peak): kind 0 = v_mfma_f32_16x16x32_bf16, 1 = v_mfma_f64_16x16x4f64,
 * 2 = v_mfma_f32_32x32x16_bf16; every wave runs `iters` trips of 8 (kind 2: 4) independent
 * accumulator chains, 8 waves per SIMD on every CU. */
hh_status hh_util_mfma_peak(int32_t kind, int32_t iters, double* tflops, void* stream);
hh_status hh_util_gemm_nt_choice(int64_t M, int64_t N, int64_t K, int32_t* config, int32_t* ksplit);
/* The same product from FP64 operands in split-BF16 x3 (the RI-K hot product's arithmetic: A and
 * B split into bf16 hi / lo, three MFMA products, FP32 accumulation, split-K partials summed in
 * FP64; no reference counterpart).  config 0 / 1: 208 x 256 tiles as 1 x 8 or 2 x 4 waves; 2: 400 x 128
 * tiles (2 x 4 waves, swizzled LDS rows); 3: 128 x 128 tiles (2 x 2 waves, swizzled LDS rows, two
 * workgroups per CU); -1: the
 * autotuner (its choice for the shape class: hh_util_gemm_nt_bf16x3_choice).  lda, ldb >= K. */
hh_status hh_util_gemm_nt_bf16x3(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                                 const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int32_t config,
                                 int32_t ksplit, void* stream);
hh_status hh_util_gemm_nt_bf16x3_choice(int64_t M, int64_t N, int64_t K, int32_t* config, int32_t* ksplit);

/* Device-to-device copy of `bytes` on `stream` (no reference counterpart; lets foreign
 * callers of hh_krylov_block move data in and out of the solver's buffers). */
hh_status hh_util_copy(void* dst, const void* src, int64_t bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HESSHIP_H */
