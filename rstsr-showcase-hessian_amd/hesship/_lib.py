"""ctypes binding of libhesship.so (the C ABI declared in include/hesship.h).

The library is loaded from this directory.  There is no fallback: if the library is missing
or fails to load, every entry point raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhesship.so")

HH_OK = 0
HH_ERR_INVALID = 1
HH_ERR_HIP = 2
HH_ERR_NOT_CONVERGED = 3
HH_ERR_NOMEM = 4
HH_ERR_STATE = 5
HH_ERR_INTERNAL = 6

PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}

_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32
_d = C.c_double
_pp = C.POINTER(C.c_void_p)
_pi64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes)
SIGNATURES = {
    "hh_last_error": (C.c_char_p, []),
    "hh_version": (C.c_char_p, []),
    "hh_pack_triu_tilde": (_i32, [_p, _i64, _i64, _p, _p]),
    "hh_unpack_triu": (_i32, [_p, _i64, _i64, _p, _p]),
    "hh_rijk_create": (_i32, [_p, _i64, _i64, _d, _d, _i32, _p, C.POINTER(C.c_void_p)]),
    "hh_rijk_destroy": (_i32, [_p]),
    "hh_rijk_create_from_j3c": (_i32, [_p, _p, _i64, _i64, _d, _d, _d, _i32, _p, C.POINTER(C.c_void_p)]),
    "hh_rijk_create_from_j3c_decomp": (_i32, [_p, _p, _i64, _i64, _i32, _d, _d, _d, _i32, _p,
                                              C.POINTER(C.c_void_p)]),
    "hh_rijk_prepare_response": (_i32, [_p, _pp, _pp, _pi64]),
    "hh_rijk_nocc": (_i32, [_p, _i32, _pi64]),
    "hh_rijk_get_response_bra": (_i32, [_p, _pp, _i64, _pp, _i32]),
    "hh_rijk_response_bra_separated": (_i32, [_p, _pp, _i64, _p, _pp, _i32]),
    "hh_rijk_set_batch": (_i32, [_p, _i64]),
    "hh_rijk_set_profiling": (_i32, [_p, _i32]),
    "hh_rijk_set_k_path": (_i32, [_p, _i32]),
    "hh_rijk_set_response_path": (_i32, [_p, _i32, _i64]),
    "hh_rijk_get_timing": (_i32, [_p, C.POINTER(C.c_double)]),
    "hh_rijk_reset_timing": (_i32, [_p]),
    "hh_util_gemm": (_i32, [_i64, _i64, _i64, _d, _p, _i64, _i32, _p, _i64, _i32, _d, _p, _i64, _i64, _i64,
                            _pi64, _i32, _p]),
}


class KrylovStats(C.Structure):
    _fields_ = [("cycles", C.c_int32), ("restarts", C.c_int32), ("aop_calls", C.c_int32),
                ("refinements", C.c_int32), ("max_innerprod", C.c_double), ("residual", C.c_double),
                ("fp64_fallback", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


AOP_FN = C.CFUNCTYPE(C.c_int32, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p)
SIGNATURES.update({
    "hh_krylov_block": (_i32, [AOP_FN, _p, _p, _p, _i64, _i64, _d, _i32, _i32, _d, _p, C.POINTER(KrylovStats), _p]),
    "hh_krylov_orth_block": (_i32, [_p, _i64, _i64, _d, _p, _p, _pi64, _p]),
    "hh_cphf_create": (_i32, [_p, _i64, _i64, _p, _p, _d, _d, _i32, _i32, _d, _p, C.POINTER(C.c_void_p)]),
    "hh_cphf_destroy": (_i32, [_p]),
    "hh_cphf_add_rijk": (_i32, [_p, _p]),
    "hh_cphf_dimless_rhs": (_i32, [_p, _p, _p, _i64, _p]),
    "hh_cphf_response_mo": (_i32, [_p, _p, _i64, _p, _i32]),
    "hh_cphf_response_dimless": (_i32, [_p, _p, _i64, _p, _i32]),
    "hh_cphf_solve_dimless": (_i32, [_p, _p, _i64, _p, _i32, _p, C.POINTER(KrylovStats)]),
    "hh_cphf_finalize": (_i32, [_p, _p, _p, _p, _i64, _p, _p]),
    "hh_cphf_hess": (_i32, [_p, _p, _p, _p, _p, _i64, _p]),
    "hh_ucphf_create": (_i32, [_pp, _i64, _pi64, _pp, _pp, _d, _d, _i32, _i32, _d, _p, C.POINTER(C.c_void_p)]),
    "hh_ucphf_dimless_rhs": (_i32, [_p, _pp, _pp, _i64, _pp]),
    "hh_ucphf_response_mo": (_i32, [_p, _pp, _i64, _pp, _i32]),
    "hh_ucphf_response_dimless": (_i32, [_p, _pp, _i64, _pp, _i32]),
    "hh_ucphf_solve_dimless": (_i32, [_p, _pp, _i64, _pp, _i32, C.POINTER(KrylovStats)]),
    "hh_ucphf_finalize": (_i32, [_p, _pp, _pp, _pp, _i64, _pp, _pp]),
    "hh_ucphf_hess": (_i32, [_p, _pp, _pp, _pp, _pp, _i64, _p]),
    "hh_linalg_solve": (_i32, [_p, _i64, _p, _i64, _p]),
    "hh_linalg_solve_noalloc": (_i32, [_p, _i64, _p, _i64, _p]),
    "hh_linalg_cholesky": (_i32, [_p, _i64, _p]),
    "hh_linalg_eigh": (_i32, [_p, _i64, _p, _p]),
    "hh_cint_int1e": (_i32, [_p, _i32, _i32, _i32, _p, _p]),
    "hh_hess_nuc_repl": (_i32, [_p, _p, _p]),
    "hh_scf_rhf": (_i32, [_p, _p, _i32, _d, _i32, _p, _p, _p, _p, _p]),
    "hh_scf_rks": (_i32, [_p, _p, _p, _i32, _i32, _p, _p, _d, _i32, _d, _i32, _p, _p, _p, _p, _p]),
    "hh_numint_vxc_matrix": (_i32, [_p, _i32, _p, _p]),
    "hh_scf_uhf": (_i32, [_p, _p, _p, _i32, _i32, _p, _p, _d, _i32, _i32, _d, _i32, _p, _p, _p, _p, _p]),
    "hh_hess_hcore": (_i32, [_p, _p, _p, _p]),
    "hh_hess_ovlp": (_i32, [_p, _p, _p, _p]),
    "hh_hess_hcore_deriv1": (_i32, [_p, _p, _p]),
    "hh_hess_ovlp_deriv1": (_i32, [_p, _p, _p]),
    "hh_hess_weighted_density": (_i32, [_p, _i64, _i64, _p, _p, _p]),
    "hh_linalg_trsm_lower": (_i32, [_p, _i64, _p, _i64, _i64, _i32, _p]),
    "hh_util_copy": (_i32, [_p, _p, _i64, _p]),
    "hh_vib_harmonic_analysis": (_i32, [_p, _p, _p, _i64, _i32, _i32, _p]),
    "hh_vib_rotation_const": (_i32, [_p, _p, _i64, _i32, _p]),
    "hh_vib_rotor_type": (_i32, [_p, C.POINTER(C.c_int32)]),
    "hh_vib_tr_space": (_i32, [_p, _p, _i64, _i32, _p, _pi64]),
    "hh_mlp_create": (_i32, [C.c_uint64, _p, C.POINTER(C.c_void_p)]),
    "hh_mlp_destroy": (_i32, [_p]),
    "hh_mlp_forward": (_i32, [_p, _p, _i64, _p]),
    "hh_mlp_loss_and_grad": (_i32, [_p, _p, _p, _p, _i64, _p, C.POINTER(C.c_double)]),
    "hh_mlp_train": (_i32, [_p, _p, _p, _p, _i64, _i32, C.c_float, C.POINTER(C.c_double)]),
    "hh_mlp_get_params": (_i32, [_p, C.POINTER(C.c_float)]),
    "hh_mlp_set_params": (_i32, [_p, C.POINTER(C.c_float)]),
    "hh_rijk_pair_norms": (_i32, [_p, _i32, _p, _i64, _p]),
    "hh_cphf_precond_features": (_i32, [_p, _p, _p, _i64, _p]),
    "hh_cphf_precond_target": (_i32, [_p, _p, _p, _i64, _p, _p]),
    "hh_cphf_precond_diag": (_i32, [_p, _p, _p]),
    "hh_cphf_add_numint": (_i32, [_p, _p, _i32, _p]),
    "hh_ucphf_add_numint": (_i32, [_p, _p, _i32, _p]),
    "hh_cphf_set_verbose": (_i32, [_p, _i32]),
    "hh_numint_create": (_i32, [_p, _i64, _i64, _i64, _p, _p, C.POINTER(C.c_void_p)]),
    "hh_numint_destroy": (_i32, [_p]),
    "hh_numint_set_batch": (_i32, [_p, _i64]),
    "hh_numint_set_fused": (_i32, [_p, _i32]),
    "hh_numint_rho1_one_bra_mult_ket": (_i32, [_p, _i32, _p, _i64, _p, _i64, _p, _i32]),
    "hh_numint_fxc_pot_bra_trans": (_i32, [_p, _i32, _p, _p, _i64, _p, _i64, _p, _i32]),
    "hh_numint_rks_response_bra": (_i32, [_p, _i32, _p, _p, _i64, _p, _i64, _p, _i32]),
    "hh_xceff_transform": (_i32, [_i32, _i32, _p, _i64, _p, _i64, _i64, _p, _p]),
    "hh_rijk_skeleton": (_i32, [_p, _pp, _pi64, _pp, _p, _p, _p, _p, _p, _p, _i64, _pi64, _i64, _pi64, _pi64, _p,
                                 _pp]),
    "hh_numint_setup_rho": (_i32, [_p, _i32, _i32, _pp, _p]),
    "hh_numint_hessian_setup": (_i32, [_p, _i32, _i32, _pp, _i64, _pi64, _p, _p, _p, _pp, _pp, _pp, _pp]),
    "hh_xceff_transform_polarized": (_i32, [_i32, _i32, _p, _i64, _i64, _p, _i64, _i64, _p, _p]),
    "hh_xceff_libxc_to_xcfun_map": (_i32, [_i32, _i32, _pi64, _i64, _pi64]),
    "hh_numint_uks_fxc_pot_bra_trans": (_i32, [_p, _i32, _p, _p, _i64, _pp, _pi64, _pp, _i32]),
    "hh_numint_uks_response_bra": (_i32, [_p, _i32, _p, _pp, _i64, _pp, _pi64, _pp, _i32]),
    "hh_util_gemm_nt": (_i32, [_i64, _i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64, _i32, _i32, _p]),
    "hh_util_set_autotune": (_i32, [_i32]),
    "hh_util_mfma_peak": (_i32, [_i32, _i32, C.POINTER(C.c_double), _p]),
    "hh_util_gemm_nt_choice": (_i32, [_i64, _i64, _i64, C.POINTER(_i32), C.POINTER(_i32)]),
    "hh_util_gemm_nt_bf16x3": (_i32, [_i64, _i64, _i64, _d, _p, _i64, _p, _i64, _d, _p, _i64, _i32, _i32, _p]),
    "hh_util_gemm_nt_bf16x3_choice": (_i32, [_i64, _i64, _i64, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "hh_cphf_set_graphs": (_i32, [_p, _i32]),
    "hh_cphf_graph_stats": (_i32, [_p, _pi64, _pi64]),
    "hh_cphf_finalize_reuses": (_i32, [_p, _pi64]),
    "hh_cint_create": (_i32, [_p, _i32, _p, _i32, _p, _i64, _i32, C.POINTER(C.c_void_p)]),
    "hh_cint_destroy": (_i32, [_p]),
    "hh_cint_nao": (_i32, [_p, _pi64, _pi64]),
    "hh_cint_eval_gto": (_i32, [_p, _i32, _p, _i64, _p, _p]),
    "hh_cint_int2c2e": (_i32, [_p, _i32, _p, _p]),
    "hh_cint_int3c2e": (_i32, [_p, _p, _i32, _i32, _i32, _p, _p]),
    "hh_cint_int3c2e_s2ij": (_i32, [_p, _p, _i32, _i32, _p, _p]),
    "hh_cint_pair_create": (_i32, [_p, _p, C.POINTER(C.c_void_p)]),
    "hh_cint_pair_destroy": (_i32, [_p]),
    "hh_cint_int3c_deriv": (_i32, [_p, _i32, _i64, _i64, _p, _p]),
    "hh_cphf_batch_create": (_i32, [_d, _i32, _d, _p, C.POINTER(C.c_void_p)]),
    "hh_cphf_batch_destroy": (_i32, [_p]),
    "hh_cphf_batch_add": (_i32, [_p, _p, _p, _pp, _i64, _pp]),
    "hh_cphf_batch_solve": (_i32, [_p, _p]),
    "hh_cphf_batch_info": (_i32, [_p, _pi64, _pi64, _pi64]),
    "hh_cphf_batch_times": (_i32, [_p, C.POINTER(C.c_double)]),
    "hh_becke_create": (_i32, [_i64, _p, _p, _p, C.POINTER(C.c_void_p)]),
    "hh_becke_destroy": (_i32, [_p]),
    "hh_becke_partition": (_i32, [_p, _p, _i64, _i32, _pi64, _p, _i32, _i32, _p, _p, _p, _p, _i64, _p, _p, _i64, _p,
                                  _p, _i64, _p]),
    "hh_numint_hessian_setup_becke": (_i32, [_p, _i32, _i32, _pp, _i64, _pi64, _p, _p, _p, _p, _p, _p, _pi64, _i32,
                                             _p, _pp, _p, _pp, _pp, _pp, _pp]),
    "hh_xc_info": (_i32, [_i32, _i32, _i32, C.POINTER(_i32), C.POINTER(_d), _pi64]),
    "hh_xc_eval": (_i32, [_i32, _i32, _i32, _p, _i64, _i64, _i64, _p, _i64, _p]),
    "hh_xc_eval_eff": (_i32, [_i32, _i32, _i32, _d, _p, _i64, _i64, _i64, _i32, _p, _p, _p, _p]),
})


class HesshipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[hh status {code}] {msg}")
        self.code = code


class KrylovNotConverged(HesshipError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libhesship.so not built at {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != HH_OK:
        msg = lib().hh_last_error().decode(errors="replace")
        if code == HH_ERR_NOT_CONVERGED:
            raise KrylovNotConverged(code, msg)
        raise HesshipError(code, msg)


def ptr_array(ptrs):
    arr = (C.c_void_p * len(ptrs))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr
