"""hesship: host-side binding of libhesship (MI355X / gfx950) for the analytic-Hessian hot path.

Mirrors the reference's (ajz34/rstsr-showcase-hessian) Rust API names:
  RHessRIJK, UHessRIJK, get_rijk_response_bra_separated     (src/ri_jk)
  krylov_block                                               (src/util/krylov_block.rs)
  BeckeMolTables.partition                                   (src/numint_matmul/becke_partition.rs)
  LibXCFunctional, libxc_eval_inner, libxc_eval_eff          (src/xceff/libxc_wrap.rs)
Device tensors are torch tensors on a ROCm device, in the reference's logical shapes; the
binding hands column-major device buffers to the C ABI.
"""
from ._lib import PREC, HesshipError, KrylovNotConverged, lib  # noqa: F401
from .tensors import fortran, empty_f  # noqa: F401
from .krylov import krylov_block, orth_block, solve, eigh, cholesky, trsm_lower  # noqa: F401
from .cphf import RHessCPHF, UHessCPHF, CPHFBatch  # noqa: F401
from .cint import CInt, Int3cGenerator, make_env  # noqa: F401
from .becke import BeckeMolTables  # noqa: F401
from .numint import NIMatmul, libxc_to_xcfun_mapping, transform_xc, transform_xc_polarized  # noqa: F401
from .xc import (LibXCFunctional, determine_den_type, determine_den_type_from_list,  # noqa: F401
                 eval_xc_eff_list, libxc_eval_eff, libxc_eval_inner)
from .rijk import RHessRIJK, UHessRIJK, get_rijk_response_bra_separated, pack_triu_tilde, unpack_triu  # noqa: F401
from .grids import Grids  # noqa: F401
from .hessian import RHF, RKS, UHF, UKS, RHessSCF, UHessSCF  # noqa: F401
