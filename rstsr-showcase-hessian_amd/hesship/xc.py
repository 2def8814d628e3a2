"""Exchange-correlation functionals on the device (reference: src/xceff/libxc_wrap.rs).

Mirrors the reference's use of LibXC: LibXCFunctional.from_identifier(name, spin),
determine_den_type, libxc_eval_inner (raw LibXC columns), libxc_eval_eff (exc, vxc_eff, fxc_eff),
and the functional-list loop of get_rho_vxc_fxc (src/numint_matmul/hess_rks.rs:174-218,
hess_uks.rs:56-108) as eval_xc_eff_list.  Functionals: LDA_X, LDA_C_VWN, LDA_C_VWN_RPA, GGA_X_B88,
GGA_C_LYP, HYB_GGA_XC_B3LYP, MGGA_X_TPSS, MGGA_C_TPSS, HYB_MGGA_XC_TPSS0.  rho: [ng, nvar]
unpolarised or [ng, nvar, 2] polarised (nvar 1 for RHO, 4 for SIGMA: density and its gradient, 5
for TAU: then the kinetic energy density tau), float64 on the device.
"""
import ctypes as C

from ._lib import check, lib
from .tensors import empty_f, fortran, stream_of

IDS = {"LDA_X": 1, "LDA_C_VWN": 7, "LDA_C_VWN_RPA": 8, "GGA_X_B88": 106, "GGA_C_LYP": 131, "HYB_GGA_XC_B3LYP": 402,
       "MGGA_X_TPSS": 202, "MGGA_C_TPSS": 231, "HYB_MGGA_XC_TPSS0": 396}
DEN_NAMES = {0: "rho", 1: "sigma", 2: "tau"}
DEN = {"rho": 0, "sigma": 1, "tau": 2}
NVAR = {"rho": 1, "sigma": 4, "tau": 5}


class LibXCFunctional:
    """A functional by LibXC identifier and spin ("unpolarized" | "polarized")."""

    def __init__(self, xc_id, spin="unpolarized"):
        self.xc_id = int(xc_id)
        if spin not in ("unpolarized", "polarized"):
            raise ValueError("spin must be 'unpolarized' or 'polarized'")
        self.spin = spin
        self.nspin = 1 if spin == "unpolarized" else 2
        den, exx = C.c_int32(), C.c_double()
        check(lib().hh_xc_info(self.xc_id, self.nspin, 0, C.byref(den), C.byref(exx), None))
        self.den_type = DEN_NAMES[den.value]
        self._exx = exx.value

    @classmethod
    def from_identifier(cls, name, spin="unpolarized"):
        key = name.upper()
        if key.startswith("XC_"):
            key = key[3:]
        if key not in IDS:
            raise ValueError(f"functional {name!r} is not evaluated by this library (have {sorted(IDS)})")
        return cls(IDS[key], spin)

    def hyb_exx_coef(self):
        return self._exx

    def ncol(self, deriv):
        n = C.c_int64()
        check(lib().hh_xc_info(self.xc_id, self.nspin, int(deriv), None, None, C.byref(n)))
        return n.value


def determine_den_type(xc_func):
    """libxc_wrap.rs:10-23."""
    return xc_func.den_type


def determine_den_type_from_list(xc_funcs):
    """libxc_wrap.rs:25-27: the family with the most density components."""
    return max((f.den_type for f in xc_funcs), key=lambda d: NVAR[d])


def _rho_layout(rho, xc_func, nvar_min):
    r = fortran(rho)
    if xc_func.nspin == 1:
        if r.dim() != 2 or r.shape[1] < nvar_min:
            raise ValueError(f"unpolarized rho must be [ng, >= {nvar_min}]")
        return r, r.shape[0], 0
    if r.dim() != 3 or r.shape[2] != 2 or r.shape[1] < nvar_min:
        raise ValueError(f"polarized rho must be [ng, >= {nvar_min}, 2]")
    return r, r.shape[0], r.shape[0] * r.shape[1]


def libxc_eval_inner(xc_func, rho, deriv):
    """libxc_wrap.rs:32-82: LibXC's columns [zk | 1st | 2nd] as [ng, ncol] (column-major)."""
    r, ld_rho, ld_spin = _rho_layout(rho, xc_func, NVAR[xc_func.den_type])
    ng = r.shape[0]
    out = empty_f((ng, xc_func.ncol(deriv)), r.device)
    check(lib().hh_xc_eval(xc_func.xc_id, xc_func.nspin, int(deriv), r.data_ptr(), ld_rho, ld_spin, ng,
                           out.data_ptr(), ng, stream_of(r.device)))
    return out


def _accumulate_eff(xc_func, rho, den_type, scale, exc, vxc, fxc, accumulate):
    r, ld_rho, ld_spin = _rho_layout(rho, xc_func, NVAR[xc_func.den_type])
    ng = r.shape[0]
    ptr = [t.data_ptr() if t is not None else None for t in (exc, vxc, fxc)]
    check(lib().hh_xc_eval_eff(xc_func.xc_id, xc_func.nspin, DEN[den_type], float(scale), r.data_ptr(), ld_rho,
                               ld_spin, ng, int(bool(accumulate)), ptr[0], ptr[1], ptr[2], stream_of(r.device)))


def _eff_shapes(ng, nvar, nspin):
    per = (nvar,) if nspin == 1 else (nvar, 2)
    return (ng,), (ng,) + per, (ng,) + per + per


def libxc_eval_eff(xc_func, rho, deriv=2):
    """libxc_wrap.rs:88-122 with transform_xc_inner (xc_deriv.rs:305-428): [exc, vxc_eff, fxc_eff][: deriv + 1];
    vxc_eff [ng, nvar(, 2)],
    fxc_eff [ng, nvar(, 2), nvar(, 2)] over the density components of the functional's family."""
    if deriv not in (0, 1, 2):
        raise ValueError("deriv must be 0, 1 or 2")
    ng = rho.shape[0]
    shapes = _eff_shapes(ng, NVAR[xc_func.den_type], xc_func.nspin)
    outs = [empty_f(s, rho.device) for s in shapes[:deriv + 1]] + [None] * (2 - deriv)
    _accumulate_eff(xc_func, rho, xc_func.den_type, 1.0, *outs, accumulate=False)
    return outs[:deriv + 1]


def eval_xc_eff_list(xc_func_list, rho):
    """The functional loop of get_rho_vxc_fxc (hess_rks.rs:193-216, hess_rks_becke.rs:192-205,
    hess_uks.rs:93-106): (exc, vxc, fxc) = sum_i scale_i * libxc_eval_eff(f_i), each term written
    into the leading slice of the list's strictest family.  xc_func_list: [(scale, functional)]."""
    if not xc_func_list:
        raise ValueError("xc_func_list must not be empty")
    nspin = {f.nspin for _, f in xc_func_list}
    if len(nspin) != 1:
        raise ValueError("all functionals of a list must share the spin setting")
    nspin = nspin.pop()
    den = determine_den_type_from_list([f for _, f in xc_func_list])
    ng = rho.shape[0]
    exc, vxc, fxc = (empty_f(s, rho.device).zero_() for s in _eff_shapes(ng, NVAR[den], nspin))
    for scale, f in xc_func_list:
        _accumulate_eff(f, rho, den, scale, exc, vxc, fxc, accumulate=True)
    return exc, vxc, fxc
