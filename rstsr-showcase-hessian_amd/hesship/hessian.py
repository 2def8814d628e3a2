"""The whole restricted analytic Hessian on the device (reference: src/hessian/rscf.rs RHessSCF
with its components: HessNucRepl, nuc_repl.rs; RHessHcore, hcore.rs; RHessOvlp, ovlp.rs; the
RI-JK electron interaction RHessRIJK, hess_r.rs; for Kohn-Sham the grid component
RHessKSNIMatmul with the Becke grid shift, hess_rks_becke.rs; the CPHF part) and the
density-fitted RHF / RKS that supplies its orbitals (the reference reads them from PySCF,
tests/hessian_rhf/mod.rs:39-41, tests/hessian_rks_becke/mod.rs:39-55).

    mol = CInt.from_basis(atoms, basis); aux = CInt.from_basis(atoms, aux_basis)
    scf = RHF(mol, aux, nocc).run()                                   # or
    scf = RKS(mol, aux, nocc, [(1.0, LibXCFunctional.from_identifier("HYB_GGA_XC_B3LYP"))],
              Grids(charges, coords)).run()
    hess = RHessSCF(mol, aux, scf.mo_coeff, scf.mo_occ, scf.mo_energy, xc=scf.xc, grids=scf.grids)
    de = hess.make_hess()             # [3, 3, natm, natm]; hess.result holds the parts

Every term runs in the library (hh_* entry points); torch holds device memory and adds the
[3, 3, natm, natm] parts.
"""
import ctypes as C
import time

import numpy as np
import torch

from ._lib import check, lib
from .cphf import RHessCPHF, UHessCPHF
from .rijk import RHessRIJK, UHessRIJK
from .tensors import empty_f, fortran, stream_of


def _dev(device):
    return device or torch.device("cuda", torch.cuda.current_device())


def _gemm(M, N, K, A, lda, a_kmajor, B, ldb, b_kmajor, Cm, ldc, alpha=1.0, beta=0.0, nb0=1, strides=None, dev=None):
    st = (C.c_int64 * 6)(*strides) if strides is not None else None
    check(lib().hh_util_gemm(M, N, K, float(alpha), A.data_ptr(), lda, int(a_kmajor), B.data_ptr(), ldb,
                             int(b_kmajor), float(beta), Cm.data_ptr(), ldc, nb0, 1, st, 0, stream_of(dev)))


def weighted_density(mo_coeff, w):
    """sum_i w_i C_i C_i^T (get_dm0_restricted with w = mo_occ, get_dme0_restricted with
    w = mo_occ mo_energy)."""
    c = fortran(mo_coeff)
    nao, nmo = c.shape
    wh = np.ascontiguousarray(np.asarray(w, dtype=np.float64))
    out = empty_f((nao, nao), c.device)
    check(lib().hh_hess_weighted_density(c.data_ptr(), nao, nmo, wh.ctypes.data, out.data_ptr(), stream_of(c.device)))
    return out


class RHF:
    """Density-fitted restricted Hartree-Fock on the device (hh_scf_rhf)."""
    xc = None
    grids = None

    def __init__(self, mol, aux, nocc, conv_tol=1e-12, max_cycle=100, device=None):
        self.mol, self.aux, self.nocc = mol, aux, int(nocc)
        self.conv_tol, self.max_cycle = float(conv_tol), int(max_cycle)
        self.device = _dev(device)
        self.mo_coeff = self.mo_energy = self.mo_occ = None
        self.e_tot = None
        self.cycles = 0

    def run(self):
        nao = self.mol.nao
        c = empty_f((nao, nao), self.device)
        e = torch.empty(nao, dtype=torch.float64, device=self.device)
        en, cy = C.c_double(), C.c_int32()
        check(lib().hh_scf_rhf(self.mol._h, self.aux._h, self.nocc, self.conv_tol, self.max_cycle, c.data_ptr(),
                               e.data_ptr(), C.byref(en), C.byref(cy), stream_of(self.device)))
        self.mo_coeff, self.mo_energy = c, e
        self.mo_occ = np.where(np.arange(nao) < self.nocc, 2.0, 0.0)
        self.e_tot, self.cycles = en.value, cy.value
        return self


def xc_list_info(xc):
    """(den_type, exx, ids, scales) of a functional list [(scale, LibXCFunctional)]: the strictest
    family (determine_den_type_from_list) and sum_i scale_i hyb_exx_coef_i."""
    from .xc import determine_den_type_from_list
    den = determine_den_type_from_list([f for _, f in xc])
    exx = sum(float(sc) * f.hyb_exx_coef() for sc, f in xc)
    return den, exx, [int(f.xc_id) for _, f in xc], [float(sc) for sc, _ in xc]


class RKS(RHF):
    """Density-fitted restricted Kohn-Sham on the device (hh_scf_rks): the functional list xc
    [(scale, LibXCFunctional)] on the partitioned grid of `grids` (hesship.grids.Grids), exact
    exchange scaled by the list's hybrid fraction."""

    def __init__(self, mol, aux, nocc, xc, grids, conv_tol=1e-12, max_cycle=100, device=None):
        super().__init__(mol, aux, nocc, conv_tol, max_cycle, device)
        self.xc, self.grids = list(xc), grids

    def run(self):
        from .numint import DEN, NIMatmul
        if self.grids.weights is None:
            self.grids.build()
        den, exx, ids, scales = xc_list_info(self.xc)
        ao = self.mol.eval_gto(self.grids.coords, 0 if den == "rho" else 1)
        ni = NIMatmul(ao, self.grids.weights)
        nao = self.mol.nao
        c = empty_f((nao, nao), self.device)
        e = torch.empty(nao, dtype=torch.float64, device=self.device)
        en, cy = C.c_double(), C.c_int32()
        idv = (C.c_int32 * len(ids))(*ids)
        scv = (C.c_double * len(scales))(*scales)
        check(lib().hh_scf_rks(self.mol._h, self.aux._h, ni._h, DEN[den], len(ids), idv, scv, exx, self.nocc,
                               self.conv_tol, self.max_cycle, c.data_ptr(), e.data_ptr(), C.byref(en), C.byref(cy),
                               stream_of(self.device)))
        self.mo_coeff, self.mo_energy = c, e
        self.mo_occ = np.where(np.arange(nao) < self.nocc, 2.0, 0.0)
        self.e_tot, self.cycles = en.value, cy.value
        return self


class UHF:
    """Density-fitted unrestricted Hartree-Fock on the device (hh_scf_uhf): nalpha / nbeta
    singly occupied orbitals; mo_coeff, mo_energy, mo_occ as (alpha, beta) pairs."""
    xc = None
    grids = None

    def __init__(self, mol, aux, nalpha, nbeta, conv_tol=1e-12, max_cycle=100, device=None):
        self.mol, self.aux = mol, aux
        self.nocc = (int(nalpha), int(nbeta))
        self.conv_tol, self.max_cycle = float(conv_tol), int(max_cycle)
        self.device = _dev(device)
        self.mo_coeff = self.mo_energy = self.mo_occ = None
        self.e_tot = None
        self.cycles = 0

    def _run(self, grid_h=None, den=None, ids=(), scales=(), exx=1.0):
        from .numint import DEN
        from ._lib import ptr_array
        nao = self.mol.nao
        cs = [empty_f((nao, nao), self.device) for _ in range(2)]
        es = [torch.empty(nao, dtype=torch.float64, device=self.device) for _ in range(2)]
        en, cy = C.c_double(), C.c_int32()
        idv = (C.c_int32 * max(1, len(ids)))(*ids)
        scv = (C.c_double * max(1, len(scales)))(*scales)
        check(lib().hh_scf_uhf(self.mol._h, self.aux._h, grid_h, DEN[den] if den else 0, len(ids), idv, scv, exx,
                               self.nocc[0], self.nocc[1], self.conv_tol, self.max_cycle,
                               ptr_array([c.data_ptr() for c in cs]), ptr_array([e.data_ptr() for e in es]),
                               C.byref(en), C.byref(cy), stream_of(self.device)))
        self.mo_coeff, self.mo_energy = tuple(cs), tuple(es)
        self.mo_occ = tuple(np.where(np.arange(nao) < n, 1.0, 0.0) for n in self.nocc)
        self.e_tot, self.cycles = en.value, cy.value
        return self

    def run(self):
        return self._run()


class UKS(UHF):
    """Density-fitted unrestricted Kohn-Sham on the device (hh_scf_uhf with a grid): xc
    [(scale, LibXCFunctional)] with polarised functionals, on the grid of `grids`."""

    def __init__(self, mol, aux, nalpha, nbeta, xc, grids, conv_tol=1e-12, max_cycle=100, device=None):
        super().__init__(mol, aux, nalpha, nbeta, conv_tol, max_cycle, device)
        if any(f.nspin != 2 for _, f in xc):
            raise ValueError("UKS needs polarized functionals")
        self.xc, self.grids = list(xc), grids

    def run(self):
        from .numint import NIMatmul
        if self.grids.weights is None:
            self.grids.build()
        den, exx, ids, scales = xc_list_info(self.xc)
        ao = self.mol.eval_gto(self.grids.coords, 0 if den == "rho" else 1)
        ni = NIMatmul(ao, self.grids.weights)
        return self._run(ni._h, den, ids, scales, exx)


def nuc_repl_hess(mol, device=None):
    """get_nuc_repl_hess (nuc_repl.rs:14-58): [3, 3, natm, natm]."""
    dev = _dev(device)
    natm = len(mol.atm)
    out = empty_f((3, 3, natm, natm), dev)
    check(lib().hh_hess_nuc_repl(mol._h, out.data_ptr(), stream_of(dev)))
    return out


def hcore_hess(mol, dm0):
    """get_hess_hcore (hcore.rs:105-125)."""
    d = fortran(dm0)
    natm = len(mol.atm)
    out = empty_f((3, 3, natm, natm), d.device)
    check(lib().hh_hess_hcore(mol._h, d.data_ptr(), out.data_ptr(), stream_of(d.device)))
    return out


def ovlp_hess(mol, dme0):
    """get_hess_ovlp (ovlp.rs:24-52)."""
    d = fortran(dme0)
    natm = len(mol.atm)
    out = empty_f((3, 3, natm, natm), d.device)
    check(lib().hh_hess_ovlp(mol._h, d.data_ptr(), out.data_ptr(), stream_of(d.device)))
    return out


def hcore_deriv1(mol, device=None):
    """generator_hcore_deriv1 (hcore.rs:140-161) for every atom: [nao, nao, 3, natm]."""
    dev = _dev(device)
    out = empty_f((mol.nao, mol.nao, 3, len(mol.atm)), dev)
    check(lib().hh_hess_hcore_deriv1(mol._h, out.data_ptr(), stream_of(dev)))
    return out


def ovlp_deriv1(mol, device=None):
    """generator_ovlp_deriv1 (ovlp.rs:67-82) for every atom: [nao, nao, 3, natm]."""
    dev = _dev(device)
    out = empty_f((mol.nao, mol.nao, 3, len(mol.atm)), dev)
    check(lib().hh_hess_ovlp_deriv1(mol._h, out.data_ptr(), stream_of(dev)))
    return out


def rijk_for(mol, aux, device=None, scale_k=1.0):
    """RHessRIJK::new_without_cderi (hess_r.rs:1751-1773): the metric of (mol, aux) with the
    Cholesky policy and threshold 1e-14, scale_j = 1, scale_k = 1 (RHF) or the hybrid fraction."""
    dev = _dev(device)
    j3c = mol.int3c2e_s2ij(aux, device=dev)
    j2c = aux.int2c2e("plain", device=dev)
    return RHessRIJK.from_j3c(j3c, j2c, 1.0, float(scale_k), threshold=1e-14, policy="cd")


class RHessSCF:
    """RHessSCF (rscf.rs:21-506) for RHF or RKS with the RI-JK electron interaction: nuclear
    repulsion, core Hamiltonian, overlap, RI-JK skeleton (K scaled by the hybrid fraction), for
    RKS the grid component with the Becke grid shift (de_xc_skeleton, vmat_deriv1 and the CPKS
    response, hess_rks_becke.rs), and the CPHF part, all on the device."""

    def __init__(self, mol, aux, mo_coeff, mo_occ, mo_energy, config=None, rijk=None, xc=None, grids=None):
        self.mol, self.aux = mol, aux
        self.mo_coeff = fortran(mo_coeff)
        self.device = self.mo_coeff.device
        self.mo_occ = np.asarray(mo_occ.cpu() if isinstance(mo_occ, torch.Tensor) else mo_occ, dtype=np.float64)
        self.mo_energy = np.asarray(mo_energy.cpu() if isinstance(mo_energy, torch.Tensor) else mo_energy,
                                    dtype=np.float64)
        self.config = dict(config or {})
        self.xc, self.grids = (list(xc) if xc else None), grids
        if self.xc and grids is None:
            raise ValueError("a Kohn-Sham Hessian needs the SCF's grids")
        exx = xc_list_info(self.xc)[1] if self.xc else 1.0
        self.rijk = rijk if rijk is not None else rijk_for(mol, aux, self.device, scale_k=exx)
        self.result = {}
        self.timing = {}  # seconds per stage of the last make_hess (every stage ends synchronised)

    @property
    def natm(self):
        return len(self.mol.atm)

    def _skeleton_rijk(self):
        mol, aux, dev = self.mol, self.aux, self.device

        def int3c(kind, p0, p1):
            sh0 = int(np.nonzero(aux.ao_loc == p0)[0][0])
            sh1 = int(np.nonzero(aux.ao_loc == p1)[0][0])
            return mol.int3c2e(aux, kind, shls=(sh0, sh1), device=dev)
        ranges = [(int(aux.ao_loc[0]), int(aux.ao_loc[-1]))]
        v1, v2, v3 = (aux.int2c2e(k, device=dev) for k in ("ip1", "ipip1", "ip1ip2"))
        return self.rijk.make_skeleton(self.mo_coeff, self.mo_occ, int3c, v1, v2, v3, ranges,
                                       mol.aoslice_by_atom(), aux.aoslice_by_atom())

    def _skeleton_xc(self, dm0):
        """make_skeleton_hess and get_deriv1_ao of the Becke grid component (hess_rks_becke.rs:
        1106-1266, 1403-1425): (de_xc_skeleton, vmat_deriv1 [nao, nao, 3, natm], the NIMatmul and
        the transformed kernel kept for the CPKS response)."""
        from .numint import NIMatmul
        from .xc import eval_xc_eff_list
        g = self.grids
        if g.weights is None:
            g.build()
        den = xc_list_info(self.xc)[0]
        ao = self.mol.eval_gto(g.coords, 3)
        ni = NIMatmul(ao, g.weights)
        rho = ni.setup_rho(den, dm0)
        exc, vxc, fxc = eval_xc_eff_list(self.xc, rho)
        out = ni.make_hessian_setup_becke(den, dm0, self.mol.aoslice_by_atom(), exc, vxc, fxc, g.becke, g.coords,
                                          g.quadrature_weights, g.atm_quad_split, g.hardness, grid_fixed_keys=False)
        return out["de_xc_skeleton"], out["vmat_deriv1_grid"], ni, den, fxc

    def compute_f1mo_s1mo(self, f1bra_el, vmat_deriv1=None):
        """compute_dimless_cphf_rhs's f1mo and s1mo (rscf.rs:125-164): C^T (f1bra_el + h1ao C_o)
        and C^T s1ao C_o, [nmo, nocc, 3, natm]."""
        c = self.mo_coeff
        nao, nmo = c.shape
        nocc = int((self.mo_occ > 0).sum())
        nb = 3 * self.natm
        h1 = hcore_deriv1(self.mol, self.device)
        s1 = ovlp_deriv1(self.mol, self.device)
        f1bra = fortran(f1bra_el).clone()
        s1bra = empty_f((nao, nocc, 3, self.natm), self.device)
        # (h1ao + vmat_deriv1) C_o per (x, A): batched over the 3 natm matrices
        _gemm(nao, nocc, nao, h1, nao, False, c, nao, True, f1bra, nao, 1.0, 1.0, nb,
              (nao * nao, 0, 0, 0, nao * nocc, 0), self.device)
        if vmat_deriv1 is not None:
            v1 = fortran(vmat_deriv1)
            _gemm(nao, nocc, nao, v1, nao, False, c, nao, True, f1bra, nao, 1.0, 1.0, nb,
                  (nao * nao, 0, 0, 0, nao * nocc, 0), self.device)
        _gemm(nao, nocc, nao, s1, nao, False, c, nao, True, s1bra, nao, 1.0, 0.0, nb,
              (nao * nao, 0, 0, 0, nao * nocc, 0), self.device)
        f1mo = empty_f((nmo, nocc, 3, self.natm), self.device)
        s1mo = empty_f((nmo, nocc, 3, self.natm), self.device)
        # C^T x: A[m, k] = C[k, m] (k-major), batched
        _gemm(nmo, nocc, nao, c, nao, True, f1bra, nao, True, f1mo, nmo, 1.0, 0.0, nb,
              (0, 0, nao * nocc, 0, nmo * nocc, 0), self.device)
        _gemm(nmo, nocc, nao, c, nao, True, s1bra, nao, True, s1mo, nmo, 1.0, 0.0, nb,
              (0, 0, nao * nocc, 0, nmo * nocc, 0), self.device)
        return f1mo, s1mo

    def make_hess(self, precision="fp64", stats=None):
        """rscf.rs:481-505: skeleton (nuclear, core, RI-JK) + overlap + CPHF; parts in self.result."""
        occ, en = self.mo_occ, self.mo_energy
        r, tm = self.result, self.timing
        tm.clear()
        t0 = time.perf_counter()

        def lap(name):
            nonlocal t0
            torch.cuda.synchronize(self.device)
            t1 = time.perf_counter()
            tm[name] = t1 - t0
            t0 = t1
        dm0 = weighted_density(self.mo_coeff, occ)
        dme0 = weighted_density(self.mo_coeff, occ * en)
        r["de_skeleton_nuc"] = nuc_repl_hess(self.mol, self.device)
        lap("nuc")
        r["de_skeleton_hcore"] = hcore_hess(self.mol, dm0)
        lap("hcore")
        de_jk, f1bra_el = self._skeleton_rijk()
        r["de_skeleton_rijk"] = de_jk
        lap("rijk_skeleton")
        r["de_ovlp"] = ovlp_hess(self.mol, dme0)
        lap("ovlp")
        vmat1 = ni = None
        if self.xc:
            r["de_skeleton_xc"], vmat1, ni, den, fxc = self._skeleton_xc(dm0)
            lap("xc_skeleton")
        f1mo, s1mo = self.compute_f1mo_s1mo(f1bra_el, vmat1)
        lap("f1mo_s1mo")
        # the RI-JK response as the resident (nao nocc)^2 matrix, built once for the CPHF solve
        # (the per-call algebra beyond the memory budget)
        self.rijk.set_response_path(self.config.get("rijk_response", "resident"),
                                    int(self.config.get("rijk_resident_budget", 0)))
        self.rijk.make_response_preparation(self.mo_coeff, occ)
        cfg = {k: v for k, v in self.config.items() if k.startswith("cphf") or k == "level_shift"}
        drv = RHessCPHF(self.mo_coeff, occ, en, cfg)
        drv.add_el(self.rijk)
        if ni is not None:
            drv.add_numint(ni, den, fxc)
        r["de_cphf"] = drv.make_cphf_hess(f1mo, s1mo, precision=precision, stats=stats)
        lap("cphf")
        tot = None
        for k in ("de_skeleton_nuc", "de_skeleton_hcore", "de_skeleton_rijk", "de_skeleton_xc", "de_ovlp", "de_cphf"):
            if k in r:
                tot = r[k].clone() if tot is None else tot.add_(r[k])
        r["de_tot"] = tot
        return tot


def _mo_transform(mol, c, nocc, f1bra, h1=None, vmat=None, s1=None, device=None):
    """C^T (f1bra + (h1 + vmat) C_o) and C^T s1 C_o per (x, A), [nmo, nocc, 3, natm] (f1bra is
    updated in place with the C_o products)."""
    nao, nmo = c.shape
    natm = len(mol.atm)
    nb = 3 * natm
    st = (nao * nao, 0, 0, 0, nao * nocc, 0)
    for m in (h1, vmat):
        if m is not None:
            _gemm(nao, nocc, nao, fortran(m), nao, False, c, nao, True, f1bra, nao, 1.0, 1.0, nb, st, device)
    s1bra = empty_f((nao, nocc, 3, natm), device)
    _gemm(nao, nocc, nao, s1, nao, False, c, nao, True, s1bra, nao, 1.0, 0.0, nb, st, device)
    f1mo = empty_f((nmo, nocc, 3, natm), device)
    s1mo = empty_f((nmo, nocc, 3, natm), device)
    stt = (0, 0, nao * nocc, 0, nmo * nocc, 0)
    _gemm(nmo, nocc, nao, c, nao, True, f1bra, nao, True, f1mo, nmo, 1.0, 0.0, nb, stt, device)
    _gemm(nmo, nocc, nao, c, nao, True, s1bra, nao, True, s1mo, nmo, 1.0, 0.0, nb, stt, device)
    return f1mo, s1mo


class UHessSCF:
    """UHessSCF (uscf.rs:32-540) for UHF or UKS with the RI-JK electron interaction: nuclear
    repulsion, core Hamiltonian and overlap on the summed (energy-weighted) densities
    (UHessHcore, UHessOvlp: ovlp.rs:128-131), the UHF RI-JK skeleton (hess_u.rs:107-261), for UKS
    the grid component with the Becke grid shift (hess_uks_becke.rs), and the unrestricted CPHF
    part; every orbital argument an (alpha, beta) pair."""

    def __init__(self, mol, aux, mo_coeff, mo_occ, mo_energy, config=None, rijk=None, xc=None, grids=None):
        self.mol, self.aux = mol, aux
        self.mo_coeff = tuple(fortran(c) for c in mo_coeff)
        self.device = self.mo_coeff[0].device
        host = lambda x: np.asarray(x.cpu() if isinstance(x, torch.Tensor) else x, dtype=np.float64)
        self.mo_occ = tuple(host(o) for o in mo_occ)
        self.mo_energy = tuple(host(e) for e in mo_energy)
        self.config = dict(config or {})
        self.xc, self.grids = (list(xc) if xc else None), grids
        if self.xc and grids is None:
            raise ValueError("a Kohn-Sham Hessian needs the SCF's grids")
        exx = xc_list_info(self.xc)[1] if self.xc else 1.0
        if rijk is None:
            j3c = mol.int3c2e_s2ij(aux, device=self.device)
            j2c = aux.int2c2e("plain", device=self.device)
            rijk = UHessRIJK.from_j3c(j3c, j2c, 1.0, exx, threshold=1e-14, policy="cd")
        self.rijk = rijk
        self.result = {}
        self.timing = {}

    @property
    def natm(self):
        return len(self.mol.atm)

    def _total(self, weights):
        """sum over both channels of sum_i w_i C_i C_i^T (one product over the joined columns)."""
        c = torch.cat(self.mo_coeff, dim=1)
        return weighted_density(c, np.concatenate(weights))

    def make_hess(self, precision="fp64", stats=None):
        """uscf.rs:514-540: skeleton (nuclear, core, RI-JK, grid) + overlap + CPHF."""
        from .numint import NIMatmul
        from .xc import eval_xc_eff_list
        mol, aux, dev = self.mol, self.aux, self.device
        occ, en = self.mo_occ, self.mo_energy
        r, tm = self.result, self.timing
        tm.clear()
        t0 = time.perf_counter()

        def lap(name):
            nonlocal t0
            torch.cuda.synchronize(dev)
            t1 = time.perf_counter()
            tm[name] = t1 - t0
            t0 = t1
        dm0 = [weighted_density(c, o) for c, o in zip(self.mo_coeff, occ)]
        r["de_skeleton_nuc"] = nuc_repl_hess(mol, dev)
        lap("nuc")
        r["de_skeleton_hcore"] = hcore_hess(mol, self._total(occ))
        lap("hcore")

        def int3c(kind, p0, p1):
            sh0 = int(np.nonzero(aux.ao_loc == p0)[0][0])
            sh1 = int(np.nonzero(aux.ao_loc == p1)[0][0])
            return mol.int3c2e(aux, kind, shls=(sh0, sh1), device=dev)
        v1, v2, v3 = (aux.int2c2e(k, device=dev) for k in ("ip1", "ipip1", "ip1ip2"))
        de_jk, f1bra = self.rijk.make_skeleton(self.mo_coeff, occ, int3c, v1, v2, v3,
                                               [(int(aux.ao_loc[0]), int(aux.ao_loc[-1]))], mol.aoslice_by_atom(),
                                               aux.aoslice_by_atom())
        r["de_skeleton_rijk"] = de_jk
        lap("rijk_skeleton")
        r["de_ovlp"] = ovlp_hess(mol, self._total([o * e for o, e in zip(occ, en)]))
        lap("ovlp")
        vmat = (None, None)
        ni = None
        if self.xc:
            g = self.grids
            if g.weights is None:
                g.build()
            den = xc_list_info(self.xc)[0]
            ni = NIMatmul(mol.eval_gto(g.coords, 3), g.weights)
            rho = ni.setup_rho(den, tuple(dm0))
            exc, vxc, fxc = eval_xc_eff_list(self.xc, rho)
            out = ni.make_hessian_setup_becke(den, tuple(dm0), mol.aoslice_by_atom(), exc, vxc, fxc, g.becke,
                                              g.coords, g.quadrature_weights, g.atm_quad_split, g.hardness,
                                              grid_fixed_keys=False)
            r["de_skeleton_xc"] = out["de_xc_skeleton"]
            vmat = (out["vmat_deriv1_grid_a"], out["vmat_deriv1_grid_b"])
            lap("xc_skeleton")
        h1 = hcore_deriv1(mol, dev)
        s1 = ovlp_deriv1(mol, dev)
        f1mo, s1mo = [], []
        for sp in range(2):
            no = int((occ[sp] > 0).sum())
            f, s_ = _mo_transform(mol, self.mo_coeff[sp], no, fortran(f1bra[sp]).clone(), h1, vmat[sp], s1, dev)
            f1mo.append(f)
            s1mo.append(s_)
        lap("f1mo_s1mo")
        self.rijk.set_response_path(self.config.get("rijk_response", "resident"),
                                    int(self.config.get("rijk_resident_budget", 0)))
        self.rijk.make_response_preparation(self.mo_coeff, occ)
        cfg = {k: v for k, v in self.config.items() if k.startswith("cphf") or k == "level_shift"}
        drv = UHessCPHF(self.mo_coeff, occ, en, cfg)
        drv.add_el(self.rijk)
        if ni is not None:
            drv.add_numint(ni, den, fxc)
        r["de_cphf"] = drv.make_cphf_hess(f1mo, s1mo, precision=precision, stats=stats)
        lap("cphf")
        tot = None
        for k in ("de_skeleton_nuc", "de_skeleton_hcore", "de_skeleton_rijk", "de_skeleton_xc", "de_ovlp", "de_cphf"):
            if k in r:
                tot = r[k].clone() if tot is None else tot.add_(r[k])
        r["de_tot"] = tot
        return tot
