"""Becke partition weights and their nuclear derivatives (reference:
src/numint_matmul/becke_partition.rs) on the device.

Grid points belong to their generating atoms and move with them; w = wquad P_A / Z with Becke's
cell functions (J. Chem. Phys. 88, 2547 (1988)); dw / ddw are total derivatives with respect to
the atom positions.  Output layouts follow the reference's flat C-order buffers: w [ng],
dw [natm, 3, ng], ddw [natm, 3, natm, 3, ng], c [nset], dc [natm, 3, nset],
ddc [natm, 3, natm, 3, nset].
"""
import ctypes as C

import numpy as np
import torch

from ._lib import check, lib
from .tensors import stream_of


class BeckeMolTables:
    """BeckeMolTables (becke_partition.rs:146-205): atom positions [natm, 3] and the antisymmetric
    radii-adjustment table [natm, natm] (entry (A, B) = a_AB), kept on `device`."""

    def __init__(self, atm_coords, adjustment_factor, device):
        R = np.ascontiguousarray(np.asarray(atm_coords, dtype=np.float64).reshape(-1, 3))
        a = np.ascontiguousarray(np.asarray(adjustment_factor, dtype=np.float64))
        self.natm = R.shape[0]
        if a.shape != (self.natm, self.natm):
            raise ValueError("adjustment_factor must be [natm, natm]")
        self.device = torch.device(device)
        h = C.c_void_p()
        check(lib().hh_becke_create(self.natm, R.ctypes.data, a.ctypes.data, stream_of(self.device), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().hh_becke_destroy(h)
            except Exception:
                pass
            self._h = None

    def partition(self, grid_coords, atm_indices, quadrature_weights, hardness=3, deriv=2, by_atom=False,
                  output=(True, True, True), contract_w=None, contract_dw=None, contract_ddw=None):
        """becke_partition_with_tables (becke_partition.rs:259-340).  grid_coords [ng, 3] and
        quadrature_weights [ng] on the device; atm_indices: ByAtom boundaries [natm + 1] when
        by_atom, else the per-grid generating atoms [ng] (values >= natm are padding).
        output: which of (w, dw, ddw) to return; contract_*: device [nset, ng] weights.
        Returns a dict with the produced keys of w, dw, ddw, c, dc, ddc."""
        dev = self.device
        natm = self.natm
        coords = grid_coords.to(device=dev, dtype=torch.float64).contiguous()
        wq = quadrature_weights.to(device=dev, dtype=torch.float64).contiguous()
        ng = coords.shape[0]
        if coords.shape != (ng, 3) or wq.shape != (ng,):
            raise ValueError("grid_coords must be [ng, 3] and quadrature_weights [ng]")
        idx = np.ascontiguousarray(np.asarray(atm_indices, dtype=np.int64))
        if idx.shape != ((natm + 1,) if by_atom else (ng,)):
            raise ValueError("ByAtom atm_indices must have length natm + 1" if by_atom
                             else "ByGrid atm_indices must have length ngrids")
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)
        out = {}
        if output[0]:
            out["w"] = z(ng)
        if output[1] and deriv >= 1:
            out["dw"] = z(natm, 3, ng)
        if output[2] and deriv >= 2:
            out["ddw"] = z(natm, 3, natm, 3, ng)
        cws = []
        for key, cwt, order in (("c", contract_w, 0), ("dc", contract_dw, 1), ("ddc", contract_ddw, 2)):
            if cwt is None or deriv < order:
                cws.append((None, 0))
                continue
            cwt = cwt.to(device=dev, dtype=torch.float64).reshape(-1, ng).contiguous()
            nset = cwt.shape[0]
            cws.append((cwt, nset))
            out[key] = z(*([natm, 3] * order + [nset]))
        ptr = lambda k: out[k].data_ptr() if k in out else None
        cp = lambda i: cws[i][0].data_ptr() if cws[i][0] is not None else None
        check(lib().hh_becke_partition(self._h, coords.data_ptr(), ng, 1 if by_atom else 0,
                                       idx.ctypes.data_as(C.POINTER(C.c_int64)), wq.data_ptr(), int(hardness),
                                       int(deriv), ptr("w"), ptr("dw"), ptr("ddw"), cp(0), cws[0][1], ptr("c"), cp(1),
                                       cws[1][1], ptr("dc"), cp(2), cws[2][1], ptr("ddc")))
        return out
