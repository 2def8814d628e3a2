"""An analytic stand-in for the 2- and 3-centre Coulomb integrals (CPU, FP64) for the RI-JK
skeleton tests.

Every AO and auxiliary function sits on an atom.  The model integrals
    (uv|P) = s_u s_v s_P exp(-a_u |c_u - c_P|^2 - a_v |c_v - c_P|^2 - b |c_u - c_v|^2)
    (P|Q)  = s_P s_Q exp(-g |c_P - c_Q|^2) + l_P delta_PQ
are symmetric in u, v, translation invariant, and (P|Q) is positive definite.  Their derivatives
with respect to the function centres are closed forms, laid out as libcint's derivative integrals
(a derivative on a function is minus the derivative with respect to its centre; the 9-component
second derivatives reshaped to [.., 3, 3] with the first index on the later function).  The
same integrals as torch functions of the atom positions give the energies whose autograd
derivatives the skeleton terms must reproduce.
"""
import numpy as np
import torch


class Model:
    def __init__(self, nbas=(3, 2, 3), naux_per=(4, 3, 5), seed=0):
        rng = np.random.default_rng(seed)
        natm = len(nbas)
        self.natm = natm
        self.R = rng.uniform(-1.0, 1.0, (natm, 3))
        self.ao_atom = np.concatenate([[A] * nbas[A] for A in range(natm)])
        self.aux_atom = np.concatenate([[A] * naux_per[A] for A in range(natm)])
        self.nao, self.naux = len(self.ao_atom), len(self.aux_atom)
        self.a = rng.uniform(0.2, 0.5, self.nao)
        self.s = rng.uniform(0.5, 1.0, self.nao)
        self.sP = rng.uniform(0.5, 1.0, self.naux)
        self.b = 0.15
        self.g = 0.3
        self.lP = rng.uniform(0.5, 1.0, self.naux)
        off = np.concatenate([[0], np.cumsum(nbas)])
        self.aoslices = [(int(off[A]), int(off[A + 1])) for A in range(natm)]
        off = np.concatenate([[0], np.cumsum(naux_per)])
        self.auxslices = [(int(off[A]), int(off[A + 1])) for A in range(natm)]

    # ---- torch functions of the atom positions ----
    def j3c_t(self, R):
        R = torch.as_tensor(R, dtype=torch.float64)
        cu = R[torch.as_tensor(self.ao_atom)]
        cP = R[torch.as_tensor(self.aux_atom)]
        a = torch.as_tensor(self.a)
        dup = ((cu[:, None, :] - cP[None, :, :]) ** 2).sum(-1)        # [u, P]
        duv = ((cu[:, None, :] - cu[None, :, :]) ** 2).sum(-1)        # [u, v]
        E = a[:, None, None] * dup[:, None, :] + a[None, :, None] * dup[None, :, :] + self.b * duv[:, :, None]
        S = torch.as_tensor(self.s)
        return S[:, None, None] * S[None, :, None] * torch.as_tensor(self.sP)[None, None, :] * torch.exp(-E)

    def j3c_centres(self, cu, cv, cP):
        """The 3-centre model with independent centres per index (for derivative checks)."""
        a = torch.as_tensor(self.a)
        dup = ((cu[:, None, :] - cP[None, :, :]) ** 2).sum(-1)
        dvp = ((cv[:, None, :] - cP[None, :, :]) ** 2).sum(-1)
        duv = ((cu[:, None, :] - cv[None, :, :]) ** 2).sum(-1)
        E = a[:, None, None] * dup[:, None, :] + a[None, :, None] * dvp[None, :, :] + self.b * duv[:, :, None]
        S = torch.as_tensor(self.s)
        return S[:, None, None] * S[None, :, None] * torch.as_tensor(self.sP)[None, None, :] * torch.exp(-E)

    def j2c_centres(self, cP, cQ):
        d = ((cP[:, None, :] - cQ[None, :, :]) ** 2).sum(-1)
        sP = torch.as_tensor(self.sP)
        return sP[:, None] * sP[None, :] * torch.exp(-self.g * d)

    def j2c_t(self, R):
        R = torch.as_tensor(R, dtype=torch.float64)
        cP = R[torch.as_tensor(self.aux_atom)]
        d = ((cP[:, None, :] - cP[None, :, :]) ** 2).sum(-1)
        sP = torch.as_tensor(self.sP)
        return sP[:, None] * sP[None, :] * torch.exp(-self.g * d) + torch.diag(torch.as_tensor(self.lP))

    # ---- integrals and closed-form derivative integrals at the model geometry ----
    def integrals(self):
        R = self.R
        cu, cP = R[self.ao_atom], R[self.aux_atom]
        f = self.j3c_t(R).numpy()
        au = self.a[:, None, None, None]
        av = self.a[None, :, None, None]
        du = cu[:, None, None, :] - cP[None, None, :, :]      # c_u - c_P   [u, 1, P, 3]
        dv = cu[None, :, None, :] - cP[None, None, :, :]      # c_v - c_P   [1, v, P, 3]
        duv = cu[:, None, None, :] - cu[None, :, None, :]     # c_u - c_v   [u, v, 1, 3]
        Eu = 2 * au * du + 2 * self.b * duv                   # dE/dc_u
        Ev = 2 * av * dv - 2 * self.b * duv                   # dE/dc_v
        EP = -2 * au * du - 2 * av * dv                       # dE/dc_P
        Eu, Ev, EP = (np.broadcast_to(x, f.shape + (3,)) for x in (Eu, Ev, EP))
        eye = np.eye(3)
        F = f[..., None, None]
        # d2 f / dx_a dy_b = f (dE/dx_a dE/dy_b - d2E/dx_a dy_b); second index b on the earlier function
        second = lambda Ea, Eb, h: F * (Ea[..., :, None] * Eb[..., None, :] - h[..., None, None] * eye)
        hu = np.broadcast_to(2 * (au + self.b), f.shape + (1,))[..., 0]
        huv = np.broadcast_to(-2 * self.b * np.ones_like(au), f.shape + (1,))[..., 0]
        huP = np.broadcast_to(-2 * au, f.shape + (1,))[..., 0]
        hP = np.broadcast_to(2 * (au + av), f.shape + (1,))[..., 0]
        I = dict(j3c=f, ip1=f[..., None] * Eu, ip2=f[..., None] * EP)
        I["ipip1"] = second(Eu, Eu, hu)
        I["ipvip1"] = second(Ev, Eu, huv)          # [.., a (v), b (u)]
        I["ip1ip2"] = second(EP, Eu, huP)          # [.., a (P), b (u)]
        I["ipip2"] = second(EP, EP, hP)
        g = (self.j2c_t(R).numpy() - np.diag(self.lP))
        dPQ = cP[:, None, :] - cP[None, :, :]
        GP = 2 * self.g * dPQ                      # dE/dc_P of exp(-g |c_P - c_Q|^2)
        GQ = -GP
        G = g[..., None, None]
        h2 = 2 * self.g
        I["j2c"] = self.j2c_t(R).numpy()
        I["j2c_ip1"] = g[..., None] * GP
        I["j2c_ipip1"] = G * (GP[..., :, None] * GP[..., None, :] - h2 * eye)
        I["j2c_ip1ip2"] = G * (GQ[..., :, None] * GP[..., None, :] + h2 * eye)
        return I

    # ---- energies / Fock matrices as functions of R (for autograd) ----
    def e_j(self, R, D):
        f, V = self.j3c_t(R), self.j2c_t(R)
        a = torch.einsum("uvP,uv->P", f, torch.as_tensor(D))
        return 0.5 * a @ torch.linalg.solve(V, a)

    def e_k(self, R, C):
        f, V = self.j3c_t(R), self.j2c_t(R)
        t = torch.einsum("uvP,ui,vj->ijP", f, torch.as_tensor(C), torch.as_tensor(C)).reshape(-1, self.naux)
        return 0.5 * (t * torch.linalg.solve(V, t.T).T).sum()

    def fock_j(self, R, D):
        f, V = self.j3c_t(R), self.j2c_t(R)
        a = torch.einsum("uvP,uv->P", f, torch.as_tensor(D))
        return torch.einsum("uvP,P->uv", f, torch.linalg.solve(V, a))

    def fock_k(self, R, C):
        f, V = self.j3c_t(R), self.j2c_t(R)
        T = torch.einsum("ukP,kj->ujP", f, torch.as_tensor(C))
        Tt = torch.linalg.solve(V, T.reshape(-1, self.naux).T).T.reshape(T.shape)
        return torch.einsum("ujP,vjP->uv", T, Tt)

    def orbitals(self, nocc, seed, occ_value=2.0):
        rng = np.random.default_rng(seed)
        C = np.linalg.qr(rng.standard_normal((self.nao, self.nao)))[0]
        occ = np.where(np.arange(self.nao) < nocc, occ_value, 0.0)
        return C, occ
