"""The analytic DFT model of dft_model with an atom-centred grid glued to its atoms and Becke
partition weights (CPU, torch FP64).

Each grid point r_g = R_{A(g)} + delta_g moves with its generating atom A(g); its weight is
w_g = wquad_g P_{A(g)}(r_g) / sum_B P_B(r_g) with Becke's cell functions
P_A = prod_{B != A} (1 - f(nu_AB)) / 2, nu_AB = mu_AB + a_AB (1 - mu_AB^2),
mu_AB = (|r - R_A| - |r - R_B|) / |R_A - R_B|, f = p^hardness, p(x) = 3x/2 - x^3/2
(Becke, J. Chem. Phys. 88, 2547 (1988)), written directly as torch functions of the atom
positions.  The XC energy E(R) = sum_g w_g(R) eps(rho(r_g(R); R)) at a fixed density matrix is
translation invariant; its autograd derivatives are what the Becke-augmented setup computes.
"""
import numpy as np
import torch

from tests import dft_model as dm


class BeckeModel(dm.Model):
    def __init__(self, natm=3, nbas=(3, 2, 3), npa=(22, 18, 26), hardness=3, seed=0):
        super().__init__(natm=natm, nbas=nbas, ng=sum(npa), seed=seed)
        rng = np.random.default_rng(seed + 100)
        self.npa = npa
        self.split = np.concatenate([[0], np.cumsum(npa)]).astype(np.int64)
        self.attr = np.repeat(np.arange(natm), npa)
        ng = int(self.split[-1])
        u = rng.standard_normal((ng, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        self.delta = u * rng.uniform(0.25, 2.2, (ng, 1))
        self.wquad = rng.uniform(0.05, 0.25, ng)
        a = rng.uniform(-0.35, 0.35, (natm, natm))
        self.adjust = np.triu(a, 1) - np.triu(a, 1).T
        self.hardness = hardness
        self.grid = self.coords(self.R)
        self.weights = self.becke_w(torch.tensor(self.R)).numpy()

    def coords(self, R):
        if isinstance(R, torch.Tensor):
            return R[torch.as_tensor(self.attr)] + torch.as_tensor(self.delta)
        return R[self.attr] + self.delta

    def ao(self, R, grid=None):
        R = torch.as_tensor(R, dtype=torch.float64)
        return super().ao(R, self.coords(R) if grid is None else grid)

    def becke_w(self, R):
        R = torch.as_tensor(R, dtype=torch.float64)
        r = self.coords(R)
        natm = self.natm
        d = [torch.linalg.norm(r - R[A], dim=1) for A in range(natm)]
        P = []
        for A in range(natm):
            p = torch.ones_like(d[0])
            for B in range(natm):
                if B == A:
                    continue
                mu = (d[A] - d[B]) / torch.linalg.norm(R[A] - R[B])
                f = mu + float(self.adjust[A, B]) * (1.0 - mu * mu)
                for _ in range(self.hardness):
                    f = 1.5 * f - 0.5 * f ** 3
                p = p * 0.5 * (1.0 - f)
            P.append(p)
        P = torch.stack(P, dim=1)
        Pg = P[torch.arange(len(self.attr)), torch.as_tensor(self.attr)]
        return torch.as_tensor(self.wquad) * Pg / P.sum(dim=1)


def energy(model, R, D, den):
    r = dm.rho_vars(model.ao(R), torch.as_tensor(D), den)
    return (model.becke_w(R) * dm.eps_unpol(*dm.libxc_vars_unpol(r, den))).sum()


def energy_uks(model, R, Da, Db, den):
    ao = model.ao(R)
    ra, rb = dm.rho_vars(ao, torch.as_tensor(Da), den), dm.rho_vars(ao, torch.as_tensor(Db), den)
    return (model.becke_w(R) * dm.eps_pol(*dm.libxc_vars_pol(ra, rb, den))).sum()


def exc_rks(model, D, den):
    """Per-particle energy density at the model geometry (eps / rho[0])."""
    r = dm.rho_vars(model.ao(model.R), torch.as_tensor(D), den)
    return (dm.eps_unpol(*dm.libxc_vars_unpol(r, den)) / r[:, 0]).numpy()


def exc_uks(model, Da, Db, den):
    ao = model.ao(model.R)
    ra, rb = dm.rho_vars(ao, torch.as_tensor(Da), den), dm.rho_vars(ao, torch.as_tensor(Db), den)
    return (dm.eps_pol(*dm.libxc_vars_pol(ra, rb, den)) / (ra[:, 0] + rb[:, 0])).numpy()
