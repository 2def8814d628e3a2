"""A small analytic molecular model for the DFT setup tests (CPU, torch FP64).

Atoms carry contracted-free Gaussian functions phi(r) = (a0 + a.d) exp(-alpha |d|^2), d = r - R_A,
whose Cartesian derivatives to third order are closed forms; moving an atom moves its functions.
The exchange-correlation energy E = sum_g w_g eps(rho, sigma, tau) of a smooth analytic energy
density (not a physical functional) at a fixed density matrix is then a differentiable function
of the nuclear coordinates, and its autograd derivatives are what the setup passes compute.
"""
import itertools

import numpy as np
import torch

from oracle import dft_setup as ods
from oracle import xceff as oxc

PAIRS = list(itertools.combinations_with_replacement(range(3), 2))
TRIPLES = list(itertools.combinations_with_replacement(range(3), 3))


class Model:
    def __init__(self, natm=3, nbas=(3, 2, 3), ng=64, seed=0):
        rng = np.random.default_rng(seed)
        self.natm = natm
        self.R = rng.uniform(-1.2, 1.2, (natm, 3))
        self.owner = np.concatenate([[A] * nbas[A] for A in range(natm)])
        self.nao = len(self.owner)
        self.alpha = rng.uniform(0.5, 1.1, self.nao)
        self.a0 = rng.uniform(0.5, 1.5, self.nao)
        self.a = rng.standard_normal((self.nao, 3)) * 0.6
        self.grid = rng.uniform(-2.5, 2.5, (ng, 3))
        self.weights = rng.uniform(0.05, 0.2, ng)
        off = np.concatenate([[0], np.cumsum(nbas)])
        self.aoslices = [(int(off[A]), int(off[A + 1])) for A in range(natm)]

    # -- AO values and r-derivatives to third order, [ng, nao, 20]; differentiable in R --
    def ao(self, R, grid=None):
        R = torch.as_tensor(R, dtype=torch.float64)
        g = torch.as_tensor(self.grid) if grid is None else grid
        al = torch.as_tensor(self.alpha)
        a0 = torch.as_tensor(self.a0)
        a = torch.as_tensor(self.a)
        d = g[:, None, :] - R[torch.as_tensor(self.owner)][None, :, :]  # [ng, nao, 3]
        G = torch.exp(-al * (d * d).sum(-1))
        P = a0 + (d * a).sum(-1)
        dl = lambda i, j: 1.0 if i == j else 0.0
        G1 = [-2.0 * al * d[..., i] * G for i in range(3)]
        G2 = {(i, j): (4.0 * al ** 2 * d[..., i] * d[..., j] - 2.0 * al * dl(i, j)) * G
              for i in range(3) for j in range(3)}
        G3 = {(i, j, k): (-8.0 * al ** 3 * d[..., i] * d[..., j] * d[..., k]
                          + 4.0 * al ** 2 * (dl(i, j) * d[..., k] + dl(i, k) * d[..., j] + dl(j, k) * d[..., i])) * G
              for i in range(3) for j in range(3) for k in range(3)}
        out = [P * G]
        out += [a[:, i] * G + P * G1[i] for i in range(3)]
        out += [a[:, i] * G1[j] + a[:, j] * G1[i] + P * G2[i, j] for i, j in PAIRS]
        out += [a[:, i] * G2[j, k] + a[:, j] * G2[i, k] + a[:, k] * G2[i, j] + P * G3[i, j, k] for i, j, k in TRIPLES]
        return torch.stack(out, dim=-1)

    def ao_np(self):
        return self.ao(self.R).numpy()

    def density_matrix(self, nocc, seed, scale=2.0):
        C = np.random.default_rng(seed).standard_normal((self.nao, nocc)) * 0.5
        return scale * C @ C.T


def rho_vars(ao, D, den):
    """rho [ng, nvar] in the rho_t variables (rho, grad rho, tau) from AO values (torch)."""
    phi = ao[..., 0]
    dphi = ao[..., 1:4]
    out = [torch.einsum("gu,uv,gv->g", phi, D, phi)]
    if den in ("sigma", "tau"):
        out += [2.0 * torch.einsum("gu,uv,gv->g", dphi[..., t], D, phi) for t in range(3)]
    if den == "tau":
        out += [0.5 * sum(torch.einsum("gu,uv,gv->g", dphi[..., t], D, dphi[..., t]) for t in range(3))]
    return torch.stack(out, dim=-1)


def eps_unpol(rho, sigma=None, tau=None):
    e = -0.6 * rho ** 2 / (0.5 + rho)
    if sigma is not None:
        e = e + 0.04 * sigma / (1.0 + rho) ** 2
    if tau is not None:
        e = e + 0.03 * tau ** 2 / (1.0 + tau + rho) + 0.01 * tau * sigma / (1.0 + rho)
    return e


def libxc_vars_unpol(r, den):
    v = [r[:, 0]]
    if den in ("sigma", "tau"):
        v.append((r[:, 1:4] ** 2).sum(-1))
    if den == "tau":
        v.append(r[:, 4])
    return v


def energy(model, R, D, den):
    r = rho_vars(model.ao(R), torch.as_tensor(D), den)
    return (torch.as_tensor(model.weights) * eps_unpol(*libxc_vars_unpol(r, den))).sum()


def xc_columns(eps, u):
    """[exc | first | unique second] derivatives of eps(*u) per point (u: list of [ng] tensors)."""
    u = [x.detach().clone().requires_grad_(True) for x in u]
    e = eps(*u)
    d1 = torch.autograd.grad(e.sum(), u, create_graph=True)
    cols = [e.detach()] + [x.detach() for x in d1]
    # second derivatives: a first derivative that is constant in u[q] has no graph to it
    for p, q in oxc.uniq_pairs(len(u)):
        g = torch.autograd.grad(d1[p].sum(), u[q], retain_graph=True, allow_unused=True)[0]
        cols.append(torch.zeros_like(u[q]) if g is None else g.detach())
    return torch.stack(cols, dim=-1).numpy()


def rks_inputs(model, D, den):
    """(rho, vxc_eff, fxc_eff) at the model's geometry through the oracle's transform."""
    r = rho_vars(model.ao(model.R), torch.as_tensor(D), den)
    xc = xc_columns(eps_unpol, libxc_vars_unpol(r, den))
    rn = r.numpy()
    if den == "rho":  # LDA: the raw columns
        return rn, xc[:, 1:2], xc[:, 2:3, None]
    return rn, oxc.transform(rn, xc, den, 1), oxc.transform(rn, xc, den, 2)


# ---- spin-polarised ----

def eps_pol(ra, rb, saa=None, sab=None, sbb=None, ta=None, tb=None):
    e = -0.6 * (ra ** 2 / (0.5 + ra) + rb ** 2 / (0.5 + rb)) + 0.02 * ra * rb / (1.0 + ra + rb)
    if saa is not None:
        e = e + 0.04 * (saa + sbb + 0.5 * sab) / (1.0 + ra + rb) ** 2
    if ta is not None:
        e = e + 0.03 * (ta ** 2 + tb ** 2) / (1.0 + ta + tb + ra + rb) + 0.01 * ta * sbb / (1.0 + ra)
    return e


def libxc_vars_pol(ra, rb, den):
    v = [ra[:, 0], rb[:, 0]]
    if den in ("sigma", "tau"):
        ga, gb = ra[:, 1:4], rb[:, 1:4]
        v += [(ga * ga).sum(-1), (ga * gb).sum(-1), (gb * gb).sum(-1)]
    if den == "tau":
        v += [ra[:, 4], rb[:, 4]]
    return v


def energy_uks(model, R, Da, Db, den):
    ao = model.ao(R)
    ra, rb = rho_vars(ao, torch.as_tensor(Da), den), rho_vars(ao, torch.as_tensor(Db), den)
    return (torch.as_tensor(model.weights) * eps_pol(*libxc_vars_pol(ra, rb, den))).sum()


def uks_inputs(model, Da, Db, den):
    ao = model.ao(model.R)
    ra, rb = rho_vars(ao, torch.as_tensor(Da), den), rho_vars(ao, torch.as_tensor(Db), den)
    xc = xc_columns(eps_pol, libxc_vars_pol(ra, rb, den))
    rho = np.stack([ra.numpy(), rb.numpy()], axis=2)
    return rho, oxc.transform_pol(rho, xc, den, 1), oxc.transform_pol(rho, xc, den, 2)


def ao_for(model, den):
    return model.ao_np()[:, :, :ods.NCOMP_AO[ods.AO_DERIV[den]]]
