"""Oracle: dimensionless CPHF driver pieces (RHF).  Test infrastructure only.

Restates src/hessian/rscf.rs (ajz34/rstsr-showcase-hessian):
  compute_dimless_cphf_rhs  104-182 (final rhs assembly 173-178)
  response_mo               210-220
  response_dimless_cphf     237-263
  solve_dimless_cphf        278-300
  finalize_cphf             328-353
  get_cphf_hess             371-399
and the configuration defaults of HessSCFConfig (rscf.rs:14-18).
"""
import numpy as np

from . import krylov

DEFAULTS = dict(level_shift=0.0, cphf_tol=1e-8, cphf_max_cycle=42, cphf_max_space=14, cphf_lindep=1e-14)

# The reference's literal factors, named so that the tests can plant a wrong one (monkeypatch) and
# show that the mathematical pins catch it.
RHS_OCC = -0.5                        # rscf.rs:178   rhs[occ] = -0.5 s1mo[occ]
HESS_R = (4.0, -4.0, -2.0)            # rscf.rs:387-391  f1.mo1, s1.mo1.e_occ, s1oo.mo_e1
HESS_U = (2.0, -2.0, -1.0)            # uscf.rs:409-440  the same terms per spin


def split(mo_occ, mo_energy):
    occ = np.asarray(mo_occ) > 0
    nocc = int(occ.sum())
    eocc = np.asarray(mo_energy)[occ]
    evir = np.asarray(mo_energy)[~occ]
    return nocc, eocc, evir


def dimless_rhs(f1mo, s1mo, mo_occ, mo_energy, level_shift=0.0):
    """rscf.rs:170-178: rhs[vir] = -(f1mo - s1mo e_occ)/(e_a - e_i + shift); rhs[occ] = -s1mo/2."""
    nocc, eocc, evir = split(mo_occ, mo_energy)
    e_ai = evir[:, None] - eocc[None, :] + level_shift
    tail = (1,) * (f1mo.ndim - 2)
    b1 = f1mo - s1mo * eocc.reshape((1, -1) + tail)
    rhs = np.zeros_like(f1mo)
    rhs[nocc:] = -b1[nocc:] / e_ai.reshape(e_ai.shape + tail)
    rhs[:nocc] = RHS_OCC * s1mo[:nocc]
    return rhs


def response_mo(mo_coeff, response_bras, mo1):
    """rscf.rs:210-220: resp = sum_el C^T el.get_response_bra(C mo1)."""
    shape = mo1.shape
    m = mo1.reshape(shape[0], shape[1], -1)
    ubra = np.einsum("um,mia->uia", mo_coeff, m)
    resp = np.zeros_like(m)
    for rb in response_bras:
        resp += np.einsum("um,uia->mia", mo_coeff, rb(ubra).reshape(ubra.shape))
    return resp.reshape(shape)


def response_dimless(mo_coeff, mo_occ, mo_energy, response_bras, mo1, level_shift=0.0):
    """rscf.rs:237-263: resp[vir] = (resp - shift mo1)[vir] / (e_ai + shift); resp[occ] = 0."""
    nocc, eocc, evir = split(mo_occ, mo_energy)
    e_ai = (evir[:, None] - eocc[None, :] + level_shift)
    resp = response_mo(mo_coeff, response_bras, mo1)
    if level_shift != 0.0:
        resp = resp - level_shift * mo1
    tail = (1,) * (mo1.ndim - 2)
    resp[nocc:] /= e_ai.reshape(e_ai.shape + tail)
    resp[:nocc] = 0.0
    return resp


def solve_dimless(mo_coeff, mo_occ, mo_energy, response_bras, rhs, cfg=None, stats=None):
    """rscf.rs:278-300: Krylov on the flattened [nmo*nocc, nprop] system."""
    cfg = dict(DEFAULTS, **(cfg or {}))
    shape = rhs.shape
    nmo, nocc = shape[:2]

    def aop(x):
        y = response_dimless(mo_coeff, mo_occ, mo_energy, response_bras,
                             x.reshape(nmo, nocc, -1), cfg["level_shift"])
        return y.reshape(nmo * nocc, -1)

    x = krylov.krylov_block(aop, rhs.reshape(nmo * nocc, -1), None, cfg["cphf_tol"],
                            cfg["cphf_max_cycle"], cfg["cphf_max_space"], cfg["cphf_lindep"], stats=stats)
    return x.reshape(shape)


def finalize(mo_coeff, mo_occ, mo_energy, response_bras, f1mo, s1mo, mo1):
    """rscf.rs:328-353: one more response, mo1[vir] = -b1mo[vir]/e_ai, mo_e1 = b1mo[occ] + mo1[occ] e_ij."""
    nocc, eocc, evir = split(mo_occ, mo_energy)
    tail = (1,) * (mo1.ndim - 2)
    e_ai = (evir[:, None] - eocc[None, :]).reshape((len(evir), nocc) + tail)
    e_ij = (eocc[:, None] - eocc[None, :]).reshape((nocc, nocc) + tail)
    b1 = f1mo - s1mo * eocc.reshape((1, -1) + tail) + response_mo(mo_coeff, response_bras, mo1)
    mo1 = mo1.copy()
    mo1[nocc:] = -b1[nocc:] / e_ai
    mo_e1 = mo_e1_of(b1, mo1, e_ij, nocc)
    return mo1, mo_e1


def mo_e1_of(b1, mo1, e_ij, nocc):
    """rscf.rs:350: mo_e1 = b1mo[occ] + mo1[occ] * (e_i - e_j)."""
    return b1[:nocc] + mo1[:nocc] * e_ij


def cphf_hess(f1mo, s1mo, mo1, mo_e1, mo_occ, mo_energy):
    """rscf.rs:371-399: de[s,t,B,A] (B <= A, then mirrored with the 3x3 transpose).

    f1mo, s1mo, mo1: [nmo, nocc, 3, natm]; mo_e1: [nocc, nocc, 3, natm].
    """
    nocc, eocc, _ = split(mo_occ, mo_energy)
    natm = f1mo.shape[3]
    de = np.zeros((3, 3, natm, natm))
    s1oo = s1mo[:nocc]
    for A in range(natm):
        for B in range(A + 1):
            # f1mo[.., .., None, .., A] * mo1[.., .., .., None, B] -> [nmo, nocc, 3(s from mo1), 3(t from f1mo)]
            # summed over axes 0,1 gives de_BA[s, t] with s <-> mo1's xyz, t <-> f1mo's xyz.
            c_f, c_s, c_e = HESS_R
            blk = c_f * np.einsum("mit,mis->st", f1mo[:, :, :, A], mo1[:, :, :, B])
            blk += c_s * np.einsum("mit,mis,i->st", s1mo[:, :, :, A], mo1[:, :, :, B], eocc)
            blk += c_e * np.einsum("jit,jis->st", s1oo[:, :, :, A], mo_e1[:, :, :, B])
            de[:, :, B, A] += blk
        for B in range(A):
            de[:, :, A, B] += de[:, :, B, A].T
    return de


# ---------------------------------------------------------------------------------------------
# Unrestricted driver: src/hessian/uscf.rs (every argument a pair (alpha, beta))
#   compute_dimless_cphf_rhs 78-182 (rhs assembly 165-174), response_mo 203-218,
#   response_dimless_cphf 229-265, solve_dimless_cphf 278-349 (spin blocks stacked per column,
#   pack_flattened / unpack_flattened 292-316), finalize_cphf 369-407, get_cphf_hess 409-440.

def u_dimless_rhs(f1mo, s1mo, mo_occ, mo_energy, level_shift=0.0):
    """uscf.rs:165-174: the restricted formula per spin."""
    return [dimless_rhs(f1mo[s], s1mo[s], mo_occ[s], mo_energy[s], level_shift) for s in range(2)]


def u_response_mo(mo_coeff, response_bras, mo1):
    """uscf.rs:203-218: resp_s = sum_el C_s^T el.get_response_bra([C_a mo1_a, C_b mo1_b])[s]."""
    shapes = [m.shape for m in mo1]
    ms = [m.reshape(m.shape[0], m.shape[1], -1) for m in mo1]
    ubra = [np.einsum("um,mia->uia", mo_coeff[s], ms[s]) for s in range(2)]
    resp = [np.zeros_like(m) for m in ms]
    for rb in response_bras:
        r = rb(ubra)
        for s in range(2):
            resp[s] += np.einsum("um,uia->mia", mo_coeff[s], np.asarray(r[s]).reshape(ubra[s].shape))
    return [r.reshape(sh) for r, sh in zip(resp, shapes)]


def u_response_dimless(mo_coeff, mo_occ, mo_energy, response_bras, mo1, level_shift=0.0):
    """uscf.rs:229-265: the restricted scaling per spin."""
    resp = u_response_mo(mo_coeff, response_bras, mo1)
    out = []
    for s in range(2):
        nocc, eocc, evir = split(mo_occ[s], mo_energy[s])
        e_ai = evir[:, None] - eocc[None, :] + level_shift
        tail = (1,) * (mo1[s].ndim - 2)
        r = resp[s] - level_shift * mo1[s] if level_shift != 0.0 else resp[s]
        r[nocc:] /= e_ai.reshape(e_ai.shape + tail)
        r[:nocc] = 0.0
        out.append(r)
    return out


def u_solve_dimless(mo_coeff, mo_occ, mo_energy, response_bras, rhs, cfg=None, stats=None):
    """uscf.rs:278-349: one Krylov solve over [nmo_a nocc_a + nmo_b nocc_b, nprop]."""
    cfg = dict(DEFAULTS, **(cfg or {}))
    shapes = [r.shape for r in rhs]
    dims = [(sh[0], sh[1]) for sh in shapes]
    split_at = dims[0][0] * dims[0][1]

    def unpack(x):
        return [x[:split_at].reshape(dims[0] + (-1,)), x[split_at:].reshape(dims[1] + (-1,))]

    def pack(v):
        return np.concatenate([v[0].reshape(split_at, -1), v[1].reshape(dims[1][0] * dims[1][1], -1)], axis=0)

    def aop(x):
        return pack(u_response_dimless(mo_coeff, mo_occ, mo_energy, response_bras, unpack(x), cfg["level_shift"]))

    x = krylov.krylov_block(aop, pack([r.reshape(d + (-1,)) for r, d in zip(rhs, dims)]), None, cfg["cphf_tol"],
                            cfg["cphf_max_cycle"], cfg["cphf_max_space"], cfg["cphf_lindep"], stats=stats)
    return [v.reshape(sh) for v, sh in zip(unpack(x), shapes)]


def u_finalize(mo_coeff, mo_occ, mo_energy, response_bras, f1mo, s1mo, mo1):
    """uscf.rs:369-407: the restricted finalisation per spin with the two-spin response."""
    last = u_response_mo(mo_coeff, response_bras, mo1)
    mo1_out, mo_e1 = [], []
    for s in range(2):
        nocc, eocc, evir = split(mo_occ[s], mo_energy[s])
        tail = (1,) * (mo1[s].ndim - 2)
        e_ai = (evir[:, None] - eocc[None, :]).reshape((len(evir), nocc) + tail)
        e_ij = (eocc[:, None] - eocc[None, :]).reshape((nocc, nocc) + tail)
        b1 = f1mo[s] - s1mo[s] * eocc.reshape((1, -1) + tail) + last[s]
        m = mo1[s].copy()
        m[nocc:] = -b1[nocc:] / e_ai
        mo1_out.append(m)
        mo_e1.append(mo_e1_of(b1, m, e_ij, nocc))
    return mo1_out, mo_e1


def u_cphf_hess(f1mo, s1mo, mo1, mo_e1, mo_occ, mo_energy):
    """uscf.rs:409-440: the restricted assembly with factors 2, -2, -1 summed over spins."""
    natm = f1mo[0].shape[3]
    de = np.zeros((3, 3, natm, natm))
    for A in range(natm):
        for B in range(A + 1):
            for s in range(2):
                nocc, eocc, _ = split(mo_occ[s], mo_energy[s])
                c_f, c_s, c_e = HESS_U
                blk = c_f * np.einsum("mit,mis->st", f1mo[s][:, :, :, A], mo1[s][:, :, :, B])
                blk += c_s * np.einsum("mit,mis,i->st", s1mo[s][:, :, :, A], mo1[s][:, :, :, B], eocc)
                blk += c_e * np.einsum("jit,jis->st", s1mo[s][:nocc, :, :, A], mo_e1[s][:, :, :, B])
                de[:, :, B, A] += blk
        for B in range(A):
            de[:, :, A, B] += de[:, :, B, A].T
    return de
