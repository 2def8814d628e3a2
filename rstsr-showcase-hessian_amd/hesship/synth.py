"""Seeded synthetic inputs for the RI-JK response workload (NumPy, vectorised).

Generator: SplitMix64 -> 53-bit uniform -> Box-Muller.  Output k of stream `seed` is
mix(seed + (k+1) * 0x9E3779B97F4A7C15); normals pair up outputs (2m, 2m+1) as
(r cos(2 pi u2), r sin(2 pi u2)) with r = sqrt(-2 ln(1 - u1)).

Workload "RI-JK response, synthetic" (naux=2400, nao=400, nocc=40, nprop=90 by default):
  B[mu,nu,P] = B[nu,mu,P] = z * exp(-|mu-nu|/25), z ~ N(0,1) drawn in packed-upper column-major
               order (P slowest)                                               seed 0x5EED0001
  C          = Q of QR(N(0,1) [nao, nao]) with diag(R) > 0                     seed 0x5EED0002
  U          = N(0,1) [nmo, nocc, nprop];  bra = C U                           seed 0x5EED0003
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
SEED_B, SEED_C, SEED_U = 0x5EED0001, 0x5EED0002, 0x5EED0003


def splitmix64(seed, start, count):
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, start, count):
    return (splitmix64(seed, start, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, start, count):
    """Normals [start, start+count) of stream `seed`."""
    p0 = start // 2
    p1 = (start + count + 1) // 2
    u = uniform(seed, 2 * p0, 2 * (p1 - p0)).reshape(-1, 2)
    r = np.sqrt(-2.0 * np.log(1.0 - u[:, 0]))
    th = 2.0 * np.pi * u[:, 1]
    z = np.empty((p1 - p0, 2))
    z[:, 0] = r * np.cos(th)
    z[:, 1] = r * np.sin(th)
    z = z.reshape(-1)
    off = start - 2 * p0
    return z[off:off + count]


def cderi(nao, naux, seed=SEED_B, chunk_aux=64):
    """Packed cderi [nao_tp, naux] (float64, Fortran order)."""
    ntp = nao * (nao + 1) // 2
    mu, nu = np.triu_indices(nao)
    order = np.argsort(nu * (nu + 1) // 2 + mu)
    decay = np.exp(-np.abs(mu - nu)[order] / 25.0)
    out = np.empty((ntp, naux), order="F")
    for p0 in range(0, naux, chunk_aux):
        p1 = min(naux, p0 + chunk_aux)
        z = normal(seed, p0 * ntp, (p1 - p0) * ntp).reshape(p1 - p0, ntp)
        out[:, p0:p1] = (z * decay[None, :]).T
    return out


def mo_coeff(nao, seed=SEED_C):
    z = normal(seed, 0, nao * nao).reshape(nao, nao).T  # column-major fill
    q, r = np.linalg.qr(z)
    s = np.sign(np.diag(r))
    s[s == 0] = 1.0
    return np.asfortranarray(q * s[None, :])


def mo1(nmo, nocc, nprop, seed=SEED_U):
    return normal(seed, 0, nmo * nocc * nprop).reshape(nprop, nocc, nmo).T.copy()


def rijk_workload(nao=400, naux=2400, nocc=40, nprop=90):
    """(cderi [nao_tp,naux], C [nao,nao], mo_occ [nao], bra [nao,nocc,nprop]) as float64 numpy."""
    B = cderi(nao, naux)
    Cm = mo_coeff(nao)
    occ = np.zeros(nao)
    occ[:nocc] = 2.0
    U = mo1(nao, nocc, nprop)
    bra = np.einsum("um,mia->uia", Cm, U, optimize=True)
    return B, Cm, occ, bra


def water_cluster(nmol, box_ang=7.0, seed=0, min_oo_ang=2.7):
    """Seeded synthetic (H2O)_n: O-H 0.96 A, H-O-H 104.5 deg, random orientations, oxygens
    uniform in a cubic box of box_ang with every O-O distance >= min_oo_ang.  Returns (charges
    [3n], coordinates [3n, 3] in bohr), each water as O, H, H."""
    rng = np.random.default_rng(seed)
    oxy = []
    while len(oxy) < nmol:
        p = rng.uniform(0.0, box_ang, 3)
        if all(np.linalg.norm(p - q) >= min_oo_ang for q in oxy):
            oxy.append(p)
    half = np.deg2rad(104.5) / 2
    h_local = 0.96 * np.array([[np.sin(half), 0.0, np.cos(half)], [-np.sin(half), 0.0, np.cos(half)]])
    coords, charges = [], []
    for p in oxy:
        q, r = np.linalg.qr(rng.standard_normal((3, 3)))
        q = q * np.sign(np.diag(r))[None, :]
        coords += [p, p + h_local[0] @ q.T, p + h_local[1] @ q.T]
        charges += [8, 1, 1]
    return np.array(charges), np.array(coords) / 0.52917721092
