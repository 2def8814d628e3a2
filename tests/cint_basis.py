"""The synthetic basis of the integral-engine tests (no basis-set files exist here): three atoms,
an AO basis with a general contraction and s..f shells, an auxiliary basis with s..g shells, and
the shell-block selection the golden vectors cover."""
import numpy as np

ATOMS = [("X", (0.0, 0.0, 0.0)), ("H", (0.0, 1.43, 1.11)), ("H", (0.31, -1.37, 0.98))]
AO = {"X": [(0, [[11.0, 0.15, 0.0], [2.4, 0.55, 0.0], [0.55, 0.45, 1.0]]), (1, [[3.1, 0.4], [0.7, 0.7]]),
            (2, [[1.1, 1.0]]), (3, [[0.9, 1.0]])],
      "H": [(0, [[4.2, 0.3], [0.65, 0.8]]), (1, [[0.8, 1.0]])]}
AUX = {"X": [(0, [[3.5, 1.0]]), (1, [[2.2, 1.0]]), (2, [[1.6, 1.0]]), (3, [[1.2, 1.0]]), (4, [[0.9, 1.0]])],
       "H": [(0, [[2.0, 1.0]]), (1, [[1.1, 1.0]]), (2, [[0.8, 1.0]])]}

# kind -> largest la + lb + lc covered by the golden blocks (the oracle's cost grows with it)
KINDS3 = {"plain": 4, "ip1": 3, "ip2": 3, "ipip1": 2, "ipvip1": 2, "ip1ip2": 2, "ipip2": 2}
KINDS2 = {"plain": 4, "ip1": 3, "ipip1": 2, "ip1ip2": 2}
MAX_BLOCKS = 36


def shells_of(atm, bas, env, cart=False):
    """(atom, centre, l, exps, coefs, first AO) per contracted shell (libcint AO order)."""
    out = []
    for b in bas:
        A, l, nprim, nctr, pe, pc = (int(v) for v in (b[0], b[1], b[2], b[3], b[5], b[6]))
        cen = np.array(env[atm[A][1]:atm[A][1] + 3])
        ex = np.array(env[pe:pe + nprim])
        for c in range(nctr):
            out.append((A, cen, l, ex, np.array(env[pc + nprim * c:pc + nprim * (c + 1)])))
    ao0 = np.cumsum([0] + [(2 * s[2] + 1) if not cart else (s[2] + 1) * (s[2] + 2) // 2 for s in out])
    return [s + (int(a),) for s, a in zip(out, ao0)]


def blocks3(sa, sc, maxl):
    allb = [(a, b, c) for a in sa for b in sa for c in sc if a[2] + b[2] + c[2] <= maxl]
    step = max(1, -(-len(allb) // MAX_BLOCKS))
    return allb[::step]


def blocks2(sc, maxl):
    allb = [(a, c) for a in sc for c in sc if a[2] + c[2] <= maxl]
    step = max(1, -(-len(allb) // MAX_BLOCKS))
    return allb[::step]
