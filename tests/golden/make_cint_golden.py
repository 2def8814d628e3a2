"""Writes tests/golden/cint_blocks.npz: oracle (oracle/cint.py) values of int3c2e / int2c2e
shell blocks of every kind for the synthetic basis of tests/cint_basis.py (the integral engine's
GPU tests compare against these; the oracle is too slow to run inside them).

Run from the repository root: python tests/golden/make_cint_golden.py"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "rstsr-showcase-hessian_amd")]

from hesship.cint import make_env  # noqa: E402  (pure Python: no device needed)
from oracle import cint as oc  # noqa: E402
from tests import cint_basis as cb  # noqa: E402


def main():
    mol = make_env(cb.ATOMS, cb.AO)
    aux = make_env(cb.ATOMS, cb.AUX)
    sa, sc = cb.shells_of(*mol), cb.shells_of(*aux)
    out = {}
    for kind, maxl in cb.KINDS3.items():
        for n, (a, b, c) in enumerate(cb.blocks3(sa, sc, maxl)):
            out[f"3c_{kind}_{n}"] = oc.block3c([a[1], b[1], c[1]], [(s[2], s[3], s[4]) for s in (a, b, c)], kind)
            out[f"3c_{kind}_{n}_at"] = np.array([a[5], b[5], c[5]])
        print(kind, n + 1, flush=True)
    for kind, maxl in cb.KINDS2.items():
        for n, (a, c) in enumerate(cb.blocks2(sc, maxl)):
            out[f"2c_{kind}_{n}"] = oc.block2c([a[1], c[1]], [(s[2], s[3], s[4]) for s in (a, c)], kind)
            out[f"2c_{kind}_{n}_at"] = np.array([a[5], c[5]])
        print("2c", kind, n + 1, flush=True)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "cint_blocks.npz"), **out)


if __name__ == "__main__":
    main()
