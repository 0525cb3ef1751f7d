"""Write tests/golden/*.npz: seeded inputs and the oracle's outputs for a few
small cases (regression fixtures of the restatement; the reference ships no
golden vectors and cannot run here).  Run from the repository root:

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import cpu_ref as R  # noqa: E402

CASES = {
    # name: (n, rs, basis, grid_length, kwargs, seed)
    "hf_szv_n8": (8, 1.25, "gth-szv", 0.168, dict(dft=False, tol=1e-10), 7),
    "dft_dzv_n4_fermi": (4, 1.5, "gth-dzv", 0.25,
                         dict(dft=True, smearing=True, smearing_sigma=0.1, tol=1e-10), 7),
    "hf_dzvp_n4_newton": (4, 1.5, "gth-dzvp", 0.3,
                          dict(dft=False, smearing=True, smearing_sigma=0.1,
                               search_method="newton", tol=1e-10), 7),
}


def run(name):
    n, rs, basis, gl, kw, seed = CASES[name]
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    xp = np.random.default_rng(seed).uniform(0, L, (n, 3))
    out = R.Solver(n, L, rs, basis, grid_length=gl, **kw).solve(xp)
    return dict(xp=xp, L=L, rs=rs, mo=out[0], dm=out[1], bands=out[2], E=out[3], Ki=out[4],
                Vep=out[5], Vee=out[6], Se=out[7], conv=out[8], Vpp=R.ewald_pp(xp, L, rs))


if __name__ == "__main__":
    for name in CASES:
        np.savez(os.path.join(HERE, name + ".npz"), **run(name))
        print("wrote", name)
