"""make_lcao: drop-in for hqc.pbc.lcao.make_lcao (hqc/pbc/lcao.py:11-104).

lcao(xp) -> (mo_coeff (n_ao, n_mo), bands (n_mo,) [Ry], E [Ry, without V_pp]),
the 3-tuple of hqc/pbc/lcao.py:77-78.  search_method defaults to 'newton' as
in the reference.  xp may also be (B, n, 3): the B geometries run together
and every output gains a leading batch axis.  gamma=False gives lcao(xp, kpt)
(hqc/pbc/lcao.py:80-94) with complex mo_coeff.
"""
import torch

from .solver import make_solver


def make_lcao(n, L, rs, basis, rcut=24, tol=1e-7, max_cycle=100, grid_length=0.12, diis=True,
              diis_space=8, diis_start_cycle=1, diis_damp=0.0, use_jit=True, dft=False,
              xc="lda,vwn", smearing=False, smearing_method="fermi", smearing_sigma=0.0,
              search_method="newton", search_cycle=100, search_tol=1e-7, gamma=True,
              precision="fp64", ovlp_cut=0.0, tuning=None):
    solver = make_solver(n, L, rs, basis, rcut=rcut, tol=tol, max_cycle=max_cycle,
                         grid_length=grid_length, diis=diis, diis_space=diis_space,
                         diis_start_cycle=diis_start_cycle, diis_damp=diis_damp,
                         use_jit=use_jit, dft=dft, xc=xc, smearing=smearing,
                         smearing_method=smearing_method, smearing_sigma=smearing_sigma,
                         search_method=search_method, search_cycle=search_cycle,
                         search_tol=search_tol, gamma=gamma, precision=precision,
                         ovlp_cut=ovlp_cut, tuning=tuning)

    def lcao(xp, kpt=None):
        out = solver.batch(xp, kpt)
        mo, bands, E = out[0], out[2], out[3]
        if torch.as_tensor(xp).dim() == 2:
            return mo[0], bands[0], E[0]
        return mo, bands, E

    lcao.solver = solver
    return lcao
