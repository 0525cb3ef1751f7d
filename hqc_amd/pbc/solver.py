"""make_solver: drop-in for hqc.pbc.solver.make_solver (hqc/pbc/solver.py:14-978).

Same signature, defaults and units: xp and L in rs units, grid_length and rcut
in Bohr, smearing_sigma and tol in Hartree; the returned solver(xp) gives the
9-tuple (mo_coeff, dm, bands [Ry], E [Ry], Ki [Ry], Vep [Ry], Vee [Ry], Se,
converged) of hqc/pbc/solver.py:767.  Results are FP64 torch tensors on the
GPU.  `use_jit` is accepted and ignored (there is no tracing step).

Beyond the reference: solver.batch(xp) takes (B, n, 3) and runs the B SCFs
together; `precision='mixed'` runs the AO x grid contraction increments on
FP32 MFMA (Gamma only).  gamma=False gives solver(xp, kpt) at one k-point
(hqc/pbc/solver.py:769-837, 901-962); kpt (3,) or (B, 3) in 1/Bohr, so a
Monkhorst-Pack mesh of one geometry is a batch of repeated positions.
"""
import numpy as np
import torch

from .engine import RY, Cell, Engine
from .kengine import KEngine


def _as_batch(xp, n, device):
    x = torch.as_tensor(np.asarray(xp) if not torch.is_tensor(xp) else xp,
                        dtype=torch.float64).to(device)
    single = x.dim() == 2
    if single:
        x = x.unsqueeze(0)
    assert x.shape[1] == n                                   # solver.py:722
    return x.contiguous(), single


class GPWSolver:
    """Callable returned by make_solver."""

    def __init__(self, n, L, rs, basis, rcut=24, tol=1e-7, max_cycle=100, grid_length=0.12,
                 diis=True, diis_space=8, diis_start_cycle=1, diis_damp=0.0, use_jit=True,
                 dft=False, xc="lda,vwn", smearing=False, smearing_method="fermi",
                 smearing_sigma=0.0, search_method="bisect", search_cycle=100,
                 search_tol=1e-7, gamma=True, precision="fp64", device="cuda", ovlp_cut=0.0,
                 tuning=None):
        self.n, self.L, self.rs = n, L, rs
        self.gamma = gamma
        self.cell = Cell(n, L, rs, basis, rcut=rcut, grid_length=grid_length, device=device)
        self.engine = (Engine if gamma else KEngine)(self.cell, dft=dft, xc=xc, smearing=smearing,
                             smearing_method=smearing_method, smearing_sigma=smearing_sigma,
                             search_method=search_method, search_cycle=search_cycle,
                             search_tol=search_tol, tol=tol, max_cycle=max_cycle, diis=diis,
                             diis_space=diis_space, diis_start_cycle=diis_start_cycle,
                             diis_damp=diis_damp, precision=precision, ovlp_cut=ovlp_cut,
                             tuning=tuning)
        self.last = None

    def _kpts(self, kpt, B):
        if self.gamma:
            if kpt is not None:
                raise TypeError("a Gamma-point solver takes no kpt (make it with gamma=False)")
            return None
        if kpt is None:
            raise TypeError("gamma=False: call solver(xp, kpt)")
        k = torch.as_tensor(np.asarray(kpt) if not torch.is_tensor(kpt) else kpt,
                            dtype=torch.float64).to(self.cell.device)
        if k.dim() == 1:
            k = k.unsqueeze(0).expand(B, 3)
        assert k.shape == (B, 3)
        return k.contiguous()

    def run(self, xp, kpt=None, forces=False):
        """Solve a batch; returns (state dict, single flag, xp tensor in rs units).
        forces=True keeps what the force pass reuses (S/T centre derivatives)."""
        x, single = _as_batch(xp, self.n, self.cell.device)
        st = self.engine.solve(x * self.rs, self._kpts(kpt, x.shape[0]), forces=forces)
        self.last = st
        return st, single, x

    def batch(self, xp, kpt=None):
        st, _, _ = self.run(xp, kpt)
        return (st["mo"], st["D"], st["w1"] * RY, st["E"] * RY, st["Ki"] * RY,
                st["Vep"] * RY, st["Vee"] * RY, st["Se"], st["conv"])

    def __call__(self, xp, kpt=None):
        out = self.batch(xp, kpt)
        if torch.as_tensor(xp).dim() == 2:
            return tuple(o[0] for o in out)
        return out


def make_solver(n, L, rs, basis, rcut=24, tol=1e-7, max_cycle=100, grid_length=0.12, diis=True,
                diis_space=8, diis_start_cycle=1, diis_damp=0.0, use_jit=True, dft=False,
                xc="lda,vwn", smearing=False, smearing_method="fermi", smearing_sigma=0.0,
                search_method="bisect", search_cycle=100, search_tol=1e-7, gamma=True,
                precision="fp64", ovlp_cut=0.0, tuning=None):
    """hqc/pbc/solver.py:14 signature (search_method default 'bisect' there)."""
    return GPWSolver(n, L, rs, basis, rcut=rcut, tol=tol, max_cycle=max_cycle,
                     grid_length=grid_length, diis=diis, diis_space=diis_space,
                     diis_start_cycle=diis_start_cycle, diis_damp=diis_damp, use_jit=use_jit,
                     dft=dft, xc=xc, smearing=smearing, smearing_method=smearing_method,
                     smearing_sigma=smearing_sigma, search_method=search_method,
                     search_cycle=search_cycle, search_tol=search_tol, gamma=gamma,
                     precision=precision, ovlp_cut=ovlp_cut, tuning=tuning)
