"""make_pes: drop-in for hqc.pbc.pes.make_pes (hqc/pbc/pes.py:10-154).

pes(xp) -> E + V_pp [Ry] (mode 'default') or the 7-tuple
(E+V_pp, Ki, Vep, Vee, Vpp, Se, converged) (mode 'dev', pes.py:99-115).
search_method defaults to 'bisect' as in the reference.

Additions (not in the reference): pes.energy_and_forces(xp) returns
(E+V_pp [Ry], F [Ha/Bohr]) with F = -dA/dR of the free energy
A = E - sigma*S_e + V_pp (equal to E+V_pp without smearing), and
pes.free_energy(xp) is a torch.autograd-differentiable A [Ry] whose gradient
comes from the hand-written adjoint.  xp may be (B, n, 3) everywhere.  With
gamma=False every call takes kpt as well (pes_kpt, hqc/pbc/pes.py:88-97).
"""
import torch

from .engine import RY
from .forces import FreeEnergy, check_supported, compute_forces
from .potential import ewald_pp
from .solver import _as_batch, make_solver


class PES:
    def __init__(self, solver, kappa, Gmax, mode):
        if mode not in ("default", "dev"):
            raise ValueError(f"Invalid mode: {mode}.")
        self.solver, self.kappa, self.Gmax, self.mode = solver, kappa, Gmax, mode
        self.rs, self.L = solver.rs, solver.L

    def _vpp(self, x):
        return ewald_pp(x, self.L, self.rs, self.kappa, self.Gmax)

    def __call__(self, xp, kpt=None):
        st, single, x = self.solver.run(xp, kpt)
        Vpp = self._vpp(x)
        Etot = st["E"] * RY + Vpp
        if self.mode == "default":
            out = (Etot,)
        else:
            out = (Etot, st["Ki"] * RY, st["Vep"] * RY, st["Vee"] * RY, Vpp, st["Se"], st["conv"])
        if single:
            out = tuple(o[0] for o in out)
        return out[0] if self.mode == "default" else out

    def energy_and_forces(self, xp, kpt=None):
        check_supported(self.solver.engine, kpt is not None)
        st, single, x = self.solver.run(xp, kpt, forces=True)
        F = compute_forces(self.solver.engine, st, self.kappa, self.Gmax)
        Etot = st["E"] * RY + st["Epp"]
        return (Etot[0], F[0]) if single else (Etot, F)

    def free_energy_and_forces(self, x, kpt=None):
        check_supported(self.solver.engine, kpt is not None)
        st, _, x = self.solver.run(x, kpt, forces=True)
        F = compute_forces(self.solver.engine, st, self.kappa, self.Gmax)
        A = st["E"] * RY - RY * self.solver.engine.sigma * st["Se"] + st["Epp"]
        return A, F

    def free_energy(self, xp, kpt=None):
        check_supported(self.solver.engine, kpt is not None)
        x, single = _as_batch(xp.detach() if torch.is_tensor(xp) else xp, self.solver.n,
                              self.solver.cell.device)
        if torch.is_tensor(xp) and xp.requires_grad:
            x = xp if xp.dim() == 3 else xp.unsqueeze(0)
        A = FreeEnergy.apply(x, self, kpt)
        return A[0] if single else A


def make_pes(n, L, rs, basis, rcut=24, tol=1e-7, max_cycle=100, grid_length=0.12, diis=True,
             diis_space=8, diis_start_cycle=1, diis_damp=0.0, use_jit=True, dft=False,
             xc="lda,vwn", smearing=False, smearing_method="fermi", smearing_sigma=0.0,
             search_method="bisect", search_cycle=100, search_tol=1e-7, gamma=True, Gmax=15,
             kappa=10, mode="default", precision="fp64", ovlp_cut=0.0, tuning=None):
    if mode not in ("default", "dev"):
        raise ValueError(f"Invalid mode: {mode}.")
    solver = make_solver(n, L, rs, basis, rcut=rcut, tol=tol, max_cycle=max_cycle,
                         grid_length=grid_length, diis=diis, diis_space=diis_space,
                         diis_start_cycle=diis_start_cycle, diis_damp=diis_damp,
                         use_jit=use_jit, dft=dft, xc=xc, smearing=smearing,
                         smearing_method=smearing_method, smearing_sigma=smearing_sigma,
                         search_method=search_method, search_cycle=search_cycle,
                         search_tol=search_tol, gamma=gamma, precision=precision,
                         ovlp_cut=ovlp_cut, tuning=tuning)
    return PES(solver, kappa, Gmax, mode)
