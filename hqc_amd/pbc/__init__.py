"""Periodic (PBC) GPW solver at Gamma or one k-point: make_solver, make_lcao,
make_pes (hqc/pbc/solver.py, lcao.py, pes.py), make_pbc_gto / make_slater
(hqc/pbc/gto.py, slater.py), plus make_pes_batch for lists of systems of
different sizes."""
from .batch import make_pes_batch
from .lcao import make_lcao
from .pes import make_pes
from .potential import potential_energy, potential_energy_pp
from .slater import make_pbc_gto, make_slater
from .solver import make_solver

__all__ = ["make_lcao", "make_pbc_gto", "make_pes", "make_pes_batch", "make_slater",
           "make_solver", "potential_energy", "potential_energy_pp"]
