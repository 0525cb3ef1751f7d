"""Periodic (PBC) Gamma-point GPW solver: make_solver, make_lcao, make_pes."""
from .lcao import make_lcao
from .pes import make_pes
from .potential import potential_energy_pp
from .solver import make_solver

__all__ = ["make_lcao", "make_pes", "make_solver", "potential_energy_pp"]
