"""hqc_amd: MI355X-native (gfx950) periodic GPW SCF engine, a drop-in for the
hqc.pbc factories make_lcao / make_pes / make_solver on one GPU."""
__version__ = "0.1.0"
