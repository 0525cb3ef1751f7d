"""CPU oracle for the periodic GPW SCF path -- TEST INFRASTRUCTURE ONLY.

`oracle.cpu_ref` restates the reference algorithm (hqc/pbc/solver.py, scf.py,
gto.py, hqc/tools/smearing.py, excor.py, hqc/pbc/potential.py,
hqc/basis/parse.py) in NumPy/SciPy FP64.  Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may import
it, and only as the checker; the product package `hqc_amd` never does.
"""
