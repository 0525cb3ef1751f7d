"""The C SCF driver (hqc_scf_create / hqc_scf_run, called through ctypes as a
host without Python would) against the Python engine's FP64 make_solver on the
same geometries, and against the oracle on one of them.  The driver takes its own
orthogonaliser (Jacobi eigenpairs of S instead of rocSOLVER's) and its own DIIS
solve, so the two runs agree to the SCF tolerance, not bit for bit."""
import numpy as np
import pytest
import torch

from oracle import cpu_ref as R

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _geoms(n, B, seed):
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(0, L, (n, 3)) for _ in range(B)]), L


def _driver(sol, rs, kw):
    from hqc_amd import _lib as K
    c = sol.cell
    smear = kw.get("smearing", False)
    return K.ScfDriver(c.n, c.L, c.g, c.basis.alpha, c.basis.coef, c.basis.comp, c.nl,
                       rcut=c.rcut, xc_id=sol.engine.xc_id, smearing=smear,
                       smearing_method=kw.get("smearing_method", "fermi"),
                       sigma=kw.get("smearing_sigma", 0.0),
                       search=kw.get("search_method", "bisect"),
                       search_cycle=kw.get("search_cycle", 100),
                       search_tol=kw.get("search_tol", 1e-7), tol=kw["tol"],
                       max_cycle=kw.get("max_cycle", 100))


CASES = [
    # name, n, rs, basis, grid_length, B, kwargs
    ("dzv_fermi_n8", 8, 1.25, "gth-dzv", 0.3, 3,
     dict(smearing=True, smearing_sigma=0.02, tol=1e-9, search_tol=1e-12, search_cycle=200)),
    ("szv_aufbau_n8", 8, 1.25, "gth-szv", 0.3, 2, dict(tol=1e-9)),
    ("dzv_gauss_n4", 4, 1.5, "gth-dzv", 0.25, 2,
     dict(smearing=True, smearing_method="gauss", smearing_sigma=0.05, tol=1e-9,
          search_tol=1e-12, search_cycle=200, max_cycle=200)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_c_driver_matches_engine(case):
    from hqc_amd.pbc import make_solver
    name, n, rs, basis, gl, B, kw = case
    print(f"[C SCF driver] {name}", flush=True)
    xp, L = _geoms(n, B, 11)
    sol = make_solver(n, L, rs, basis, grid_length=gl, dft=True, **kw)
    ref = sol.batch(xp)
    nit_ref = sol.last["nit"].cpu().numpy()
    drv = _driver(sol, rs, kw)
    out = drv.run(torch.tensor(xp * rs, device=DEV))
    drv.close()
    conv = out["converged"].cpu().numpy()
    assert conv.tolist() == ref[8].cpu().numpy().astype(bool).tolist()
    print("n_iter C", out["n_iter"].tolist(), "engine", nit_ref.tolist(),
          "dE", (out["E"] - ref[3]).abs().max().item(), flush=True)
    assert torch.allclose(out["E"], ref[3], atol=1e-8, rtol=0), (out["E"], ref[3])
    assert torch.allclose(out["bands"], ref[2], atol=1e-7, rtol=0)
    for k, j in (("Ki", 4), ("Vep", 5), ("Vee", 6)):
        assert torch.allclose(out[k], ref[j], atol=1e-7, rtol=0), k
    assert torch.allclose(out["Se"], ref[7], atol=1e-7, rtol=0)
    assert torch.allclose(out["dm"], ref[1], atol=1e-6, rtol=0)
    assert out["mo"].shape == (B, sol.cell.nao, sol.cell.nao) and torch.isfinite(out["mo"]).all()
    assert np.all(np.abs(out["n_iter"].cpu().numpy() - nit_ref) <= 2), \
        (out["n_iter"].tolist(), nit_ref.tolist())


def test_c_driver_matches_oracle():
    """One geometry of the 8 H dzv Fermi case against the oracle (the reference
    algorithm on the CPU): energy, bands and density matrix."""
    from hqc_amd.pbc import make_solver
    n, rs, basis, gl = 8, 1.25, "gth-dzv", 0.3
    kw = dict(smearing=True, smearing_sigma=0.02, tol=1e-10, search_tol=1e-12,
              search_cycle=200)
    print("[C SCF driver] oracle", flush=True)
    xp, L = _geoms(n, 1, 5)
    sol = make_solver(n, L, rs, basis, grid_length=gl, dft=True, **kw)
    drv = _driver(sol, rs, kw)
    out = drv.run(torch.tensor(xp * rs, device=DEV))
    drv.close()
    ref = R.Solver(n, L, rs, basis, grid_length=gl, dft=True, **kw).solve(xp[0])
    assert bool(out["converged"][0]) == bool(ref[8])
    assert abs(out["E"][0].item() - ref[3]) < 1e-8, (out["E"][0].item(), ref[3])
    assert np.allclose(out["bands"][0].cpu().numpy(), ref[2], atol=1e-7)
    assert np.allclose(out["dm"][0].cpu().numpy(), ref[1], atol=1e-6)


def test_c_driver_headline_workload():
    """54 H gth-dzv at 49^3, LDA, Fermi 10^4 K (the benchmark's system; the mu search
    run to 1e-12 so both runs take the same iteration counts): 4 geometries against
    the Python engine."""
    from hqc_amd.pbc import make_solver
    n, rs, basis, gl, B = 54, 1.25, "gth-dzv", 0.16, 4
    sigma = 1.0 / (157888.088922572 / 1e4) / 2.0          # 1/(2 beta), T = 10^4 K
    kw = dict(smearing=True, smearing_sigma=sigma, tol=1e-8, search_tol=1e-12,
              search_cycle=200)
    print("[C SCF driver] 54 H dzv 49^3", flush=True)
    xp, L = _geoms(n, B, 0)
    sol = make_solver(n, L, rs, basis, grid_length=gl, dft=True, **kw)
    assert sol.cell.g == 49
    ref = sol.batch(xp)
    drv = _driver(sol, rs, kw)
    out = drv.run(torch.tensor(xp * rs, device=DEV))
    drv.close()
    print("n_iter C", out["n_iter"].tolist(), "engine", sol.last["nit"].tolist(), flush=True)
    assert out["converged"].tolist() == ref[8].bool().tolist()
    assert torch.allclose(out["E"], ref[3], atol=4e-8 * n, rtol=0), (out["E"], ref[3])
    assert torch.allclose(out["bands"], ref[2], atol=1e-6, rtol=0)
    assert torch.allclose(out["dm"], ref[1], atol=1e-5, rtol=0)


FORCE_CASES = [
    ("dzv_fermi_n8", 8, 1.25, "gth-dzv", 0.3, 3,
     dict(smearing=True, smearing_sigma=0.02, tol=1e-9, search_tol=1e-12, search_cycle=200)),
    ("szv_aufbau_n8", 8, 1.25, "gth-szv", 0.3, 2, dict(tol=1e-9)),
    ("dzv_54H_49", 54, 1.25, "gth-dzv", 0.16, 3,
     dict(smearing=True, smearing_sigma=1.0 / (157888.088922572 / 1e4) / 2.0, tol=1e-8,
          search_tol=1e-12, search_cycle=200)),
]


@pytest.mark.parametrize("case", FORCE_CASES, ids=[c[0] for c in FORCE_CASES])
def test_c_driver_forces_match_pes(case):
    """hqc_scf_forces against make_pes(...).energy_and_forces on the same geometries:
    total energy (with V_pp) and forces (Pulay, grid adjoint, v_loc, Ewald)."""
    from hqc_amd.pbc import make_pes
    name, n, rs, basis, gl, B, kw = case
    print(f"[C SCF forces] {name}", flush=True)
    xp, L = _geoms(n, B, 3)
    pes = make_pes(n, L, rs, basis, grid_length=gl, dft=True, **kw)
    E_ref, F_ref = pes.energy_and_forces(xp)
    drv = _driver(pes.solver, rs, kw)
    E, F, conv, nit = drv.forces(torch.tensor(xp * rs, device=DEV))
    drv.close()
    print("  dE", (E - E_ref).abs().max().item(), "dF", (F - F_ref).abs().max().item(),
          "n_iter", nit.tolist(), pes.solver.last["nit"].tolist(), flush=True)
    assert conv.tolist() == pes.solver.last["conv"].bool().tolist()
    assert torch.allclose(E, E_ref, atol=1e-8 * n, rtol=0), (E, E_ref)
    assert torch.allclose(F, F_ref, atol=1e-7, rtol=0)


def test_c_driver_forces_refuse_gauss():
    from hqc_amd import _lib as K
    from hqc_amd.pbc import make_solver
    n, rs = 4, 1.5
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    kw = dict(smearing=True, smearing_method="gauss", smearing_sigma=0.05, tol=1e-8)
    sol = make_solver(n, L, rs, "gth-dzv", grid_length=0.3, dft=True, **kw)
    drv = _driver(sol, rs, kw)
    with pytest.raises(RuntimeError, match="bad argument"):
        drv.forces(torch.zeros(1, n, 3, dtype=torch.float64, device=DEV))
    drv.close()
