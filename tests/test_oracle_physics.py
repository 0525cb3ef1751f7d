"""Oracle pins for integrals, the two two-electron routes, xc, occupations,
Ewald, SCF invariants and forces (CPU only, a few seconds each)."""
import math

import numpy as np
import pytest

from oracle import cpu_ref as R


def _rand_sys(n, rs, basis, grid_length, seed=0, L=None):
    L = (4 / 3 * np.pi * n) ** (1 / 3) if L is None else L
    s = R.System(n, L, rs, basis, grid_length=grid_length)
    xp = np.random.default_rng(seed).uniform(0, s.L, (n, 3))
    return s, xp


# ------------------------------------------------------------------ integrals
@pytest.mark.parametrize("basis", ["gth-dzv", "gth-dzvp"])
def test_overlap_kinetic_vs_grid_quadrature(basis):
    """Analytic lattice sums (Hermite-derivative route) against an independent
    quadrature of the grid AO values: S = dV sum phi phi, T = 1/2 dV sum grad.grad."""
    s, xp = _rand_sys(3, 1.4, basis, 0.09, seed=2)
    S, T = R.overlap_kinetic(s, xp)
    phi = R.ao_values(s, xp)
    gphi = R.ao_gradients(s, xp)
    Sg = phi @ phi.T * s.dV
    Tg = 0.5 * sum(gp @ gp.T for gp in gphi) * s.dV
    assert np.allclose(S, Sg, atol=1e-9)
    assert np.allclose(T, Tg, atol=1e-7)


def test_isolated_atom_normalised():
    """At large L only T=0 survives: S_ii = 1 (radial norm x solid-harmonic constant)."""
    s = R.System(1, 30.0, 1.0, "gth-dzvp", rcut=24.0, grid_length=3.0)
    S, T = R.overlap_kinetic(s, np.zeros((1, 3)))
    assert np.allclose(np.diag(S), 1.0, atol=1e-12)
    # p functions are orthogonal to s and to each other on one centre
    assert np.allclose(S[2:, :2], 0, atol=1e-14) and np.allclose(S[2:, 2:], np.eye(3), atol=1e-12)
    # kinetic energy of a normalised s primitive is 3a/2
    b = R.AtomBasis("gth-dzv")
    a = b.alpha[3]
    assert np.isclose(T[1, 1], 1.5 * a, rtol=1e-12)


def test_overlap_derivatives_fd():
    s, xp = _rand_sys(3, 1.5, "gth-dzvp", 0.4, seed=1)
    S, T, dS, dT = R.overlap_kinetic(s, xp, deriv=True)
    h = 1e-5
    nloc = s.basis.n_ao_atom
    for c in range(3):
        xq = xp.copy(); xq[0, c] += h
        xm = xp.copy(); xm[0, c] -= h
        Sp, Tp = R.overlap_kinetic(s, xq)
        Sm, Tm = R.overlap_kinetic(s, xm)
        # rows of atom 0, columns of other atoms: first-argument derivative
        assert np.allclose((Sp - Sm)[:nloc, nloc:] / (2 * h), dS[:nloc, nloc:, c], atol=1e-8)
        assert np.allclose((Tp - Tm)[:nloc, nloc:] / (2 * h), dT[:nloc, nloc:, c], atol=1e-8)


# ---------------------------------------------------------- two-electron routes
ROUTES = [
    ("gth-szv", True, False, True),
    ("gth-dzv", True, True, True),
    ("gth-dzvp", False, True, True),
    ("gth-dzv", False, False, False),
    ("gth-dzvp", False, True, False),
]


@pytest.mark.parametrize("basis,gamma,dft,smearing", ROUTES)
def test_literal_rhoG_route_equals_fft_route(basis, gamma, dft, smearing):
    """The reference's rhoG/eris route (restated literally) and the FFT-Poisson
    / pair-FFT restatement give the same SCF on the symmetric odd G-set."""
    n, rs = 3, 1.5
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    rng = np.random.default_rng(5)
    xp = rng.uniform(0, L, (n, 3))
    kpt = None if gamma else rng.uniform(-np.pi / L / rs, np.pi / L / rs, 3)
    out = []
    for lit in (True, False):
        sv = R.Solver(n, L, rs, basis, grid_length=0.35, dft=dft, tol=1e-11, max_cycle=200,
                      smearing=smearing, smearing_sigma=0.1, gamma=gamma, literal=lit,
                      search_tol=1e-13, search_cycle=200)
        out.append(sv.solve(xp, kpt))
    a, b = out
    assert abs(a[3] - b[3]) < 1e-10
    assert np.allclose(a[2], b[2], atol=1e-9)
    assert np.allclose(a[1], b[1], atol=1e-8)


# ------------------------------------------------------------------------ xc
@pytest.mark.parametrize("xc", ["lda,vwn", "lda,", "lda,hl", "lda,bh", "lda,gl"])
def test_vxc_is_derivative_of_rho_eps(xc):
    rho = np.geomspace(1e-4, 10.0, 40)
    eps, v = R.xc_eps_v(rho, xc)
    h = 1e-6 * rho
    ep, _ = R.xc_eps_v(rho + h, xc)
    em, _ = R.xc_eps_v(rho - h, xc)
    fd = ((rho + h) * ep - (rho - h) * em) / (2 * h)
    assert np.allclose(v, fd, rtol=1e-7, atol=1e-10)


def test_lda_exchange_uniform_gas():
    rs = np.array([0.5, 1.0, 2.0, 5.0])
    rho = 3 / (4 * np.pi * rs ** 3)
    eps, _ = R.xc_eps_v(rho, "lda,")
    # Slater/Dirac exchange per electron: -(3/4)(9/(4 pi^2))^(1/3) / rs = -0.458165293 / rs Ha
    assert np.allclose(eps, -0.4581652932831429 / rs, rtol=1e-13)


def test_vwn_high_density_limit():
    """VWN interpolates the Gell-Mann-Brueckner limit eps_c ~ (A/2) ln rs, A/2 = 0.0310907."""
    rs = np.array([1e-6, 2e-6])
    rho = 3 / (4 * np.pi * rs ** 3)
    ec = R.xc_eps_v(rho, "lda,vwn")[0] - R.xc_eps_v(rho, "lda,")[0]
    slope = (ec[1] - ec[0]) / np.log(2.0)
    assert abs(slope - 0.0310907) < 2e-4


def test_unsupported_xc():
    with pytest.raises(ValueError):
        R.xc_eps_v(np.array([1.0]), "pbe,vwn")
    with pytest.raises(NameError):
        R.xc_eps_v(np.array([1.0]), "lda,pz")


# --------------------------------------------------------------- occupations
@pytest.mark.parametrize("method,search", [("fermi", "bisect"), ("fermi", "newton"),
                                           ("gauss", "bisect")])
def test_occupations_conserve_electrons(method, search):
    rng = np.random.default_rng(0)
    w = np.sort(rng.normal(0.3, 0.4, 30))     # mu > 0 and mu < 0 both reached below
    for shift in (0.0, -1.0):
        occ = R.make_occupation(12, 30, True, method, 0.05, search, 100, 1e-9)(w + shift)
        assert abs(occ.sum() - 12) < 1e-6
        # gauss lacks the closed-shell factor 2 (hqc/tools/smearing.py:52): f <= 1
        assert np.all(occ >= 0) and np.all(occ <= (2 if method == "fermi" else 1))


def test_small_sigma_gives_aufbau():
    w = np.sort(np.random.default_rng(1).normal(0, 1, 20))
    occ = R.make_occupation(10, 20, True, "fermi", 1e-4, "bisect", 200, 1e-12)(w)
    assert np.allclose(occ, [2.0] * 5 + [0.0] * 15, atol=1e-12)


def test_entropy_limits():
    assert R.entropy(np.array([2.0, 2.0, 0.0])) == 0.0
    assert np.isclose(R.entropy(np.array([1.0, 1.0])), 2 * 2 * np.log(2.0))


# ---------------------------------------------------------------------- Ewald
ALPHA_SC = 2.837297479480620


def test_ewald_simple_cubic():
    L, rs = 1.7, 1.3
    E = R.ewald_pp(np.zeros((1, 3)), L, rs)
    assert np.isclose(E, -ALPHA_SC / (L * rs), rtol=1e-10)


def test_ewald_bcc_wigner():
    rs = 1.0
    L = (4 * np.pi * 2 / 3) ** (1 / 3)
    xp = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]]) * L
    E = R.ewald_pp(xp, L, rs)
    # bcc Wigner crystal: -0.895929255682 / rs Hartree per particle
    assert np.isclose(E, 2 * 2 * -0.895929255682 / rs, rtol=1e-9)


def test_ewald_kappa_independent():
    n, rs = 10, 1.2
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    xp = np.random.default_rng(2).uniform(0, L, (n, 3))
    # nearest-image real space and |G| <= 15 make the sum kappa-independent only
    # in a window around the reference's kappa = 10
    assert np.isclose(R.ewald_pp(xp, L, rs, kappa=8), R.ewald_pp(xp, L, rs, kappa=10),
                      rtol=1e-9)


def test_ewald_forces_fd():
    n, rs = 5, 1.2
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    xp = np.random.default_rng(3).uniform(0, L, (n, 3))
    _, F = R.ewald_pp(xp, L, rs, forces=True)
    h = 1e-5
    for c in range(3):
        xq = xp.copy(); xq[1, c] += h / rs
        xm = xp.copy(); xm[1, c] -= h / rs
        fd = -(R.ewald_pp(xq, L, rs) - R.ewald_pp(xm, L, rs)) / 2 / (2 * h)
        assert abs(F[1, c] - fd) < 1e-7


# ------------------------------------------------------------------------ SCF
def test_scf_invariants_and_diis_vs_fixed_point():
    n, rs = 4, 1.5
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    xp = np.random.default_rng(1).uniform(0, L, (n, 3))    # a geometry plain fixed point converges on
    kw = dict(grid_length=0.3, dft=True, tol=1e-11, max_cycle=400)
    a = R.Solver(n, L, rs, "gth-dzv", diis=True, **kw)
    oa = a.solve(xp)
    b = R.Solver(n, L, rs, "gth-dzv", diis=False, **kw)
    ob = b.solve(xp)
    assert oa[8] and ob[8]
    assert abs(oa[3] - ob[3]) < 1e-8
    S = a.last["S"]
    D = oa[1]
    assert abs(np.trace(D @ S) - n) < 1e-10                     # electron count
    assert np.allclose(D @ S @ D, 2 * D, atol=1e-8)              # idempotent (aufbau)


def test_smearing_electron_count():
    n, rs = 4, 1.5
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    xp = np.random.default_rng(6).uniform(0, L, (n, 3))
    sv = R.Solver(n, L, rs, "gth-dzv", grid_length=0.3, dft=True, tol=1e-10,
                  smearing=True, smearing_sigma=0.05, search_tol=1e-12)
    out = sv.solve(xp)
    assert abs(np.trace(out[1] @ sv.last["S"]) - n) < 1e-9


# --------------------------------------------------------------------- forces
@pytest.mark.parametrize("dft,basis", [(True, "gth-dzv"), (False, "gth-szv")])
def test_forces_match_finite_differences(dft, basis):
    n, rs = 3, 1.5
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    xp = np.random.default_rng(8).uniform(0, L, (n, 3))
    sv = R.Solver(n, L, rs, basis, grid_length=0.35, dft=dft, tol=1e-12, max_cycle=400,
                  smearing=True, smearing_sigma=0.1, search_tol=1e-14, search_cycle=200)
    sv.solve(xp)
    F = R.forces(sv, xp)
    h = 1e-4
    for c in range(3):
        xq = xp.copy(); xq[0, c] += h / rs
        xm = xp.copy(); xm[0, c] -= h / rs
        fd = -(R.free_energy(sv, xq) - R.free_energy(sv, xm)) / (2 * h)
        assert abs(F[0, c] - fd) < 1e-7, (c, F[0, c], fd)
