"""Oracle pins for AO values at electron positions and the Slater matrices of
make_slater (hqc/pbc/slater.py): point values must equal the grid AO values at
grid points (Gamma and k), and a brute-force 3-D image sum at random points."""
import numpy as np
import pytest

from oracle import cpu_ref as R


@pytest.mark.parametrize("kpt", [None, np.array([0.2, -0.35, 0.5])])
def test_points_equal_grid_values(kpt):
    s = R.System(3, 2.2, 1.4, "gth-dzvp", grid_length=0.4)
    xp = np.random.default_rng(1).uniform(0, s.L, (3, 3))
    grid = R.ao_values(s, xp, kpt)
    sel = np.random.default_rng(2).choice(s.N, 40, replace=False)
    pts = R.ao_points(s.basis, s.L, xp, s.R[sel], kpt=kpt)
    assert np.allclose(pts.T, grid[:, sel], rtol=1e-13, atol=1e-15)


def test_points_brute_force_image_sum():
    """phi(x) = sum over the cube of images T of the isolated contracted function."""
    b = R.AtomBasis("gth-dzvp")
    L = 4.1
    rng = np.random.default_rng(3)
    xp = rng.uniform(0, L, (2, 3))
    xe = rng.uniform(-L, 2 * L, (5, 3))
    got = R.ao_points(b, L, xp, xe)
    tmax = int(24 // L)
    t1 = np.arange(-tmax, tmax + 1) * L
    T = np.array(np.meshgrid(t1, t1, t1, indexing="ij")).reshape(3, -1).T
    ref = np.zeros_like(got)
    for e in range(5):
        for A in range(2):
            d = xe[e] - xp[A] - T                                    # (nT, 3)
            g = np.exp(-b.alpha[None, :] * (d ** 2).sum(1)[:, None])  # (nT, a)
            for o in range(b.n_ao_atom):
                ang = 1.0 if b.comp[o] < 0 else d[:, b.comp[o]]
                ref[e, A * b.n_ao_atom + o] = ((g * ang[..., None] if b.comp[o] >= 0 else g)
                                               @ b.coef[:, o]).sum()
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-15)


def test_slater_groundstate_shapes_and_orthogonality():
    n, rs = 4, 1.3
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    rng = np.random.default_rng(4)
    xp = rng.uniform(0, L, (n, 3))
    sv = R.Solver(n, L, rs, "gth-dzv", grid_length=0.35, tol=1e-9)
    mo = sv.solve(xp)[0]
    xe = rng.uniform(0, L, (n, 3))
    up, dn = R.slater(n, L, rs, "gth-dzv", xp, xe, mo)
    assert up.shape == (2, 2) and dn.shape == (2, 2)
    # exchanging two same-spin electrons swaps rows
    up2, _ = R.slater(n, L, rs, "gth-dzv", xp, xe[[1, 0, 2, 3]], mo)
    assert np.allclose(up2, up[[1, 0]])


def test_potential_energy_neutral_total_kappa_independent():
    """Protons + electrons: sum_ij q_i q_j psi_ij + (n/2) Madelung (each charge's
    self term) is the Ewald energy of the neutral cell and cannot depend on kappa;
    the like-charge parts reproduce the proton Ewald sum."""
    rng = np.random.default_rng(5)
    n, rs = 8, 1.3
    L = (4 / 3 * np.pi * (n // 2)) ** (1 / 3)
    x = rng.uniform(0, L, (n, 3))
    G = R.ewald_kpoints(15)
    tot = []
    for kappa in (9.0, 9.5, 10.0):      # window where both truncations are < 1e-9
        vpp, vep, vee = R.potential_energy(x, kappa, G, L, rs)
        tot.append(vpp + vep + vee + n * rs / L * R.madelung(kappa, G))
    assert np.allclose(tot, tot[0], rtol=1e-9, atol=1e-12)
    vpp, _, _ = R.potential_energy(x, 10.0, G, L, rs)
    Epp = R.ewald_pp(x[:n // 2], L, rs)                  # Ry, with the Madelung term
    assert np.isclose(vpp / rs ** 2 + (n // 2) * R.madelung(10.0, G) / L / rs, Epp, rtol=1e-12)
