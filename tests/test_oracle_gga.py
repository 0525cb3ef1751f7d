"""Oracle pins for the PBE (GGA) functional and its grid discretisation: closed
limits of the functional, derivatives against finite differences, the spectral
gradient against an analytic one, variational consistency of v_xc, and SCF
forces against finite differences of the free energy (CPU only)."""
import numpy as np
import pytest

from oracle import cpu_ref as R


def _kF(rho):
    return (3 * np.pi ** 2 * rho) ** (1 / 3)


def test_lda_limit():
    """sigma = 0: exchange is Slater's, correlation is PW92, which fits the same
    uniform-gas data as VWN (agreement to a few 1e-4 Ha over 0.5 <= r_s <= 6)."""
    rs = np.linspace(0.5, 6.0, 23)
    rho = 3 / (4 * np.pi * rs ** 3)
    ex = R.xc_eps_v(rho, "lda,")[0]
    assert np.allclose(R.pbe_x(rho, 0 * rho) / rho, ex, rtol=1e-14, atol=0)
    ec_pw = R.pbe_c(rho, 0 * rho) / rho
    ec_vwn = R.xc_eps_v(rho, "lda,vwn")[0] - ex
    assert np.max(np.abs(ec_pw - ec_vwn)) < 5e-4, np.max(np.abs(ec_pw - ec_vwn))
    assert np.all(np.abs(ec_pw - ec_vwn) / np.abs(ec_vwn) < 1e-2)


def test_exchange_enhancement_limits():
    """F_x(s) = 1 + mu s^2 + O(s^4) with mu = beta pi^2/3 = 0.21951; F_x -> 1 + kappa."""
    rho = np.array([0.02, 0.3, 2.0])
    exl = rho * R.xc_eps_v(rho, "lda,")[0]
    for s_val, check in [(1e-3, "small"), (1e5, "large")]:
        sigma = (2 * _kF(rho) * rho * s_val) ** 2
        Fx = R.pbe_x(rho, sigma) / exl
        if check == "small":
            assert np.allclose((Fx - 1) / s_val ** 2, 0.2195149727645171, rtol=1e-5)
        else:
            assert np.allclose(Fx, 1.804, rtol=1e-9)


def test_correlation_gradient_limits():
    """H = beta t^2 + O(t^4) with beta = 0.066725; eps_c + H -> 0 as t -> infinity."""
    rho = np.array([0.02, 0.3, 2.0])
    ec = R.pbe_c(rho, 0 * rho) / rho
    ks = np.sqrt(4 * _kF(rho) / np.pi)
    for t in (1e-4, 1e5):
        sigma = (2 * ks * rho * t) ** 2
        H = R.pbe_c(rho, sigma) / rho - ec
        if t < 1:
            assert np.allclose(H / t ** 2, 0.06672455060314922, rtol=1e-5)
        else:
            assert np.all(np.abs(ec + H) < 1e-9 * np.abs(ec))


def test_derivatives_vs_finite_differences():
    rng = np.random.default_rng(3)
    rho = rng.uniform(1e-3, 3.0, 50)
    sigma = rng.uniform(0, 4.0, 50) * rho ** (8 / 3)
    e, vr, vs = R.pbe_derivs(rho, sigma)
    f = R.pbe_energy_density

    def d4(fun, x, h):    # fourth-order central difference
        return (-fun(x + 2 * h) + 8 * fun(x + h) - 8 * fun(x - h) + fun(x - 2 * h)) / (12 * h)
    fr = d4(lambda r: f(r, sigma), rho, 1e-3 * rho)
    fs = d4(lambda sg: f(rho, sg), sigma, 1e-3 * sigma)
    assert np.allclose(vr, fr, rtol=1e-9, atol=1e-13)
    assert np.allclose(vs, fs, rtol=1e-8, atol=1e-13)
    e0, vr0, vs0 = R.pbe_derivs(np.array([1e-13, 0.0]), np.array([1e-20, 0.0]))
    assert not np.any(e0) and not np.any(vr0) and not np.any(vs0)


def test_spectral_gradient_of_periodic_gaussian():
    s = R.System(1, 3.0, 1.0, "gth-szv", grid_length=0.12)
    a, R0 = 2.5, np.array([0.7, 1.9, 2.2])
    f = np.zeros(s.N)
    gf = np.zeros((3, s.N))
    for t in np.array(np.meshgrid(*[np.arange(-3, 4)] * 3, indexing="ij")).reshape(3, -1).T:
        d = s.R - R0 + t * s.L
        e = np.exp(-a * (d ** 2).sum(1))
        f += e
        gf += -2 * a * d.T * e
    assert np.allclose(R.spectral_grad(s, f), gf, atol=1e-9 * np.abs(gf).max())


def test_gga_potential_is_variational_derivative():
    s = R.System(2, 2.0, 1.3, "gth-dzv", grid_length=0.3)
    xp = np.random.default_rng(1).uniform(0, s.L, (2, 3))
    phi = R.ao_values(s, xp)
    rho = (phi ** 2).sum(0) * 0.7 + 0.01
    E0, v = R.xc_gga(s, rho)
    for r0 in (0, 17, s.N // 2, s.N - 5):
        d = 1e-5 * rho[r0]
        rp, rm = rho.copy(), rho.copy()
        rp[r0] += d
        rm[r0] -= d
        fd = (R.xc_gga(s, rp)[0] - R.xc_gga(s, rm)[0]) / (2 * d)
        assert abs(fd - s.dV * v[r0]) < 1e-8 * abs(s.dV * v[r0]) + 1e-13, (r0, fd, s.dV * v[r0])


def test_gga_name_rules():
    with pytest.raises(ValueError):
        R.Solver(2, 2.0, 1.3, "gth-dzv", dft=True, xc="pbe,vwn")
    assert R.is_gga("pbe,pbe") and R.is_gga("gga_x_pbe,gga_c_pbe") and not R.is_gga("lda,vwn")


def test_gga_scf_forces_match_finite_differences():
    n, rs = 3, 1.5
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    xp = np.random.default_rng(8).uniform(0, L, (n, 3))
    sv = R.Solver(n, L, rs, "gth-dzv", grid_length=0.35, dft=True, xc="pbe,pbe", tol=1e-12,
                  max_cycle=400, smearing=True, smearing_sigma=0.1, search_tol=1e-14,
                  search_cycle=200)
    out = sv.solve(xp)
    assert out[8]
    F = R.forces(sv, xp)
    h = 1e-4
    for c in range(3):
        xq = xp.copy(); xq[1, c] += h / rs
        xm = xp.copy(); xm[1, c] -= h / rs
        fd = -(R.free_energy(sv, xq) - R.free_energy(sv, xm)) / (2 * h)
        assert abs(F[1, c] - fd) < 1e-7, (c, F[1, c], fd)


@pytest.mark.parametrize("xc", ["lda,vwn", "pbe,pbe"])
def test_kpoint_forces_match_finite_differences(xc):
    """Forces of the free energy at a fixed k-point (complex Bloch AOs) against
    central differences."""
    n, rs = 3, 1.5
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    rng = np.random.default_rng(9)
    xp = rng.uniform(0, L, (n, 3))
    kpt = rng.uniform(-np.pi / L / rs, np.pi / L / rs, 3)
    sv = R.Solver(n, L, rs, "gth-dzv", grid_length=0.35, dft=True, xc=xc, tol=1e-12,
                  max_cycle=400, smearing=True, smearing_sigma=0.1, search_tol=1e-14,
                  search_cycle=200, gamma=False)
    out = sv.solve(xp, kpt)
    assert out[8]
    F = R.forces(sv, xp)
    h = 1e-4
    for c in range(3):
        xq = xp.copy(); xq[2, c] += h / rs
        xm = xp.copy(); xm[2, c] -= h / rs
        fd = -(R.free_energy(sv, xq, kpt=kpt) - R.free_energy(sv, xm, kpt=kpt)) / (2 * h)
        assert abs(F[2, c] - fd) < 1e-7, (c, F[2, c], fd)
