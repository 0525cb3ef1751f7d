"""Oracle pins for the basis, grid and GTO tables (CPU only).

Each check is chosen so that a plausible transcription mistake fails it:
normalisation by the closed-form self-overlap, the odd-grid rule by Python's
half-even rounding, the image-summed tables against a brute-force 3-D image
sum, and AO periodicity."""
import math

import numpy as np
import pytest

from oracle import cpu_ref as R


def _radial_self_overlap(l, a, c):
    # int_0^inf (sum_i c_i r^l e^{-a_i r^2})^2 r^2 dr, closed form
    A = a[:, None] + a[None, :]
    return np.sum(np.outer(c, c) * math.gamma(l + 1.5) / (2 * A ** (l + 1.5)))


@pytest.mark.parametrize("basis", ["gth-szv", "gth-dzv", "gth-dzvp"])
def test_normalisation_closed_form(basis):
    sets = R.normalize_sets(R.read_basis_sets(basis))
    for quant, tab in sets:
        lo = int(quant[1])
        col = 1
        for l in range(lo, int(quant[2]) + 1):
            for _ in range(int(quant[4 + l - lo])):
                assert abs(_radial_self_overlap(l, tab[:, 0], tab[:, col]) - 1.0) < 1e-12
                col += 1


def test_basis_structure():
    b = R.AtomBasis("gth-dzvp")
    assert b.n_ao_atom == 5 and list(b.comp) == [-1, -1, 0, 1, 2]
    assert np.allclose(b.alpha, [8.3744350009, 1.8058681460, 0.4852528328, 0.1658236932, 0.727])
    # p coefficients live only on the p primitive, s only on the s set
    assert np.all(b.coef[4, :2] == 0) and np.all(b.coef[:4, 2:] == 0)
    assert R.AtomBasis("gth-szv").n_ao_atom == 1 and R.AtomBasis("gth-dzv").n_ao_atom == 2
    with pytest.raises(FileNotFoundError):
        R.AtomBasis("gth-nonexistent")


def test_product_basis_matches_oracle():
    from hqc_amd.basis import HBasis
    for basis in ("gth-szv", "gth-dzv", "gth-dzvp"):
        h, o = HBasis(basis), R.AtomBasis(basis)
        assert np.allclose(h.alpha, o.alpha, rtol=0, atol=0)
        assert np.allclose(h.coef, o.coef, rtol=1e-15, atol=1e-16)
        assert list(h.comp) == list(o.comp)
    with pytest.raises(FileNotFoundError):
        HBasis("gth-nonexistent")


def test_grid_rule_half_even():
    # L/grid_length/2 = 12.5 exactly -> Python round gives 12 -> g = 25
    s = R.System(1, 5.0, 1.0, "gth-szv", grid_length=0.2)
    assert s.g == 25
    s = R.System(1, 5.4, 1.0, "gth-szv", grid_length=0.2)   # 13.5 -> 14 -> 29
    assert s.g == 29
    n = 54
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    s2 = R.System(n, L, 1.25, "gth-dzv", grid_length=0.16)
    assert s2.g == 49 and abs(s2.L - 7.6162) < 1e-4
    assert len(s2.lattice) == 123


def test_lattice_counts():
    for n, rs, count in ((8, 1.25, 895), (128, 1.25, 57)):
        L = (4 / 3 * np.pi * n) ** (1 / 3)
        s = R.System(n, L, rs, "gth-szv", grid_length=1.0)
        assert len(s.lattice) == count


def test_VG_closed_form():
    s = R.System(2, 3.0, 1.0, "gth-szv", grid_length=0.5)
    g = s.g
    k = np.fft.fftfreq(g, 1.0 / g)
    i, j, l = 1, g - 2, 3
    G2 = (2 * np.pi / s.L) ** 2 * (k[i] ** 2 + k[j] ** 2 + k[l] ** 2)
    assert np.isclose(s.VG[i, j, l], 4 * np.pi / s.L ** 3 / G2, rtol=1e-14)
    assert s.VG[0, 0, 0] == 0.0


def test_tables_bruteforce_images():
    n, rs = 3, 1.3
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    s = R.System(n, L, rs, "gth-dzvp", grid_length=0.4)
    xp = np.random.default_rng(0).uniform(0, L * rs, (n, 3))
    phi = R.ao_values(s, xp)
    b = s.basis
    t = int(24 // s.L)
    imgs = np.arange(-t, t + 1) * s.L
    T3 = np.array(np.meshgrid(imgs, imgs, imgs, indexing="ij")).reshape(3, -1).T
    rng = np.random.default_rng(1)
    for r in rng.integers(0, s.N, 5):
        pos = s.R[r]
        for A in range(n):
            for o in range(b.n_ao_atom):
                d = pos - xp[A] - T3                        # all cube images
                rad = np.exp(-b.alpha[:, None] * (d ** 2).sum(1)[None, :])
                ang = np.ones(len(T3)) if b.comp[o] < 0 else d[:, b.comp[o]]
                val = np.sum(b.coef[:, o][:, None] * rad * ang[None, :])
                assert np.isclose(phi[A * b.n_ao_atom + o, r], val, rtol=1e-12, atol=1e-14)


def test_ao_periodicity():
    n, rs = 2, 1.5
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    s = R.System(n, L, rs, "gth-dzvp", grid_length=0.4)
    xp = np.random.default_rng(3).uniform(0, L * rs, (n, 3))
    # shifting an atom by a lattice vector leaves its AOs on the grid unchanged
    # up to the truncation of the image cube (negligible for these exponents)
    xq = xp.copy()
    xq[0, 1] += s.L
    a, b = R.ao_values(s, xp), R.ao_values(s, xq)
    assert np.allclose(a, b, atol=1e-10)
