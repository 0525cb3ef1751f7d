"""The host restatement of the int8-digit density (tests/q8_digits.py) pinned
against what the mathematics fixes: digit expansions reconstruct their value to
2^(e - 7n), enough digits and levels reproduce the FP64 orbital density, and
the kept digits (4 of phi, levels <= 5) stay within the stated error bound."""
import numpy as np
import pytest

from q8_digits import q8_density_digits, q8_digits, q8_exponent


def test_exponent_brackets_the_maximum():
    m = np.array([0.0, 1.0, 0.75, 2.0 ** -40, 3.0e5, 1.0 - 2.0 ** -52])
    e = q8_exponent(m)
    nz = m > 0
    assert (m[nz] < np.ldexp(1.0, e[nz])).all()
    assert (m[nz] >= np.ldexp(1.0, e[nz] - 1)).all()
    assert e[0] == 0


@pytest.mark.parametrize("n", [1, 3, 4, 7])
def test_digits_reconstruct(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((5, 200)) * np.exp2(rng.integers(-30, 10, (5, 1)))
    e = q8_exponent(np.abs(x).max(1))[:, None]
    d = q8_digits(x, e, n)
    rec = sum(np.ldexp(ds, e - 6 - 7 * s) for s, ds in enumerate(d))
    assert (np.abs(rec - x) <= np.ldexp(1.0, e - 7 * n)).all()
    # the leading digit carries the sign and at most 7 bits
    assert (np.abs(d[0]) <= 64).all()


def _system(seed, B=2, nao=40, m=24, ld=96):
    rng = np.random.default_rng(seed)
    phi = rng.standard_normal((B, nao, ld)) * np.exp2(-rng.integers(0, 25, (B, 1, ld)))
    phi[:, :, 3:5] = 0.0
    At = rng.standard_normal((B, nao, m)) * np.exp(-0.3 * np.arange(m))
    At[1] *= 1e-3
    return phi, At


def test_all_digits_reproduce_fp64_density():
    phi, At = _system(1)
    rho = ((np.matmul(At.transpose(0, 2, 1), phi)) ** 2).sum(1)
    r = q8_density_digits(phi, At, 8, top=14)
    assert np.allclose(r, rho, rtol=1e-12, atol=0.0)


def test_kept_digits_within_bound():
    phi, At = _system(2)
    psi = np.matmul(At.transpose(0, 2, 1), phi)
    rho = (psi ** 2).sum(1)
    ns = 4
    r = q8_density_digits(phi, At, ns)
    nao = phi.shape[1]
    d = (nao * np.abs(At).max(1)[:, :, None] * np.abs(phi).max(1)[:, None, :]
         * (2.0 ** (1 - 7 * ns) + 2.0 ** -36))
    bound = ((2 * np.abs(psi) + d) * d).sum(1)
    assert (np.abs(r - rho) <= bound).all()
    # four phi digits are a visible approximation (the bound is not vacuous) ...
    nz = rho > 0
    assert np.abs(r - rho)[nz].max() / rho[nz].max() > 1e-14
    # ... and the coefficient levels matter: dropping level 5 moves the result
    assert not np.array_equal(q8_density_digits(phi, At, ns, top=4), r)
