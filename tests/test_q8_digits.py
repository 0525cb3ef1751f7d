"""The host restatement of the int8-digit density (tests/q8_digits.py) pinned
against what the mathematics fixes: base-256 digit expansions stay in int8 and
reconstruct their value to 2^(e - 7 - 8(n-1)) (the carries included), enough
digits and levels reproduce the FP64 orbital density, and the kept digits (4 of
phi, levels <= 5 or <= 4) stay within the stated error bound."""
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


@pytest.mark.parametrize("n", [1, 3, 4, 6])
def test_digits_reconstruct(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((5, 200)) * np.exp2(rng.integers(-30, 10, (5, 1)))
    e = q8_exponent(np.abs(x).max(1))[:, None]
    d = q8_digits(x, e, n)
    rec = sum(np.ldexp(ds, e - 6 - 8 * s) for s, ds in enumerate(d))
    assert (np.abs(rec - x) <= np.ldexp(1.0, e - 7 - 8 * (n - 1))).all()
    # the leading digit carries the sign and at most 7 bits (plus a carry)
    assert (np.abs(d[0]) <= 65).all()
    assert all(((ds >= -128) & (ds <= 127)).all() for ds in d)


def test_digit_prefix():
    """The first n of N > n digits differ from the n-digit expansion only by a carry
    from the dropped digits into the last kept one, so the prefix stays within
    2^(e - 7 - 8(n-1)) (1 + 2^-8) of the value."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal((4, 500))
    e = q8_exponent(np.abs(x).max(1))[:, None]
    for n, N in ((3, 4), (3, 5), (4, 5)):
        full = q8_digits(x, e, N)
        own = q8_digits(x, e, n)
        assert all(np.array_equal(a, b) for a, b in zip(full[:n - 1], own[:n - 1]))
        assert np.abs(full[n - 1] - own[n - 1]).max() <= 1
        rec = sum(np.ldexp(ds, e - 6 - 8 * s) for s, ds in enumerate(full[:n]))
        assert (np.abs(rec - x) <= np.ldexp(1.0 + 2.0 ** -8, e - 7 - 8 * (n - 1))).all()


def test_digit_carries():
    """Values whose expansion hits a remainder of exactly 1/2: the digit that would
    be +128 is stored as -128 with a carry, and the expansion still sums exactly."""
    e = np.zeros(6, dtype=np.int64)
    # y_0 = 64 x; a remainder of exactly 1/2 after digit 0 or 1 gives y = 128 next
    x = np.ldexp(np.array([32.5, -32.5, 2.5, 0.5, 127.5 / 256, 1.0 + 2.0 ** -9]), -6)
    d = q8_digits(x, e, 3)
    rec = sum(np.ldexp(ds, -6 - 8 * s) for s, ds in enumerate(d))
    assert np.array_equal(rec, x)
    assert (d[1] == -128).sum() >= 2                 # carries happened
    assert all(((ds >= -128) & (ds <= 127)).all() for ds in d)
    # a plain rint expansion of the same values needs the forbidden digit +128
    y = np.ldexp(x, 6)
    assert (np.rint((y - np.rint(y)) * 256.0) == 128).any()


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


@pytest.mark.parametrize("top", [5, 4])
def test_kept_digits_within_bound(top):
    """|psi_q - psi| <= n_ao max|At_i| max|phi(r)| (2^(-6-8(ns-1)) + ns 2^(-3-8 top)):
    the phi digits' residual, plus the dropped digit products of level > top and the
    coefficient residual (each tail of a carried expansion below 2^(-8(j-1)))."""
    phi, At = _system(2)
    psi = np.matmul(At.transpose(0, 2, 1), phi)
    rho = (psi ** 2).sum(1)
    ns = 4
    r = q8_density_digits(phi, At, ns, top)
    nao = phi.shape[1]
    d = (nao * np.abs(At).max(1)[:, :, None] * np.abs(phi).max(1)[:, None, :]
         * (2.0 ** (-6 - 8 * (ns - 1)) + ns * 2.0 ** (-3 - 8 * top)))
    bound = ((2 * np.abs(psi) + d) * d).sum(1)
    assert (np.abs(r - rho) <= bound).all()
    # four phi digits are a visible approximation (the bound is not vacuous) ...
    nz = rho > 0
    assert np.abs(r - rho)[nz].max() / rho[nz].max() > 1e-16
    # ... and the coefficient levels matter: one level fewer moves the result
    assert not np.array_equal(q8_density_digits(phi, At, ns, top=top - 1), r)
