"""Host restatement of the int8-digit orbital density (hqc_orb_density_q8), used
by the tests as the exact expectation of the kernel's digit algorithm.  Test
infrastructure only."""
import numpy as np


def q8_exponent(m):
    """Smallest e with m < 2^e (0 for m = 0), elementwise."""
    e = np.zeros(m.shape, dtype=np.int64)
    nz = m > 0
    e[nz] = np.frexp(m[nz])[1]
    return e


def q8_digits(x, e, n):
    """Base-256 int8 digits of x under the power-of-two exponent e: y_0 = x 2^(6-e),
    d_s = rint(y_s), y_(s+1) = 256 (y_s - d_s), all steps exact in FP64; then a
    digit +128 becomes -128 with a carry of 1 into the digit above (last to first),
    so every digit lies in [-128, 127] and sum_s 2^(-8s) d_s is unchanged."""
    y = np.ldexp(x, 6 - e)
    out = []
    for _ in range(n):
        d = np.rint(y)
        out.append(d)
        y = (y - d) * 256.0
    for s in range(n - 1, 0, -1):
        c = out[s] == 128
        out[s] = np.where(c, -128.0, out[s])
        out[s - 1] = out[s - 1] + c
    assert all(((d >= -128) & (d <= 127)).all() for d in out)
    return out


def q8_density_digits(phi, At, ns, top=5, nstored=None):
    """The int8-digit density restated on the host: psi_i(r) = 2^(e_i+e_r-12)
    sum_d 2^(-8d) P_d with P_d the exact sum over digit pairs k+l = d (k <= top),
    rho = sum_i psi_i^2, from the first ns of nstored (default ns) phi digits: the
    prefix of a longer expansion, whose last kept digit may hold a carry from the
    dropped ones.  Digit products are integers below 2^23 for up to 128 AOs, so
    FP64 matrix products give them exactly."""
    er = q8_exponent(np.abs(phi).max(1))[:, None, :]          # (B, 1, ld)
    ei = q8_exponent(np.abs(At).max(1))[:, None, :]           # (B, 1, m)
    xd = q8_digits(phi, er, ns if nstored is None else nstored)[:ns]
    ad = q8_digits(At, ei, top + 1)
    P = [np.zeros((phi.shape[0], At.shape[2], phi.shape[2])) for _ in range(top + 1)]
    for l in range(ns):
        for k in range(top + 1 - l):
            P[k + l] += np.matmul(ad[k].transpose(0, 2, 1), xd[l])
    Q = P[top]
    for d in range(top - 1, -1, -1):
        Q = Q * 2.0 ** -8 + P[d]
    w = np.ldexp(1.0, 2 * ei[:, 0, :])[:, :, None]               # (B, m, 1)
    return np.ldexp((w * Q * Q).sum(1), 2 * er[:, 0, :] - 24)
