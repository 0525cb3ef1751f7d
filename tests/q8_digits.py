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
    """int8 digits of x under the power-of-two exponent e: y_0 = x 2^(6-e),
    d_s = rint(y_s), y_(s+1) = 128 (y_s - d_s); all steps exact in FP64."""
    y = np.ldexp(x, 6 - e)
    out = []
    for _ in range(n):
        d = np.rint(y)
        assert np.abs(d).max() <= 64
        out.append(d)
        y = (y - d) * 128.0
    return out


def q8_density_digits(phi, At, ns, top=5):
    """The int8-digit density restated on the host: psi_i(r) = 2^(e_i+e_r-12)
    sum_d 2^(-7d) P_d with P_d the exact sum over digit pairs k+l = d (k <= top),
    rho = sum_i psi_i^2.  Digit products are integers below 2^22, so FP64
    matrix products give them exactly."""
    er = q8_exponent(np.abs(phi).max(1))[:, None, :]          # (B, 1, ld)
    ei = q8_exponent(np.abs(At).max(1))[:, None, :]           # (B, 1, m)
    xd = q8_digits(phi, er, ns)
    ad = q8_digits(At, ei, top + 1)
    P = [np.zeros((phi.shape[0], At.shape[2], phi.shape[2])) for _ in range(top + 1)]
    for l in range(ns):
        for k in range(top + 1 - l):
            P[k + l] += np.matmul(ad[k].transpose(0, 2, 1), xd[l])
    Q = P[top]
    for d in range(top - 1, -1, -1):
        Q = Q * 2.0 ** -7 + P[d]
    w = np.ldexp(1.0, 2 * ei[:, 0, :])[:, :, None]               # (B, m, 1)
    return np.ldexp((w * Q * Q).sum(1), 2 * er[:, 0, :] - 24)
