"""CPU restatement (NumPy/SciPy, FP64) of the reference's periodic GPW SCF path.

TEST INFRASTRUCTURE ONLY.  Only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` may import this module, and only as the
checker; the product package `hqc_amd` never imports it.

Every function cites the reference file:line it restates.  Two independent
routes are kept for the two-electron part:

* the *literal* route (`literal=True`): the pair-density tensor
  rhoG[G, p, q] built separably from 1-D FFTs with the Hartley (Re+Im) trick at
  Gamma, then V_ep, J (and the ERIs for HF) contracted from it, exactly as
  hqc/pbc/solver.py:391-629.  Memory is O(N_g * n_ao^2), so it is only used on
  small systems.
* the *FFT-Poisson* route (default): the same quantities written on the real
  space grid, rho(r) -> FFT -> V(G)/G^2 -> IFFT, V_pq = dV sum_r phi_p v phi_q
  and exchange from pair-density FFTs.  On the reference's symmetric odd G-set
  the two routes are algebraically identical; tests pin one against the other.

Parity status of each function is stated in its docstring ("pinned by ...").
"""
import math
import os

import numpy as np
import scipy.linalg
import scipy.special

RY = 2.0                      # hqc/pbc/solver.py:153
EXCHANGE_G0_CONST = 0.22578495  # hqc/pbc/solver.py:404 (literal kept for parity)
_BASIS_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "hqc_amd", "basis", "data")


# --------------------------------------------------------------------------
# Basis: CP2K text -> normalised coefficient tables (hqc/basis/parse.py)
# --------------------------------------------------------------------------
def read_basis_sets(basis, element="H"):
    """Parse one element's CP2K basis block.

    Restates parse.load_as_str/parse_quant_num/parse_gto
    (hqc/basis/parse.py:32-175): split on '#BASIS SET', take the block whose
    first line starts with the element, line 3 is the number of sets, each set
    is a quantum-number line 'n lmin lmax nprim nfunc(l)...' followed by nprim
    rows 'alpha c_1 ... c_k'.  Pinned by: tests/test_oracle_basis_grid.py
    (test_basis_structure: known H gth-dzvp structure; test_normalisation_closed_form).
    """
    path = basis if os.path.isfile(basis) else os.path.join(_BASIS_DIR, basis + ".dat")
    if not os.path.isfile(path):
        raise FileNotFoundError(f"No '{basis}.dat' basis file found.")
    with open(path) as f:
        text = f.read()
    block = None
    for chunk in text.split("#BASIS SET\n"):
        if chunk[:len(element)] == element:
            block = chunk
            break
    if block is None:
        raise ValueError(f"No '{element}' in '{path}'.")
    lines = [ln.strip() for ln in ("#BASIS SET\n" + block).split("\n") if ln.strip()]
    nsets = int(lines[2])
    sets = []
    line = 3
    for _ in range(nsets):
        quant = np.array([int(w) for w in lines[line].split()])
        if len(quant) != 5 + quant[2] - quant[1]:
            raise ValueError(f"Line {lines[line]} has incorrect number of l orbitals.")
        rows = np.array([[float(w) for w in ln.split()]
                         for ln in lines[line + 1:line + 1 + quant[3]]])
        if rows.shape != (quant[3], int(np.sum(quant[4:])) + 1):
            raise ValueError(f"Line {lines[line + 1]} has incorrect number of coefficents.")
        sets.append((quant, rows))
        line += quant[3] + 1
    return sets


def _dfact(n):
    return 1 if n <= 0 else n * _dfact(n - 2)


def normalize_sets(sets):
    """Primitive radial normalisation times contraction normalisation.

    Restates normalize_gto_coeff (hqc/basis/parse.py:192-239):
    c_i * alpha_i^(l/2+3/4) * 2^(l+7/4) / pi^(1/4) / sqrt((2l+1)!!) * N,
    N^-2 = 2^(l+3/2) sum_ij c_i c_j (a_i+a_j)^(-l-3/2) (a_i a_j)^(l/2+3/4).
    Pinned by: closed-form self-overlap of every contracted radial function = 1
    (tests/test_oracle_basis_grid.py test_normalisation_closed_form).
    """
    out = []
    for quant, rows in sets:
        alpha = rows[:, 0]
        asum = alpha[:, None] + alpha[None, :]
        aprod = alpha[:, None] * alpha[None, :]
        cols = [rows[:, 0:1]]
        lmin, lmax = int(quant[1]), int(quant[2])
        for l in range(lmin, lmax + 1):
            const = 2.0 ** (l + 1.75) / np.pi ** 0.25 / math.sqrt(_dfact(2 * l + 1))
            i0 = int(np.sum(quant[4:l - lmin + 4])) + 1
            i1 = int(np.sum(quant[4:l - lmin + 5])) + 1
            for i in range(i0, i1):
                c = rows[:, i]
                s = 2.0 ** (l + 1.5) * np.einsum("i,j,ij,ij", c, c, asum ** (-l - 1.5),
                                                   aprod ** (l / 2 + 0.75))
                cols.append((c * alpha ** (l / 2 + 0.75) * const * s ** -0.5)[:, None])
        out.append((quant, np.concatenate(cols, axis=1)))
    return out


class AtomBasis:
    """Per-atom AO description with the solid-harmonic constants folded in.

    Restates _make_alpha / _make_alpha_coeff_cart (hqc/pbc/solver.py:88-151) and
    the cartesian->spherical tables of hqc/pbc/gto.py:108-300 for l <= 1:
    s columns are scaled by 1/sqrt(4 pi), p columns by sqrt(3/(4 pi)).
    Local AO order: for each contracted function in set order, one AO for s,
    three (x, y, z) for p (hqc/pbc/gto.py:268-300; hqc/pbc/solver.py:316-321).
      alpha  (n_alpha,)            all primitive exponents (sets concatenated)
      coef   (n_alpha, n_ao_atom)  coefficient of primitive a in local AO o
      comp   (n_ao_atom,)          -1 for s, 0/1/2 for p_x/p_y/p_z
    """

    def __init__(self, basis):
        sets = normalize_sets(read_basis_sets(basis))
        lmin = min(int(q[1]) for q, _ in sets)
        lmax = max(int(q[2]) for q, _ in sets)
        if lmin != 0:
            raise ValueError("l_min != 0")
        if lmax > 1:
            raise ValueError("l_max > 1 is not supported yet.")
        self.l_max = lmax
        self.alpha = np.concatenate([rows[:, 0] for _, rows in sets])
        n_alpha = len(self.alpha)
        cols, comps, ls = [], [], []
        off = 0
        for quant, rows in sets:
            lo = int(quant[1])
            for l in range(lo, int(quant[2]) + 1):
                i0 = int(np.sum(quant[4:l - lo + 4])) + 1
                i1 = int(np.sum(quant[4:l - lo + 5])) + 1
                for i in range(i0, i1):
                    full = np.zeros(n_alpha)
                    full[off:off + rows.shape[0]] = rows[:, i]
                    if l == 0:
                        cols.append(full / np.sqrt(4 * np.pi)); comps.append(-1); ls.append(0)
                    else:
                        for d in range(3):
                            cols.append(full * np.sqrt(3 / (4 * np.pi))); comps.append(d); ls.append(1)
            off += rows.shape[0]
        self.coef = np.stack(cols, axis=1)
        self.comp = np.array(comps)
        self.l = np.array(ls)
        # the solver's S/T block assembly assumes all s functions come first
        # (hqc/pbc/solver.py:316-321); the H bases used here satisfy it.
        if np.any(np.diff(self.l) < 0):
            raise ValueError("basis with p before s functions is not supported")
        self.n_ao_atom = self.coef.shape[1]


# --------------------------------------------------------------------------
# Cell, grid and lattice set-up (hqc/pbc/solver.py:153-207, gto.py:47-49)
# --------------------------------------------------------------------------
class System:
    """Static set-up shared by every call, like the make_solver closure.

    n_grid = round(L_bohr/grid_length/2)*2+1 with Python's round
    (hqc/pbc/solver.py:159); R grid j*L/g (:189-191); G = k*2pi/L for
    k = -g//2..g//2; VG = 4 pi/(Omega G^2), VG(0) = 0 (:193-200); spherical
    lattice |T| <= rcut for the analytic S/T sums (:162-176); a cube of
    +-floor(24/L) images for grid AO values (hqc/pbc/gto.py:47-49, rcut fixed at
    its default 24 there).  Grid points are stored in (x, y, z) order with z
    fastest; the reference's meshgrid 'xy' order only permutes summands.
    Pinned by: tests/test_oracle_basis_grid.py (test_grid_rule_half_even,
    test_VG_closed_form, test_lattice_counts).
    """

    def __init__(self, n, L, rs, basis, rcut=24.0, grid_length=0.12):
        self.n = n
        self.rs = rs
        self.basis = AtomBasis(basis)
        self.L = L * rs
        self.omega = self.L ** 3
        self.g = round(self.L / grid_length / 2) * 2 + 1
        g = self.g
        self.N = g ** 3
        self.dV = (self.L / g) ** 3
        self.n_ao = n * self.basis.n_ao_atom
        self.r1d = np.arange(g) * self.L / g
        tmax = rcut // self.L
        nt = np.arange(-tmax, tmax + 1)
        T = np.array(np.meshgrid(nt, nt, nt, indexing="ij")).reshape(3, -1).T * self.L
        self.lattice = T[(T ** 2).sum(1) <= rcut ** 2]
        t1 = 24.0 // self.L
        self.images_1d = np.arange(-t1, t1 + 1) * self.L
        # FFT-ordered integer frequencies and G vectors
        k = np.fft.fftfreq(g, d=1.0 / g)              # 0..m, -m..-1
        self.kfreq = np.rint(k).astype(int)
        G1 = self.kfreq * 2 * np.pi / self.L
        Gx, Gy, Gz = np.meshgrid(G1, G1, G1, indexing="ij")
        self.G = np.stack([Gx, Gy, Gz], axis=-1)      # (g,g,g,3) FFT order
        G2 = Gx ** 2 + Gy ** 2 + Gz ** 2
        with np.errstate(divide="ignore"):
            VG = 4 * np.pi / self.omega / G2
        VG[0, 0, 0] = 0.0
        self.VG = VG                                   # FFT order
        # grid point coordinates (N_g, 3)
        X, Y, Z = np.meshgrid(self.r1d, self.r1d, self.r1d, indexing="ij")
        self.R = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


# --------------------------------------------------------------------------
# Analytic periodic overlap/kinetic (hqc/pbc/solver.py:233-389)
# --------------------------------------------------------------------------
def _hermite_factors(mu, R, nmax):
    """e_k(R) = d^k/dR^k exp(-mu R^2) / exp(-mu R^2), k = 0..nmax (elementwise).

    Recurrence e_{k+1} = -2 mu R e_k - 2 k mu e_{k-1} (physicists' Hermite
    polynomials in sqrt(mu) R).
    """
    e = [np.ones_like(R), -2 * mu * R]
    for k in range(1, nmax):
        e.append(-2 * mu * R * e[k] - 2 * k * mu * e[k - 1])
    return e[:nmax + 1]


def _dgauss(mu, pref, R, orders):
    """d^(n1,n2,n3)/dR^n of pref*exp(-mu |R|^2) for R (..., 3); orders (n1,n2,n3)."""
    val = pref * np.exp(-mu * np.sum(R * R, axis=-1))
    for d in range(3):
        if orders[d]:
            val = val * _hermite_factors(mu, R[..., d], orders[d])[orders[d]]
    return val


def _pair_integral(kind, alpha, beta, R, la, lb, extra=None):
    """One primitive pair, summed later over the lattice.

    s/p derivative trick of hqc/pbc/solver.py:307-341: a p_i function on A is
    (1/2a) d/dA_i of the s function, a p_j on B is (1/2b) d/dB_j; R = A-B+T so
    d/dA = d/dR and d/dB = -d/dR.  Kinetic = -1/2 Laplacian_R of the s-s
    overlap pi^1.5 (a+b)^-1.5 exp(-mu R^2), which equals the reference's
    mu(3-2mu R^2) S (:252-254).  `extra` adds one more d/dR_c (forces).
    la/lb: -1 for s, 0..2 for the p component.
    """
    mu = alpha * beta / (alpha + beta)
    pref = np.pi ** 1.5 * (alpha + beta) ** -1.5
    base = [0, 0, 0]
    fac = 1.0
    if la >= 0:
        base[la] += 1
        fac *= 1.0 / (2 * alpha)
    if lb >= 0:
        base[lb] += 1
        fac *= -1.0 / (2 * beta)
    if extra is not None:
        base[extra] += 1
    if kind == "S":
        return fac * _dgauss(mu, pref, R, base)
    tot = 0.0
    for d in range(3):
        o = list(base)
        o[d] += 2
        tot = tot + _dgauss(mu, pref, R, o)
    return -0.5 * fac * tot


def overlap_kinetic(sysm, xp_bohr, kpt=None, deriv=False):
    """S and T (n_ao, n_ao) by analytic lattice sums; optionally dS, dT.

    Restates eval_overlap_kinetic / _kpt (hqc/pbc/solver.py:233-389):
    S_pq = sum_T e^{-i k.T} sum_ab c_ap c_bq I_ab(R_A - R_B + T).
    deriv=True also returns dS[p, q, c] = d S_pq / d R_{atom(p), c} holding
    R_{atom(q)} fixed (first-argument derivative), used for forces.
    Pinned by: tests/test_oracle_physics.py (test_isolated_atom_normalised,
    test_overlap_kinetic_vs_grid_quadrature, test_overlap_derivatives_fd).
    """
    b = sysm.basis
    n, nloc = sysm.n, b.n_ao_atom
    T = sysm.lattice
    phase = None if kpt is None else np.exp(-1j * T @ np.asarray(kpt))
    dtype = float if kpt is None else complex
    S = np.zeros((n, nloc, n, nloc), dtype)
    K = np.zeros_like(S)
    dS = np.zeros((n, nloc, n, nloc, 3), dtype) if deriv else None
    dK = np.zeros_like(dS) if deriv else None
    na = len(b.alpha)
    R = xp_bohr[:, None, None, :] - xp_bohr[None, :, None, :] + T[None, None]   # (n,n,nT,3)
    for i in range(na):
        for j in range(na):
            for o in range(nloc):
                if b.coef[i, o] == 0.0:
                    continue
                for q in range(nloc):
                    if b.coef[j, q] == 0.0:
                        continue
                    w = b.coef[i, o] * b.coef[j, q]
                    args = (b.alpha[i], b.alpha[j], R, b.comp[o], b.comp[q])
                    vs = _pair_integral("S", *args)
                    vk = _pair_integral("T", *args)
                    if phase is not None:
                        vs = vs * phase; vk = vk * phase
                    S[:, o, :, q] += w * vs.sum(-1)
                    K[:, o, :, q] += w * vk.sum(-1)
                    if deriv:
                        for c in range(3):
                            ds = _pair_integral("S", *args, extra=c)
                            dk = _pair_integral("T", *args, extra=c)
                            if phase is not None:
                                ds = ds * phase; dk = dk * phase
                            dS[:, o, :, q, c] += w * ds.sum(-1)
                            dK[:, o, :, q, c] += w * dk.sum(-1)
    nao = n * nloc
    S = S.reshape(nao, nao)
    K = K.reshape(nao, nao)
    if deriv:
        return S, K, dS.reshape(nao, nao, 3), dK.reshape(nao, nao, 3)
    return S, K


# --------------------------------------------------------------------------
# Periodic GTO tables and AO values on the grid (hqc/pbc/gto.py:326-420)
# --------------------------------------------------------------------------
def gto_tables(sysm, xp_bohr, kpt=None, deriv=False):
    """P[A, d, j, a, l] = sum_t (x - tL)^l exp(-alpha_a (x - tL)^2) [e^{i k_d t L}],
    x = R1d[j] - xp[A, d], l = 0..l_max (+1 if deriv).

    Restates eval_pbc_gaussian_power_x(_kpt) (hqc/pbc/gto.py:326-339, 371-386)
    vmapped over atoms and the 1-D grid (hqc/pbc/solver.py:209-226).
    Pinned by: brute-force 3-D image sums (tests/test_oracle_basis_grid.py
    test_tables_bruteforce_images).
    """
    b = sysm.basis
    lmax = b.l_max + (2 if deriv else 0)
    x = sysm.r1d[None, None, :] - xp_bohr[:, :, None]              # (n,3,g)
    xt = x[..., None] - sysm.images_1d                               # (n,3,g,t)
    gauss = np.exp(-b.alpha[:, None] * xt[..., None, :] ** 2)        # (n,3,g,a,t)
    if kpt is not None:
        ph = np.exp(1j * np.asarray(kpt)[None, :, None, None, None] * sysm.images_1d)
        gauss = gauss * ph
    pw = np.stack([xt ** l for l in range(lmax + 1)], axis=-2)      # (n,3,g,l,t)
    return np.einsum("ndgat,ndglt->ndgal", gauss, pw)


def ao_values(sysm, xp_bohr, kpt=None):
    """phi (n_ao, N_g): AO values on the grid, separable product of 1-D sums.

    Restates eval_pbc_gto_sph / eval_pbc_ao (hqc/pbc/gto.py:341-369, 388-420):
    phi_(A,o)(r) = sum_a coef[a,o] prod_d P[A,d,r_d,a,pow_d(o)].
    Pinned by: brute-force 3-D image sum at random points, periodicity
    (tests/test_oracle_basis_grid.py test_ao_periodicity, tests/test_oracle_slater.py
    test_points_brute_force_image_sum).
    """
    P = gto_tables(sysm, xp_bohr, kpt)
    return _ao_from_tables(sysm, P)


def _ao_from_tables(sysm, P, grad_dir=None):
    b = sysm.basis
    g = sysm.g
    out = []
    for A in range(sysm.n):
        for o in range(b.n_ao_atom):
            acc = 0.0
            for a in range(len(b.alpha)):
                c = b.coef[a, o]
                if c == 0.0:
                    continue
                f = []
                for d in range(3):
                    pw = 1 if b.comp[o] == d else 0
                    if grad_dir == d:
                        # d/dx [x^l e^{-a x^2}] = l x^{l-1} e - 2a x^{l+1} e
                        v = -2 * b.alpha[a] * P[A, d, :, a, pw + 1]
                        if pw:
                            v = v + P[A, d, :, a, 0]
                        f.append(v)
                    else:
                        f.append(P[A, d, :, a, pw])
                acc = acc + c * (f[0][:, None, None] * f[1][None, :, None] * f[2][None, None, :])
            out.append(np.asarray(acc).reshape(g ** 3) if np.ndim(acc) else np.zeros(g ** 3))
    return np.array(out)


def ao_gradients(sysm, xp_bohr, kpt=None):
    """grad phi (3, n_ao, N_g) = d phi / d r on the grid (complex at a k-point),
    from the same 1-D tables with one extra power (for the forces, which the
    reference does not compute)."""
    P = gto_tables(sysm, xp_bohr, kpt, deriv=True)
    return np.array([_ao_from_tables(sysm, P, grad_dir=d) for d in range(3)])


def ao_points(basis, L_bohr, xp_bohr, xe_bohr, rcut=24.0, kpt=None):
    """phi (n_e, n_ao) at electron positions: make_pbc_gto's eval_pbc_ao /
    eval_pbc_ao_kpt (hqc/pbc/gto.py:326-369, 371-420) with the 1-D image set
    (-tmax..tmax) L, tmax = rcut // L (gto.py:48-49), and the phase
    e^{i k_d t L} at a k-point.  basis: an AtomBasis.
    Pinned by: equality with ao_values at grid points (tests/test_oracle_slater.py)."""
    b = basis
    tmax = rcut // L_bohr
    lat = np.arange(-tmax, tmax + 1) * L_bohr
    r = np.asarray(xe_bohr, float)[:, None, :] - np.asarray(xp_bohr, float)[None, :, :]
    xt = r[..., None] - lat                                          # (ne, n, 3, t)
    gauss = np.exp(-b.alpha[:, None] * xt[..., None, :] ** 2)        # (ne, n, 3, a, t)
    if kpt is not None:
        gauss = gauss * np.exp(1j * np.asarray(kpt)[None, None, :, None, None] * lat)
    P = np.stack([(gauss * xt[..., None, :] ** l).sum(-1) for l in range(b.l_max + 1)], -1)
    ne, n = r.shape[0], r.shape[1]
    out = np.zeros((ne, n, b.n_ao_atom), dtype=P.dtype)
    for o in range(b.n_ao_atom):
        pw = [1 if b.comp[o] == d else 0 for d in range(3)]
        f = P[:, :, 0, :, pw[0]] * P[:, :, 1, :, pw[1]] * P[:, :, 2, :, pw[2]]   # (ne, n, a)
        out[:, :, o] = f @ b.coef[:, o]
    return out.reshape(ne, n * b.n_ao_atom)


def slater(n, L, rs, basis, xp, xe, mo_coeff, kpt=None, state_idx=None, rcut=24.0):
    """(slater_up, slater_dn) of make_slater (hqc/pbc/slater.py:7-104): AO values at
    the electrons times the occupied columns; state_idx defaults to the ground
    state [0..n/2-1, 0..n/2-1]."""
    assert n % 2 == 0
    if state_idx is None:
        state_idx = np.concatenate([np.arange(n // 2), np.arange(n // 2)])
    ao = ao_points(AtomBasis(basis), L * rs, np.asarray(xp) * rs, np.asarray(xe) * rs, rcut, kpt)
    h = n // 2
    return ao[:h] @ mo_coeff[:, state_idx[:h]], ao[h:] @ mo_coeff[:, state_idx[h:]]


# --------------------------------------------------------------------------
# Literal pair-density route (hqc/pbc/solver.py:391-629) -- small systems only
# --------------------------------------------------------------------------
def _coef_tensor(b):
    """alpha_coeff_gto_cart2sph (n_alpha, n_cart, n_ao_atom) and power2cart
    (3, n_l, n_cart) for l_max <= 1 (hqc/pbc/gto.py:138-300)."""
    ncart = 1 if b.l_max == 0 else 4
    C = np.zeros((len(b.alpha), ncart, b.n_ao_atom))
    for o in range(b.n_ao_atom):
        C[:, b.comp[o] + 1 if b.l_max else 0, o] = b.coef[:, o]
    p2c = np.zeros((3, b.l_max + 1, ncart))
    for d in range(3):
        for c in range(ncart):
            p2c[d, 1 if c == d + 1 else 0, c] = 1.0
    return C, p2c


def rhoG_literal(sysm, xp_bohr, kpt=None):
    """rhoG (g, g, g, n_ao, n_ao), G axes in ascending -g//2..g//2 order, (x,y,z).

    Restates the lax.scan body of eval_vep_eris / eval_vep_eris_dft (and _kpt)
    (hqc/pbc/solver.py:410-423, 454-467, 500-512, 543-555): per dimension the
    1-D pair products are FFT'd and scaled by L/g (:412), fft-shifted by the
    concatenation at :413, mapped power->cartesian (:414), multiplied over the
    three dimensions (:416), Hartley Re+Im at Gamma (:418) and contracted with
    the coefficient tensor (:419).  The "one" side is conjugated at k (:501).
    """
    b = sysm.basis
    g, n, m = sysm.g, sysm.n, sysm.g // 2
    P = gto_tables(sysm, xp_bohr, kpt)                       # (n,3,g,a,l)
    C, p2c = _coef_tensor(b)
    nloc = b.n_ao_atom
    out = np.zeros((g, g, g, n, nloc, n, nloc), dtype=float if kpt is None else complex)
    for B in range(n):
        one = P[B] if kpt is None else np.conj(P[B])          # (3,g,b,k)
        pair = np.einsum("ndgal,dgbk->dgnablk", P, one)
        F = np.fft.fft(pair, axis=1) * (sysm.L / g)
        F = np.concatenate([F[:, -m:], F[:, :-m]], axis=1)   # G = -m..m
        cart = np.einsum("dgnablk,dlc,dke->dgnabce", F, p2c, p2c)
        for bb in range(len(b.alpha)):
            for e in range(C.shape[1]):
                w = C[bb, e]                                   # (nloc,)
                if not np.any(w):
                    continue
                cx, cy, cz = (cart[d][:, :, :, bb, :, e] for d in range(3))   # (g,n,a,c)
                prod = cx[:, None, None] * cy[None, :, None] * cz[None, None, :]  # (x,y,z,n,a,c)
                if kpt is None:
                    prod = prod.real + prod.imag
                tmp = np.einsum("xyznac,aco->xyzno", prod, C)
                out[:, :, :, :, :, B, :] += tmp[..., None] * w
    nao = n * nloc
    return out.reshape(g, g, g, nao, nao)


def _shifted(sysm, arr_fft_order):
    """FFT-ordered (g,g,g) array -> ascending -m..m order on every axis."""
    return np.fft.fftshift(arr_fft_order)


def vloc_G(sysm, xp_bohr):
    """v_loc(G) = -VG(G) sum_A exp(-i G.R_A), FFT order (hqc/pbc/solver.py:406-407)."""
    SI = np.exp(-1j * np.einsum("xyzd,nd->xyzn", sysm.G, xp_bohr)).sum(-1)
    return -SI * sysm.VG


def literal_integrals(sysm, xp_bohr, rhoG, kpt=None, hf=True):
    """V_ep and (HF) eris, eris0 from rhoG (hqc/pbc/solver.py:424-433, 513-524)."""
    VG = _shifted(sysm, sysm.VG)
    vl = _shifted(sysm, vloc_G(sysm, xp_bohr))
    c0 = 4 * np.pi / sysm.L * EXCHANGE_G0_CONST
    if kpt is None:
        vh = vl.real + vl.imag
        vep = np.einsum("xyz,xyzpq->pq", vh, rhoG)
    else:
        vep = np.einsum("xyz,xyzpq->pq", vl, np.conj(rhoG))
    if not hf:
        return vep, None, None
    m = sysm.g // 2
    r0 = rhoG[m, m, m]
    if kpt is None:
        eris = np.einsum("xyz,xyzrs,xyzqp->prsq", VG, rhoG, rhoG, optimize=True)
    else:
        eris = np.einsum("xyz,xyzrs,xyzpq->prsq", VG, rhoG, np.conj(rhoG), optimize=True)
    eris0 = np.einsum("rs,qp->prsq", r0, r0) * c0
    return vep, eris, eris0


def hartree_literal(sysm, rhoG, dm):
    """J from rhoG (hqc/pbc/solver.py:616-629): rho(G)=sum D rhoG, J=sum rho VG rhoG*."""
    VG = _shifted(sysm, sysm.VG)
    rho = np.einsum("rs,xyzrs->xyz", dm, rhoG)
    return np.einsum("xyz,xyz,xyzpq->pq", rho, VG, np.conj(rhoG))


# --------------------------------------------------------------------------
# FFT-Poisson route (the build's restatement; same algebra, O(n_ao N_g) memory)
# --------------------------------------------------------------------------
def vloc_real(sysm, xp_bohr):
    """v_loc(r) = sum_G v_loc(G) e^{iG.r} on the grid (real on the symmetric odd G-set).

    Equivalent to the Hartley contraction of hqc/pbc/solver.py:406-408,427.
    """
    return (np.fft.ifftn(vloc_G(sysm, xp_bohr)) * sysm.N).real.ravel()


def poisson(sysm, rho):
    """v_H(r) = sum_G VG rho~(G) e^{iG.r}, rho~ = dV * FFT(rho).

    Equivalent to J of hqc/pbc/solver.py:616-629 (and :592-602 for HF) written
    as J_pq = dV sum_r phi_p v_H phi_q.  Returns (v_H, rho~ in FFT order).
    """
    g = sysm.g
    rg = np.fft.fftn(rho.reshape(g, g, g)) * sysm.dV
    vh = np.fft.ifftn(rg * sysm.VG) * sysm.N
    return vh.real.ravel(), rg


def grid_matrix(sysm, phi, v):
    """dV * phi^* diag(v) phi^T  -> (n_ao, n_ao) ('pr,qr,r->pq' with conj on p)."""
    return (np.conj(phi) * v) @ phi.T * sysm.dV


def grid_density(phi, dm):
    """rho(r) = Re sum_pq D_pq phi_p(r) phi_q(r)^*  (hqc/pbc/solver.py:690)."""
    return np.einsum("pr,pr->r", dm.T @ phi, np.conj(phi)).real


def exchange_fft(sysm, phi, mo_coeff, occ, kpt=None):
    """K (n_ao, n_ao) via pair-density FFTs; equals sum_rs D_rs (eris+eris0)[p,q,s,r].

    K_pq = sum_i f_i sum_G VG A_qi(G) A_pi(G)^* + c0 (S~ D^T S~)_qp with
    A_pi = dV FFT[phi_p psi_i^*], S~_ab = dV sum_r phi_a phi_b^*; at Gamma the real
    part is taken (the Hartley products of hqc/pbc/solver.py:418,430 reduce to it).
    The G=0 term restates eris0 (:431) with the literal 0.22578495.
    """
    g = sysm.g
    c0 = 4 * np.pi / sysm.L * EXCHANGE_G0_CONST
    psi = mo_coeff.T @ phi                             # (n_mo, N_g): psi_i = sum_s c_si phi_s
    nao = phi.shape[0]
    K = np.zeros((nao, nao), dtype=complex)
    for i in np.nonzero(occ > 0)[0]:
        pair = phi * np.conj(psi[i])[None, :]
        A = np.fft.fftn(pair.reshape(nao, g, g, g), axes=(1, 2, 3)).reshape(nao, -1) * sysm.dV
        K += occ[i] * (A * sysm.VG.ravel()[None, :]) @ np.conj(A).T   # [q,p]
    K = K.T
    St = (phi @ np.conj(phi).T) * sysm.dV
    dm = (mo_coeff * occ) @ np.conj(mo_coeff).T
    K0 = (St @ dm.T @ St).T * c0
    K = K + K0
    return K.real if kpt is None else K


# --------------------------------------------------------------------------
# Exchange-correlation (hqc/tools/excor.py) with closed-form derivatives
# --------------------------------------------------------------------------
_WS = (3 / 4 / np.pi) ** (1 / 3)
_VWN = dict(A=0.0621814, y0=-0.10498, b=3.72744, c=12.93532)


def _vwn_consts():
    A, y0, b, c = _VWN["A"], _VWN["y0"], _VWN["b"], _VWN["c"]
    Yy0 = y0 ** 2 + b * y0 + c
    Q = (4 * c - b ** 2) ** 0.5
    c0 = A / 2
    c1 = c0 * 2 * b / Q
    c2 = c0 * b * y0 / Yy0
    c3 = c2 * 2 * (b + 2 * y0) / Q
    return y0, b, c, Q, c0, c2, c1 - c3


def xc_eps_v(rho, xc="lda,vwn"):
    """(eps_xc(rho), v_xc(rho)) pointwise; v = eps + rho d eps/d rho.

    eps restates hqc/tools/excor.py:8-12 (Slater), :76-85 (VWN), :87-112
    (HL71/BH72/GL77); v_xc is the reference's jax.grad (hqc/pbc/solver.py:673)
    written in closed form.  Pinned by: FD of rho*eps, uniform-gas values and the
    VWN high-density slope (tests/test_oracle_physics.py), the high- and
    low-density limits of HL71/BH72/GL77, which fix both constants C and A
    (tests/test_oracle_pins.py test_hl_bh_gl_limits).
    """
    xname, cname = xc.split(",")
    if xname != "lda":
        raise ValueError("exchange functional not supported yet.")
    r13 = np.cbrt(rho)
    ex = -3 / 4 * (3 / np.pi) ** (1 / 3) * r13
    eps = ex.copy()
    v = 4.0 / 3.0 * ex
    if cname == "":
        return eps, v
    rs = _WS / r13
    if cname == "vwn":
        y0, b, c, Q, c0, c2, c4 = _vwn_consts()
        y = np.sqrt(rs)
        Y = y * y + b * y + c
        ec = c0 * np.log(y * y / Y) + c4 * np.arctan(Q / (2 * y + b)) - c2 * np.log((y - y0) ** 2 / Y)
        dY = 2 * y + b
        dec = c0 * (2 / y - dY / Y) + c4 * (-Q / (2 * Y)) - c2 * (2 / (y - y0) - dY / Y)
        return eps + ec, v + ec - y / 6 * dec
    table = {"hl71": (21.0, 0.045), "hl": (21.0, 0.045), "bh72": (30.0, 0.0504),
             "bh": (30.0, 0.0504), "gl77": (11.4, 0.0666), "gl": (11.4, 0.0666)}
    if cname in ("pz81", "pz"):
        raise NameError("name 'pw91' is not defined")   # hqc/tools/excor.py:116-117
    if cname not in table:
        raise ValueError("correlation functional not supported yet.")
    Ac, Cc = table[cname]
    x = rs / Ac
    lg = np.log(1 + 1 / x)
    ec = -0.5 * Cc * ((1 + x ** 3) * lg + x / 2 - x ** 2 - 1 / 3)
    dec = -0.5 * Cc * (3 * x ** 2 * lg - 1 / x + 1.5 - 3 * x)
    return eps + ec, v + ec - x / 3 * dec


# --------------------------------------------------------------------------
# GGA (PBE) exchange-correlation on the grid.  Not in the reference, which
# accepts only LDA exchange (hqc/tools/excor.py:14-17).
# --------------------------------------------------------------------------
PBE_KAPPA = 0.804
PBE_BETA = 0.06672455060314922
PBE_MU = PBE_BETA * np.pi ** 2 / 3                  # 0.2195149727645171
PBE_GAMMA = (1 - np.log(2)) / np.pi ** 2
PW92 = dict(A=0.0310907, a1=0.21370, b1=7.5957, b2=3.5876, b3=1.6382, b4=0.49294)
RHO_CUT = 1e-12                                     # e = v = 0 below this density
GGA_NAMES = {"pbe", "gga_x_pbe", "gga_c_pbe"}


def is_gga(xc):
    xname, cname = xc.split(",")
    if xname in GGA_NAMES or cname in GGA_NAMES:
        if not (xname in GGA_NAMES and cname in GGA_NAMES):
            raise ValueError("GGA needs PBE exchange and PBE correlation ('pbe,pbe').")
        return True
    return False


def pw92_eps_c(rs):
    """Perdew-Wang 1992 correlation energy per electron, unpolarised
    (PRB 45, 13244, eq. 10 with p = 1); complex-step safe."""
    A, a1 = PW92["A"], PW92["a1"]
    q = 2 * A * (PW92["b1"] * rs ** 0.5 + PW92["b2"] * rs + PW92["b3"] * rs ** 1.5
                 + PW92["b4"] * rs ** 2)
    return -2 * A * (1 + a1 * rs) * np.log(1 + 1 / q)


def pbe_x(rho, sigma):
    """PBE exchange energy per volume rho eps_x^LDA F_x(s), s^2 = sigma / (2 k_F rho)^2."""
    kF = (3 * np.pi ** 2 * rho) ** (1 / 3)
    ex_lda = -0.75 * (3 / np.pi) ** (1 / 3) * rho ** (1 / 3)
    s2 = sigma / (4 * kF ** 2 * rho ** 2)
    Fx = 1 + PBE_KAPPA - PBE_KAPPA / (1 + PBE_MU * s2 / PBE_KAPPA)
    return rho * ex_lda * Fx


def pbe_c(rho, sigma):
    """PBE correlation energy per volume rho (eps_c^PW92 + H(t)),
    t^2 = sigma / (2 k_s rho)^2, k_s = sqrt(4 k_F / pi) (unpolarised, phi = 1)."""
    kF = (3 * np.pi ** 2 * rho) ** (1 / 3)
    rs = (3 / (4 * np.pi * rho)) ** (1 / 3)
    ec = pw92_eps_c(rs)
    ks = (4 * kF / np.pi) ** 0.5
    t2 = sigma / (4 * ks ** 2 * rho ** 2)
    bg = PBE_BETA / PBE_GAMMA
    Ac = bg / (np.exp(-ec / PBE_GAMMA) - 1)
    H = PBE_GAMMA * np.log(1 + bg * t2 * (1 + Ac * t2) / (1 + Ac * t2 + Ac ** 2 * t2 ** 2))
    return rho * (ec + H)


def pbe_energy_density(rho, sigma):
    """e(rho, sigma) = pbe_x + pbe_c per volume, PBE (Perdew, Burke, Ernzerhof,
    PRL 77, 3865 (1996)), unpolarised, sigma = |grad rho|^2.  Written with **
    powers so that it accepts complex arguments (complex-step derivatives).
    Pinned by: the LDA limit sigma -> 0, the gradient expansions F_x = 1 + mu s^2
    and H = beta t^2, the bounds F_x -> 1 + kappa and eps_c + H -> 0
    (tests/test_oracle_gga.py)."""
    return pbe_x(rho, sigma) + pbe_c(rho, sigma)


def pbe_derivs(rho, sigma):
    """(e, de/drho, de/dsigma) by complex-step differentiation (exact to rounding:
    Im f(x + ih)/h with h = 1e-30 x); zero below RHO_CUT."""
    rho = np.asarray(rho, float)
    sigma = np.asarray(sigma, float)
    ok = rho > RHO_CUT
    r = np.where(ok, rho, 1.0)
    sg = np.where(ok, sigma, 0.0)
    h = 1e-30
    e = pbe_energy_density(r, sg)
    vr = pbe_energy_density(r + 1j * h * r, sg).imag / (h * r)
    hs = h * np.maximum(sg, 1e-20)
    vs = pbe_energy_density(r, sg + 1j * hs).imag / hs
    z = np.zeros_like(rho)
    return np.where(ok, e, z), np.where(ok, vr, z), np.where(ok, vs, z)


def spectral_grad(sysm, f):
    """(3, N) gradient of a real grid function by FFT: IFFT(iG F(G)), exact for
    the band-limited interpolant on the odd grid (real, no Nyquist term)."""
    g = sysm.g
    F = np.fft.fftn(f.reshape(g, g, g))
    return np.stack([np.fft.ifftn(1j * sysm.G[..., c] * F).real.ravel() for c in range(3)])


def spectral_div(sysm, w):
    g = sysm.g
    acc = np.zeros((g, g, g), complex)
    for c in range(3):
        acc += 1j * sysm.G[..., c] * np.fft.fftn(w[c].reshape(g, g, g))
    return np.fft.ifftn(acc).real.ravel()


def xc_gga(sysm, rho):
    """(E_xc, v_xc) of the discrete PBE functional E = dV sum_r e(rho_r, |D rho|_r^2)
    with D the spectral derivative: v_xc = de/drho - D.(2 de/dsigma D rho), the exact
    derivative dE/drho_r / dV (D^T = -D).  Pinned by: finite differences of E in one
    grid value (tests/test_oracle_gga.py)."""
    grad = spectral_grad(sysm, rho)
    sigma = (grad ** 2).sum(0)
    e, vr, vs = pbe_derivs(rho, sigma)
    return e.sum() * sysm.dV, vr - spectral_div(sysm, 2 * vs * grad)


# --------------------------------------------------------------------------
# Occupations and chemical potential (hqc/tools/smearing.py)
# --------------------------------------------------------------------------
def make_occupation(n, n_mo, smearing=False, smearing_method="fermi", smearing_sigma=0.0,
                    search_method="bisect", search_cycle=100, search_tol=1e-7):
    """Restates make_occupation_func (hqc/tools/smearing.py:4-173), including the
    bisection return rule of :102 and the clipped Fermi function of :40.
    Pinned by: sum f = n, sigma -> 0 gives aufbau, entropy closed form
    (tests/test_oracle_physics.py), and hand-computed bisection end points for
    mu > 0, mu < 0, a bracket around 0 and loop == search_cycle
    (tests/test_oracle_pins.py test_bisection_return_rule).
    """
    assert n_mo >= n // 2
    if not smearing:
        occ = 2.0 * np.concatenate([np.ones(n // 2), np.zeros(n_mo - n // 2)])
        return lambda w: occ.copy()
    assert smearing_method in ("fermi", "gauss")
    sig = smearing_sigma
    if smearing_method == "fermi":
        f_func = lambda mu, w: 2.0 / (np.exp(np.clip((w - mu) / sig, -50, 50)) + 1.0)
        shift = -sig * np.log(2 * n_mo / n - 1)
    else:
        f_func = lambda mu, w: 0.5 * (1 + scipy.special.erf((mu - w) / sig))
        shift = -sig * scipy.special.erfinv(1 - n / n_mo)
    N = lambda mu, w: np.sum(f_func(mu, w)) - n

    def bisect(w):
        lo, hi = np.min(w) + shift, np.max(w) + shift
        ylo, yhi, loop = -1.0, 1.0, 0
        while ylo < -search_tol and yhi > search_tol and loop < search_cycle:
            mid = (lo + hi) / 2
            ym = N(mid, w)
            lo, hi = (mid, hi) if ym < 0 else (lo, mid)
            ylo, yhi, loop = N(lo, w), N(hi, w), loop + 1
        return ((lo > -search_tol) * lo + (hi < search_tol) * hi
                + (lo + hi) / 2 * (loop == search_cycle))

    def dN(mu, w):
        # first and second mu-derivatives of N; zero where the clip is active
        x = (w - mu) / sig
        inside = np.abs(x) < 50
        ex = np.exp(np.clip(x, -50, 50))
        if smearing_method == "fermi":
            d1 = np.where(inside, 2 * ex / (ex + 1) ** 2 / sig, 0.0)
            d2 = np.where(inside, -2 * ex * (1 - ex) / (ex + 1) ** 3 / sig ** 2, 0.0)
        else:
            t = (mu - w) / sig
            d1 = np.exp(-t * t) / np.sqrt(np.pi) / sig
            d2 = -2 * t * np.exp(-t * t) / np.sqrt(np.pi) / sig ** 2
        return d1.sum(), d2.sum()

    def newton(w):
        mu_old, mu, loop = 1.0, w[n // 2], 0
        while abs(mu - mu_old) > search_tol and loop < search_cycle:
            y = N(mu, w)
            d1, d2 = dN(mu, w)
            g1 = 2 * y * d1
            g2 = 2 * (d1 * d1 + y * d2)
            mu_old, mu, loop = mu, mu - g1 / abs(g2), loop + 1
        return mu

    search = bisect if search_method == "bisect" else newton

    def occupation(w):
        return f_func(search(w), w)
    occupation.mu = search
    return occupation


def entropy(occ, eps=1e-15):
    """S_e (hqc/pbc/solver.py:579-590)."""
    h = occ / 2
    with np.errstate(divide="ignore", invalid="ignore"):
        a = np.where(h > eps, h * np.log(np.where(h > 0, h, 1)), 0.0)
        b = np.where(1 - h > eps, (1 - h) * np.log(np.where(1 - h > 0, 1 - h, 1)), 0.0)
    return -2 * (np.sum(a) + np.sum(b))


# --------------------------------------------------------------------------
# SCF (hqc/pbc/scf.py)
# --------------------------------------------------------------------------
def _eigh_orth(F, v):
    """eigh of v^H F v as jnp.linalg.eigh does it (symmetrize_input=True: the Hermitian
    part (A + A^H)/2 is decomposed); this matters at k-points, where the DIIS
    extrapolation with complex coefficients leaves F slightly non-Hermitian."""
    A = np.conj(v.T) @ F @ v
    w, c = np.linalg.eigh(0.5 * (A + np.conj(A.T)))
    return w, v @ c


def scf_fixed_point(v, Hcore, dm0, hartree_fn, xc_fn, dm_fn, tol, max_cycle):
    """Restates fixed_point_scf (hqc/pbc/scf.py:21-88): carry starts (0, 1); the
    returned E belongs to the density that entered the last iteration."""
    E_old, E, dm, loop = 0.0, 1.0, dm0, 0
    mo, w1 = None, None
    while abs(E - E_old) > tol and loop < max_cycle:
        J = hartree_fn(dm)
        Exc, Vxc = xc_fn(dm)
        E_new = 0.5 * np.einsum("pq,qp", 2 * Hcore + J, dm).real + Exc
        w1, mo = _eigh_orth(Hcore + J + Vxc, v)
        dm = dm_fn(mo, w1)
        E_old, E, loop = E, E_new, loop + 1
    return mo, w1, E, loop != max_cycle, loop


def scf_diis(v, Hcore, dm0, hartree_fn, xc_fn, dm_fn, errvec_fn, tol, max_cycle,
             diis_space=8, diis_start_cycle=1, diis_damp=0.0):
    """Restates diis_scf (hqc/pbc/scf.py:90-242): history initialised with
    diis_space copies of F(dm0), a fixed-point warm-up of diis_start_cycle +
    diis_space iterations, then Pulay DIIS with B + 1e-12 (tr B/m + 1) I and the
    bordered solve; mo/w1 come from eigh of the extrapolated F; stop when
    |dE| <= tol and ||e_last||_F <= 10 tol, or at max_cycle.
    Pinned by: the iteration count start + space + 1 of a non-interacting problem and
    the converged flag at max_cycle, the warm-up = fixed-point iterate relation, and
    the returned (mo, w1, E) of one step from the extrapolated F
    (tests/test_oracle_pins.py); E equal to the fixed-point SCF's at convergence
    (tests/test_oracle_physics.py)."""
    F0 = Hcore + hartree_fn(dm0) + xc_fn(dm0)[1]
    e0 = errvec_fn(dm0, F0)
    Fk = [F0.copy() for _ in range(diis_space)]
    ek = [e0.copy() for _ in range(diis_space)]
    E_old, E, loop = 0.0, 1.0, 0
    mo, w1 = None, None

    def step(Fin):
        w1, mo = _eigh_orth(Fin, v)
        dm = dm_fn(mo, w1)
        J = hartree_fn(dm)
        Exc, Vxc = xc_fn(dm)
        E_new = 0.5 * np.einsum("pq,qp", 2 * Hcore + J, dm).real + Exc
        F = Hcore + J + Vxc
        return mo, w1, E_new, F, errvec_fn(dm, F)

    while loop < diis_start_cycle + diis_space:
        mo, w1, E_new, F, e = step(Fk[-1])
        Fk = Fk[1:] + [F]
        ek = ek[1:] + [e]
        E_old, E, loop = E, E_new, loop + 1
    E_old = E - 1.0
    m = diis_space
    while ((abs(E - E_old) > tol) or (np.linalg.norm(ek[-1]) > 10 * tol)) and loop < max_cycle:
        e = np.array(ek)
        B = np.einsum("imn,jmn->ij", e, e)
        B = B + 1e-12 * (np.trace(B) / m + 1.0) * np.eye(m)
        h = np.zeros((m + 1, m + 1), dtype=B.dtype)
        h[:m, :m] = B
        h[:m, m] = -1.0
        h[m, :m] = -1.0
        rhs = np.zeros(m + 1, dtype=B.dtype)
        rhs[m] = -1.0
        ck = np.linalg.solve(h, rhs)[:m]
        Fg = np.einsum("k,kab->ab", ck, np.array(Fk))
        Fg = (1 - diis_damp) * Fg + diis_damp * Fk[-1]
        mo, w1, E_new, F, e = step(Fg)
        Fk = Fk[1:] + [F]
        ek = ek[1:] + [e]
        E_old, E, loop = E, E_new, loop + 1
    return mo, w1, E, loop != max_cycle, loop


# --------------------------------------------------------------------------
# Solver (hqc/pbc/solver.py:705-962)
# --------------------------------------------------------------------------
class Solver:
    """Restates make_solver's returned function (hqc/pbc/solver.py:14-978).

    solve(xp[, kpt]) -> (mo_coeff, dm, bands*Ry, E*Ry, Ki*Ry, Vep*Ry, Vee*Ry,
    Se, converged) with xp in rs units and kpt in 1/Bohr.  literal=True uses the
    rhoG/eris route of the reference; otherwise the FFT-Poisson route.
    """

    def __init__(self, n, L, rs, basis, rcut=24, tol=1e-7, max_cycle=100, grid_length=0.12,
                 diis=True, diis_space=8, diis_start_cycle=1, diis_damp=0.0, dft=False,
                 xc="lda,vwn", smearing=False, smearing_method="fermi", smearing_sigma=0.0,
                 search_method="bisect", search_cycle=100, search_tol=1e-7, gamma=True,
                 literal=False):
        self.sys = System(n, L, rs, basis, rcut=rcut, grid_length=grid_length)
        self.n, self.rs = n, rs
        self.tol, self.max_cycle = tol, max_cycle
        self.diis, self.diis_space = diis, diis_space
        self.diis_start_cycle, self.diis_damp = diis_start_cycle, diis_damp
        self.dft, self.xc = dft, xc
        self.gamma, self.literal = gamma, literal
        self.smearing, self.smearing_method = smearing, smearing_method
        self.sigma = smearing_sigma if smearing else 0.0
        self.gga = dft and is_gga(xc)
        if dft and not self.gga:
            xc_eps_v(np.array([0.1]), xc)     # raise early on unsupported names
        self.occupation = make_occupation(n, self.sys.n_ao, smearing, smearing_method,
                                          smearing_sigma, search_method, search_cycle,
                                          search_tol)
        self.last = {}

    def _dm(self, mo, w1):
        occ = self.occupation(w1)
        return (mo * occ) @ np.conj(mo).T

    def solve(self, xp, kpt=None):
        s = self.sys
        assert xp.shape[0] == self.n
        xpb = np.asarray(xp, float) * self.rs
        k = None if self.gamma else np.asarray(kpt, float)
        S, T = overlap_kinetic(s, xpb, k)
        w, u = np.linalg.eigh(S)
        v = u @ np.diag(w ** -0.5)
        phi = ao_values(s, xpb, k)
        if self.literal:
            rhoG = rhoG_literal(s, xpb, k)
            V, eris, eris0 = literal_integrals(s, xpb, rhoG, k, hf=not self.dft)
            hartree_fn = lambda dm: hartree_literal(s, rhoG, dm)
            if not self.dft:
                hartree_fn = lambda dm: np.einsum("rs,prsq->pq", dm, eris)
                Ktot = eris + eris0
                k_fn = lambda dm: np.einsum("rs,pqsr->pq", dm, Ktot)
        else:
            vl = vloc_real(s, xpb)
            V = grid_matrix(s, phi, vl)

            def hartree_fn(dm):
                vh, _ = poisson(s, grid_density(phi, dm))
                return grid_matrix(s, phi, vh)
            if not self.dft:
                def k_fn(dm):
                    # decompose D = C f C^H exactly as the SCF built it
                    wd, cd = np.linalg.eigh(dm)
                    keep = wd > 1e-13 * np.max(np.abs(wd))
                    return exchange_fft(s, phi, cd[:, keep], wd[keep], k)
        Hcore = T + V
        w1, mo = _eigh_orth(Hcore, v)
        dm0 = self._dm(mo, w1)
        if self.dft:
            def xc_fn(dm):
                rho = grid_density(phi, dm)
                if self.gga:
                    Exc, vxc = xc_gga(s, rho)
                    return Exc, grid_matrix(s, phi, vxc)
                eps, vxc = xc_eps_v(rho, self.xc)
                return np.sum(rho * eps) * s.dV, grid_matrix(s, phi, vxc)
        else:
            def xc_fn(dm):
                Vx = -0.5 * k_fn(dm)
                return 0.5 * np.einsum("pq,qp", Vx, dm).real, Vx
        errvec = lambda dm, F: S @ dm @ F - F @ dm @ S
        if self.diis:
            mo, w1, E, conv, nit = scf_diis(v, Hcore, dm0, hartree_fn, xc_fn, self._dm, errvec,
                                            self.tol, self.max_cycle, self.diis_space,
                                            self.diis_start_cycle, self.diis_damp)
        else:
            mo, w1, E, conv, nit = scf_fixed_point(v, Hcore, dm0, hartree_fn, xc_fn, self._dm,
                                                   self.tol, self.max_cycle)
        dm = self._dm(mo, w1)
        J = hartree_fn(dm)
        Ex = xc_fn(dm)[0]
        Ki = np.einsum("pq,qp", T, dm).real
        Vep = np.einsum("pq,qp", V, dm).real
        Vee = 0.5 * np.einsum("pq,qp", J, dm).real + Ex
        occ = self.occupation(w1)
        Se = entropy(occ)
        self.last = dict(S=S, T=T, V=V, Hcore=Hcore, phi=phi, dm=dm, occ=occ, w1=w1, mo=mo,
                         n_iter=nit, xpb=xpb, E=E, kpt=k)
        return mo, dm, w1 * RY, E * RY, Ki * RY, Vep * RY, Vee * RY, Se, conv


# --------------------------------------------------------------------------
# Ewald proton-proton energy (hqc/pbc/potential.py:6-130)
# --------------------------------------------------------------------------
def ewald_kpoints(Gmax):
    """Integer vectors 0 < |G|^2 <= Gmax^2 (hqc/pbc/potential.py:6-16)."""
    r = np.arange(-Gmax, Gmax + 1)
    G = np.array(np.meshgrid(r, r, r, indexing="ij")).reshape(3, -1).T
    G2 = (G ** 2).sum(1)
    return G[(G2 <= Gmax ** 2) & (G2 > 0)].astype(float)


def madelung(kappa, G):
    """hqc/pbc/potential.py:18-33 (dim = 3)."""
    Gn = np.linalg.norm(G, axis=-1)
    gk = np.exp(-np.pi ** 2 * Gn ** 2 / kappa ** 2) / (np.pi * Gn ** 2)
    return gk.sum() - np.pi / kappa ** 2 - 2 * kappa / np.sqrt(np.pi)


def ewald_pp(xp, L, rs, kappa=10, Gmax=15, forces=False):
    """V_pp in Ry (hqc/pbc/potential.py:103-130) for xp, L in rs units; with
    forces=True also -dE/dR in Ha/Bohr (not in the reference).
    Pinned by: simple-cubic and bcc Madelung energies, kappa independence
    (tests/test_oracle_physics.py test_ewald_*)."""
    xp = np.asarray(xp, float)
    n = xp.shape[0]
    x = xp - L * np.floor(xp / L)
    i, j = np.triu_indices(n, k=1)
    rij = (x[i] - x[j]) / L
    rij = rij - np.rint(rij)
    G = ewald_kpoints(Gmax)
    Gn = np.linalg.norm(G, axis=-1)
    gk = np.exp(-np.pi ** 2 * Gn ** 2 / kappa ** 2) / (np.pi * Gn ** 2)
    d = np.linalg.norm(rij, axis=-1)
    arg = 2 * np.pi * rij @ G.T                     # (npair, nG)
    psi = scipy.special.erfc(kappa * d) / d + np.cos(arg) @ gk - np.pi / kappa ** 2
    E = 2 * psi.sum() / L / rs + n * madelung(kappa, G) / L / rs
    if not forces:
        return E
    # dpsi/drij (fractional units)
    dsr = -(scipy.special.erfc(kappa * d) / d ** 2
            + 2 * kappa / np.sqrt(np.pi) * np.exp(-kappa ** 2 * d ** 2) / d)
    grad = dsr[:, None] * rij / d[:, None] - 2 * np.pi * (np.sin(arg) * gk) @ G
    Lb = L * rs
    F = np.zeros((n, 3))
    # E_Ha = sum psi / Lb ; d rij / d x_i(Bohr) = 1/Lb
    np.add.at(F, i, -grad / Lb ** 2)
    np.add.at(F, j, grad / Lb ** 2)
    return E, F


def potential_energy(x, kappa, G, L, rs):
    """(v_pp, v_ep, v_ee) of hqc/pbc/potential.py:70-101 for one configuration x (n, 3)
    in rs units whose first n/2 rows are protons (Z = +1) and last n/2 electrons
    (Z = -1): 2 rs / L times the pair sums of psi by total pair charge, v_ep with
    a minus sign.  Pinned by: kappa independence of the neutral total
    v_pp + v_ep + v_ee + n rs/L Madelung (tests/test_oracle_slater.py)."""
    x = np.asarray(x, float)
    n = x.shape[0]
    x = x - L * np.floor(x / L)
    i, j = np.triu_indices(n, k=1)
    rij = (x[i] - x[j]) / L
    rij = rij - np.rint(rij)
    Gn = np.linalg.norm(G, axis=-1)
    gk = np.exp(-np.pi ** 2 * Gn ** 2 / kappa ** 2) / (np.pi * Gn ** 2)
    d = np.linalg.norm(rij, axis=-1)
    v = scipy.special.erfc(kappa * d) / d + np.cos(2 * np.pi * rij @ G.T) @ gk - np.pi / kappa ** 2
    Z = np.concatenate([np.ones(n // 2), -np.ones(n // 2)])
    tc = Z[i] + Z[j]
    f = 2 * rs / L
    return f * v[tc == 2].sum(), -f * v[tc == 0].sum(), f * v[tc == -2].sum()


# --------------------------------------------------------------------------
# Forces of the Mermin free energy A = E - sigma S_e + V_pp (not in the reference)
# --------------------------------------------------------------------------
def forces(solver, xp, kappa=10, Gmax=15):
    """Analytic -dA/dR (n, 3) in Ha/Bohr after solver.solve(xp[, kpt]) (Gamma or one
    k-point, DFT or HF; complex matrices enter as 2 Re Tr[conj(W) dS] etc.).

    A = E_elec - sigma S_e + E_pp (Hartree).  Terms: kinetic -Tr[dT D], Pulay
    +Tr[W dS] (W = C f eps C^T), grid adjoint dV sum_r Z_p grad phi_p with
    Z = dE/dphi (2 v_eff D phi for the local part, minus exchange pair terms),
    structure-factor derivative of v_loc, and the Ewald term.
    Pinned by: central finite differences of A, for aufbau occupations and Fermi
    smearing only (tests/test_oracle_physics.py test_forces_match_finite_differences,
    tests/test_oracle_pins.py test_hf_kpoint_forces_match_finite_differences,
    tests/test_oracle_gga.py test_gga_scf_forces_match_finite_differences and
    test_kpoint_forces_match_finite_differences).

    Gaussian smearing raises NotImplementedError: the reference's Gaussian
    occupations f = (1 + erf((mu - eps)/sigma))/2 (hqc/tools/smearing.py:52) with the
    Fermi-Dirac entropy S_e (hqc/pbc/solver.py:579-590) do not make A stationary in
    the occupations, so these terms are not -dA/dR there
    (tests/test_oracle_pins.py test_gauss_forces_are_not_the_gradient).
    """
    if solver.smearing and solver.smearing_method == "gauss":
        raise NotImplementedError("forces with Gaussian smearing: E - sigma S_e is not "
                                  "stationary in the Gaussian occupations")
    return _analytic_forces(solver, xp, kappa, Gmax)


def _analytic_forces(solver, xp, kappa=10, Gmax=15):
    """The force terms of `forces` without the smearing check."""
    s = solver.sys
    st = solver.last
    xpb = st["xpb"]
    n, nloc = s.n, s.basis.n_ao_atom
    dm, occ, w1, mo, phi = st["dm"], st["occ"], st["w1"], st["mo"], st["phi"]
    k = st.get("kpt")
    _, _, dS, dT = overlap_kinetic(s, xpb, k, deriv=True)
    W = (mo * (occ * w1)) @ np.conj(mo).T
    atom = np.repeat(np.arange(n), nloc)
    F = np.zeros((n, 3))
    other = atom[:, None] != atom[None, :]
    for c in range(3):
        tk = 2 * np.einsum("pq,pq->p", np.conj(dm * other), dT[:, :, c]).real
        ts = 2 * np.einsum("pq,pq->p", np.conj(W * other), dS[:, :, c]).real
        np.add.at(F[:, c], atom, -tk + ts)
    # grid adjoint
    rho = grid_density(phi, dm)
    vh, rg = poisson(s, rho)
    veff = vloc_real(s, xpb) + vh
    if solver.dft:
        veff = veff + (xc_gga(s, rho)[1] if solver.gga else xc_eps_v(rho, solver.xc)[1])
    X = dm @ np.conj(phi)
    Z = 2 * veff[None, :] * X
    if not solver.dft:
        # exchange: E_x = -1/4 sum_ij f_i f_j dV sum_r conj(n_ji) u_ji, n_ji = psi_j psi_i^*,
        # u = dV N IFFT[VG FFT(n)], plus -1/4 c0 Tr[S~ D S~ D], S~ = dV conj(phi) phi^T;
        # dE = -dV Re sum conj(dphi_p) Y_p, Y_p = sum_j f_j conj(C_pj) G_j,
        # G_j = sum_i f_i psi_i u_ji (and -c0 (M^T phi)_p, M = D S~ D, for G = 0)
        g = s.g
        c0 = 4 * np.pi / s.L * EXCHANGE_G0_CONST
        keep = occ > 0
        C, f = mo[:, keep], occ[keep]
        psi = C.T @ phi                                   # psi_j = sum_p C_pj phi_p
        Gm = np.zeros_like(psi)
        for i in range(len(f)):
            pair = psi * np.conj(psi[i])[None, :]
            u = np.fft.ifftn(np.fft.fftn(pair.reshape(-1, g, g, g), axes=(1, 2, 3))
                             * s.VG[None], axes=(1, 2, 3)).reshape(len(f), -1) * s.dV * s.N
            if k is None:
                u = u.real
            Gm += f[i] * psi[i][None, :] * u
        Y = np.conj(C) @ (f[:, None] * Gm)
        St = np.conj(phi) @ phi.T * s.dV
        M = dm @ St @ dm
        Z = Z - np.conj(Y) - c0 * (M @ np.conj(phi))
    gphi = ao_gradients(s, xpb, k)
    for c in range(3):
        np.add.at(F[:, c], atom, s.dV * np.einsum("pr,pr->p", Z, gphi[c]).real)
    # explicit v_loc structure-factor term: dE/dR_A = Re sum_G iG VG e^{-iG.R_A} conj(rho~)
    ph = np.exp(-1j * np.einsum("xyzd,nd->nxyz", s.G, xpb))
    for c in range(3):
        dE = np.real(np.einsum("xyz,nxyz,xyz->n", 1j * s.G[..., c] * s.VG, ph, np.conj(rg)))
        F[:, c] -= dE
    _, Fpp = ewald_pp(xp, s.L / s.rs, s.rs, kappa, Gmax, forces=True)
    return F + Fpp


def free_energy(solver, xp, kappa=10, Gmax=15, kpt=None):
    """A = E_elec - sigma S_e + V_pp in Hartree (the functional `forces` differentiates)."""
    out = solver.solve(xp, kpt)
    E = out[3] / RY
    return E - solver.sigma * out[7] + ewald_pp(xp, solver.sys.L / solver.rs, solver.rs,
                                                kappa, Gmax) / RY
