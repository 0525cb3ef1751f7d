"""Forces by a hand-written adjoint of the grid contractions (new: the
reference has none, doc/solver_algorithm.md:1210).

The differentiated functional is the Mermin free energy
    A = E_elec - sigma * S_e + E_pp            (Hartree; sigma = 0 without smearing)
which is stationary in the orbitals and occupations at SCF convergence, so
    -dA/dR_A = -Tr[dT/dR_A D] + Tr[W dS/dR_A]                  kinetic + Pulay
               + dV sum_r sum_{p in A} Z_p(r) grad phi_p(r)      grid adjoint
               - d/dR_A (structure factor of v_loc) . rho~       Hellmann-Feynman
               - dE_pp/dR_A                                      Ewald
with W = C diag(f eps) C^T and Z = dE/dphi / dV: 2 v_eff D phi for the local
terms, minus the exchange pair terms for HF.  grad phi is recomputed from the
1-D tables inside the kernel (never stored).
"""
import math

import torch

from .. import _lib as K
from .engine import EXCHANGE_G0_CONST
from .potential import ewald_pp


def check_supported(engine, kpt=False):
    """Raise before any device work where the force terms are not -dA/dR.

    Gaussian smearing: the reference's occupations f = (1 + erf((mu - eps)/sigma))/2
    (hqc/tools/smearing.py:52) together with the Fermi-Dirac entropy S_e
    (hqc/pbc/solver.py:579-590) do not make A = E - sigma S_e stationary in the
    occupations, so the Hellmann-Feynman/Pulay terms here miss an occupation
    response and would not be the gradient of A."""
    if engine.smearing and engine.method == 1:
        raise NotImplementedError("forces with Gaussian smearing: E - sigma*S_e is not "
                                  "stationary in the reference's Gaussian occupations")


def _cut_basis_term(st, D, F):
    """Pulay-type term of a basis with overlap directions cut (ovlp_cut > 0).

    The SCF is variational in the subspace of the overlap eigenvectors u_k with
    s_k > ovlp_cut, V = U_k s_k^-1/2.  When a nucleus moves, that subspace turns:
    first order perturbation gives u_c^H du_k = u_c^H dS u_k / (s_k - s_c) for a cut
    direction c, so dE/dR gains 2 Re Tr[V^H F U_c B D'] with B = U_c^H dV and
    V D' = D S V.  Written as Re Tr[X dS/dR], X = U N U^H with
        N_kc = 2 (U^H S D F U)_kc / (s_k (s_k - s_c))   (k kept, c cut; else 0),
    it enters like the energy-weighted density: W_eff = W - (X + X^H)/2.  Without a
    cut the term is absent (the full space does not turn)."""
    w, U = st["ovlp_w"], st["ovlp_u"]
    cut = st["lindep"] > 0
    S = st["S"]
    G = U.mH @ S @ D @ F @ U                                     # rows k, columns c
    wk, wc = w.unsqueeze(-1), w.unsqueeze(-2)
    mask = (~cut).unsqueeze(-1) & cut.unsqueeze(-2)
    den = torch.where(mask, wk * (wk - wc), torch.ones_like(wk * wc))
    N = torch.where(mask, 2.0 * G / den, torch.zeros_like(G))
    X = U @ N @ U.mH
    return 0.5 * (X + X.mH)


ADJOINT_BLOCK_BYTES = 4 << 30      # pair potentials of one block of orbital columns


def _exchange_adjoint(engine, st):
    """Zx = -Y - c0 (D S~ D) phi for HF, (B, nao_pad, ld), each geometry with its own
    cell constants (VG, dV, c0):
        Y_p = sum_j f_j C_pj G_j,   G_j = sum_i f_i psi_i u_ij,
        u_ij = dV N IFFT[VG FFT(psi_i psi_j)]  (real and symmetric in i, j).
    Only the pairs i <= j are transformed, two real pair densities per complex grid
    (VG is real and even, so the two potentials come back as Re and Im); a pair adds
    f_i psi_i u_ij to G_j and, off the diagonal, f_j psi_j u_ij to G_i.  The columns
    j run in blocks whose pair potentials fit ADJOINT_BLOCK_BYTES; every sum is a
    fixed-order reduction."""
    c = engine.c
    B, N, g = st["B"], c.N, c.g
    dev = c.device
    phi = st["phi"][:, :c.nao, :N]
    mo, occ = st["slot"]["mo"], st["slot"]["occ"]
    Zx = torch.zeros(B, c.nao_pad, c.ld, dtype=torch.float64, device=dev)
    Lh = st["Lb"].cpu().tolist()
    keep_h = (occ > 1e-14).cpu()
    for b in range(B):
        cst = c.consts(Lh[b])
        VG = cst.VG.reshape(1, g, g, g)
        keep = torch.nonzero(keep_h[b]).flatten()
        m = keep.numel()
        C = mo[b][:, keep.to(dev)]
        f = occ[b][keep.to(dev)]
        psi = C.transpose(0, 1) @ phi[b]                       # (m, N)
        fpsi = f[:, None] * psi
        G = torch.zeros(m, N, dtype=torch.float64, device=dev)
        jb = max(1, min(m, int(ADJOINT_BLOCK_BYTES // (3 * 8 * m * N))))
        for j0 in range(0, m, jb):
            j1 = min(m, j0 + jb)
            ii, jj = [], []
            for j in range(j0, j1):
                ii += list(range(j + 1))
                jj += [j] * (j + 1)
            P = len(ii)
            it = torch.tensor(ii, device=dev)
            jt = torch.tensor(jj, device=dev)
            n = psi[it] * psi[jt] * cst.dV                     # (P, N) pair densities
            if P % 2:
                n = torch.cat([n, torch.zeros_like(n[:1])])
            z = torch.complex(n[0::2], n[1::2]).reshape(-1, g, g, g)
            del n
            K.fft3d(z, -1)
            z *= VG
            K.fft3d(z, +1)
            u = torch.stack([z.real, z.imag], 1).reshape(-1, N)[:P]
            del z
            ub = torch.zeros(m, j1 - j0, N, dtype=torch.float64, device=dev)
            ub[it, jt - j0] = u                                # u_ij for i <= j, j in block
            del u
            G[j0:j1] += (fpsi[:, None, :] * ub).sum(0)          # G_j += f_i psi_i u_ij
            d = torch.arange(j0, j1, device=dev)
            ub[d, d - j0] = 0.0
            G += (fpsi[None, j0:j1, :] * ub).sum(1)             # G_i += f_j psi_j u_ij, j > i
            del ub
        Y = C @ (f[:, None] * G)
        D = st["slot"]["D"][b]
        M = D @ st["Sgrid"][b] @ D
        Zx[b, :c.nao, :N] = -Y - cst.c0 * (M @ phi[b])
    return Zx


def _exchange_adjoint_k(engine, st):
    """Zx for HF at one k-point per entry, stacked (B, 2 nao_pad, ld): with the pair
    densities n_ji = psi_j psi_i^* and u_ji = dV N IFFT[VG FFT(n_ji)],
        dE_x = -dV Re sum_r conj(dphi_p) Y_p,   Y_p = sum_j f_j conj(C_pj) G_j + c0 (M^T phi)_p,
        G_j = sum_i f_i psi_i u_ji,  M = D S~ D,  S~ = dV phi^H phi,
    and Zx = -(Re Y; Im Y) pairs with the derivatives of (Re phi; Im phi).  At Gamma
    it reduces to _exchange_adjoint's Y."""
    c = engine.c
    B, n, p, N, g = st["B"], c.nao, c.nao_pad, c.N, c.g
    dev = c.device
    mo, occ, D = st["slot"]["mo"], st["slot"]["occ"], st["slot"]["D"]
    phi = st["phi"]
    VG = c.VG.reshape(1, g, g, g)
    c0 = 4 * math.pi / c.L * EXCHANGE_G0_CONST
    Zx = torch.zeros(B, 2 * p, c.ld, dtype=torch.float64, device=dev)
    for b in range(B):
        phic = torch.complex(phi[b, :n, :N], phi[b, p:p + n, :N])
        keep = torch.nonzero(occ[b] > 1e-14).flatten()
        C = mo[b][:, keep]
        f = occ[b][keep].to(torch.complex128)
        psi = C.transpose(0, 1) @ phic                      # (m, N)
        G = torch.zeros_like(psi)
        for i in range(psi.shape[0]):
            pair = (psi * psi[i].conj()).reshape(-1, g, g, g).contiguous()
            K.fft3d(pair, -1)
            pair *= VG
            K.fft3d(pair, +1)
            G += f[i] * psi[i] * pair.reshape(-1, N) * c.dV
        Y = C.conj() @ (f[:, None] * G)
        M = D[b] @ st["Sgrid"][b] @ D[b]
        Y += c0 * (M.transpose(0, 1) @ phic)
        Zx[b, :n, :N] = -Y.real
        Zx[b, p:p + n, :N] = -Y.imag
    return Zx


def compute_forces(engine, st, kappa=10, Gmax=15):
    """-dA/dR (B, n, 3) in Ha/Bohr after engine.solve, in input order.  Works on
    the slot-ordered device state of the SCF and un-permutes at the end.  The
    cell-length dependent kernels run once per distinct cell length."""
    if st.get("kpt") is not None:
        return compute_forces_k(engine, st, kappa, Gmax)
    check_supported(engine)
    c = engine.c
    B, dev = st["B"], c.device
    f64 = torch.float64
    sl = st["slot"]
    D, occ, w1, mo = sl["D"], sl["occ"], sl["w1"], sl["mo"]
    groups = engine._cell_groups(B, st["Lb"] if st.get("hetero") else None)
    sub, put = engine._sub, engine._put
    if "dS" in st:                          # kept by a set-up with forces=True
        dS, dT = st["dS"], st["dT"]
    else:
        dS = torch.empty(B, c.nao, c.nao, 3, dtype=f64, device=dev)
        dT = torch.empty_like(dS)
        for cst, idx in groups:
            x = sub(st["xpb"], idx)
            nb = x.shape[0]
            S = torch.empty(nb, c.nao, c.nao, dtype=f64, device=dev)
            T = torch.empty_like(S)
            dSg = dS if idx is None else torch.empty(nb, c.nao, c.nao, 3, dtype=f64,
                                                     device=dev)
            dTg = dT if idx is None else torch.empty_like(dSg)
            K.overlap_kinetic(x, cst.lattice, c.alpha, c.coef, c.comp, S, T, dSg, dTg,
                              s_only=c.basis.l_max == 0)
            put(dS, idx, dSg)
            put(dT, idx, dTg)
    W = (mo * (occ * w1).unsqueeze(-2)) @ mo.transpose(-1, -2)
    if "lindep" in st:
        W = W - _cut_basis_term(st, D, sl["F"])
    atom = c.atom_of_ao
    other = (atom[:, None] != atom[None, :]).to(f64)
    t = 2.0 * ((W * other).unsqueeze(-1) * dS - (D * other).unsqueeze(-1) * dT).sum(2)
    # AOs are atom-major: a fixed-order sum per atom (index_add_ would sum with atomics)
    F = t.reshape(B, c.n, c.nloc, 3).sum(2)
    # grid adjoint (FP64 phi regardless of the SCF precision): X = D phi, v and rho~(G)
    # of the final density, from the solve's last pass when it made them (DFT)
    if "force_grid" in st:
        X, v, rg = st.pop("force_grid")
    else:
        # the D-form kernel writes every row and grid column of X and rho
        X = torch.empty(B, c.nao_pad, c.ld, dtype=f64, device=dev)
        rho = torch.empty(B, c.ld, dtype=f64, device=dev)
        K.ao_density(st["phi"], engine._pad(D), rho, X, bidx=st["gidx"])
        v, _, rg = engine.xc_potential(rho, save_rg=True, geo=engine._geo(st, B))
    veff = (v + st["vloc"]).contiguous()
    Zx = None if engine.dft else _exchange_adjoint(engine, st)
    Epp = torch.empty(B, dtype=f64, device=dev)
    nblk = K.force_blocks(c.g)
    for cst, idx in groups:
        x = sub(st["xpb"], idx)
        nb = x.shape[0]
        P = torch.empty(nb, c.n, 3, c.g, c.na, 3, dtype=f64, device=dev)
        K.gto_tables(x, c.g, cst.L, cst.tmax_img, c.alpha, 3, P)
        fp = torch.empty(nb, c.n, nblk, 3, dtype=f64, device=dev)
        K.grid_forces(P, c.alpha, c.coef, c.comp, c.nao_pad, c.ld, sub(X, idx), sub(veff, idx),
                      None if Zx is None else sub(Zx, idx), cst.dV, fp)
        dE = torch.empty(nb, c.n, 3, dtype=f64, device=dev)
        K.vloc_force(x, c.g, cst.L, cst.VG, sub(rg, idx), dE)
        # Ewald in Bohr units (rs = 1): the same sums as at (xp/rs, L/rs, rs)
        Eg, Fpp = ewald_pp(x, cst.L, 1.0, kappa, Gmax, forces=True)
        Fg = fp.sum(2) - dE + Fpp
        if idx is None:
            F += Fg
            Epp.copy_(Eg)
        else:
            F[idx] += Fg
            Epp[idx] = Eg
    st["Epp"] = st["unperm"](Epp)
    return st["unperm"](F)


def compute_forces_k(engine, st, kappa=10, Gmax=15):
    """-dA/dR at one k-point per batch entry (DFT or HF).  Complex S, T, D, W enter as
    2 Re sum_q [conj(W_pq) dS_pq - conj(D_pq) dT_pq]; the grid adjoint uses the
    stacked rows (Re phi; Im phi) with X = M rows, M = [[Re D, Im D], [-Im D, Re D]]."""
    check_supported(engine, kpt=True)
    c = engine.c
    B, dev = st["B"], c.device
    f64, c128 = torch.float64, torch.complex128
    sl = st["slot"]
    D, occ, w1, mo = sl["D"], sl["occ"], sl["w1"], sl["mo"]
    if "dS" in st:                          # kept by a set-up with forces=True
        dS, dT = st["dS"], st["dT"]
    else:
        S = torch.empty(B, c.nao, c.nao, dtype=c128, device=dev)
        T = torch.empty_like(S)
        dS = torch.empty(B, c.nao, c.nao, 3, dtype=c128, device=dev)
        dT = torch.empty_like(dS)
        K.overlap_kinetic_k(st["xpb"], st["kpt"], c.lattice, c.alpha, c.coef, c.comp, S, T,
                            dS, dT)
    W = (mo * (occ * w1).unsqueeze(-2)) @ mo.mH
    if "lindep" in st:
        W = W - _cut_basis_term(st, D, sl["F"])
    atom = c.atom_of_ao
    other = (atom[:, None] != atom[None, :]).to(f64)
    t = 2.0 * ((W * other).conj().unsqueeze(-1) * dS
               - (D * other).conj().unsqueeze(-1) * dT).sum(2).real
    # AOs are atom-major: a fixed-order sum per atom (index_add_ would sum with atomics)
    F = t.reshape(B, c.n, c.nloc, 3).sum(2)
    X = torch.empty(B, 2 * c.nao_pad, c.ld, dtype=f64, device=dev)
    rho = torch.empty(B, c.ld, dtype=f64, device=dev)
    K.ao_density(st["phi"], engine.stacked_dm(D), rho, X, bidx=st["gidx"])
    v, _, rg = engine.xc_potential(rho, save_rg=True)
    veff = (v + st["vloc"]).contiguous()
    P = torch.empty(B, c.n, 3, c.g, c.na, 3, dtype=c128, device=dev)
    K.gto_tables_k(st["xpb"], st["kpt"], c.g, c.L, c.tmax_img, c.alpha, 3, P)
    fp = torch.empty(B, c.n, K.force_blocks(c.g), 3, dtype=f64, device=dev)
    Zx = None if engine.dft else _exchange_adjoint_k(engine, st)
    K.grid_forces_k(P, c.alpha, c.coef, c.comp, c.nao_pad, c.ld, X, veff, Zx, c.dV, fp)
    F += fp.sum(2)
    dE = torch.empty(B, c.n, 3, dtype=f64, device=dev)
    K.vloc_force(st["xpb"], c.g, c.L, c.VG, rg, dE)
    F -= dE
    Epp, Fpp = ewald_pp(st["xpb"] / c.rs, c.L / c.rs, c.rs, kappa, Gmax, forces=True)
    st["Epp"] = st["unperm"](Epp)
    return st["unperm"](F + Fpp)


class FreeEnergy(torch.autograd.Function):
    """A(xp) in Ry for xp (B, n, 3) in rs units, with the analytic gradient.

    forward runs the SCF and the adjoint once; backward returns
    dA_Ry/dxp_rs = -2 rs F."""

    @staticmethod
    def forward(ctx, xp, runner, kpt=None):
        A, F = runner.free_energy_and_forces(xp, kpt)
        ctx.save_for_backward(F)
        ctx.rs = runner.rs
        return A

    @staticmethod
    def backward(ctx, gA):
        (F,) = ctx.saved_tensors
        return gA[:, None, None] * (-2.0 * ctx.rs) * F, None, None
