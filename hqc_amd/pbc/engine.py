"""GPU engine for the periodic GPW SCF (Gamma point), batched over geometries.

Host code only orchestrates: every grid-sized operation runs in a kernel of
libhqc_amd.so (GTO tables and values, AO x grid contractions on MFMA, 3-D
FFTs, xc, occupations, lattice sums, forces); the small n_ao x n_ao algebra
(orthogonalisation, eigh, DIIS) uses torch on the same device in FP64.

Reference semantics reproduced (file:line of hqc):
  set-up          hqc/pbc/solver.py:153-207 (grid, lattice, VG), :726-745 (S^-1/2, core guess)
  V_ep, J         hqc/pbc/solver.py:391-479, 616-629 as the FFT-Poisson equivalent
  xc              hqc/pbc/solver.py:669-694, hqc/tools/excor.py
  exchange        hqc/pbc/solver.py:430-431, 592-614, 747-751 as pair-density FFTs
  occupations     hqc/tools/smearing.py, hqc/pbc/solver.py:566-590
  SCF             hqc/pbc/scf.py:21-242
  energies        hqc/pbc/solver.py:758-767, 891-899
"""
import math

import numpy as np
import torch

from .. import _lib as K
from ..basis import HBasis

RY = 2.0
EXCHANGE_G0_CONST = 0.22578495          # hqc/pbc/solver.py:404
XC_IDS = {"": 2, "vwn": 1, "hl71": 3, "hl": 3, "bh72": 4, "bh": 4, "gl77": 5, "gl": 5}


def _round_up(x, m):
    return (x + m - 1) // m * m


class Cell:
    """Static data of one (n, L, rs, basis, grid) family, host and device copies."""

    def __init__(self, n, L, rs, basis, rcut=24.0, grid_length=0.12, device="cuda"):
        self.device = torch.device(device)
        self.n, self.rs = n, rs
        self.basis = HBasis(basis)
        b = self.basis
        self.L = L * rs                                    # Bohr (solver.py:154)
        self.omega = self.L ** 3
        self.g = round(self.L / grid_length / 2) * 2 + 1   # odd (solver.py:159)
        self.N = self.g ** 3
        self.ld = _round_up(self.N, 64)
        self.dV = (self.L / self.g) ** 3
        self.nloc, self.na = b.nloc, b.na
        self.nao = n * b.nloc
        self.nao_pad = _round_up(self.nao, 16)
        self.nl = b.l_max + 1
        self.tmax_img = int(24.0 // self.L)                # gto.py:48 (rcut fixed at 24 there)
        tmax = rcut // self.L                              # solver.py:168
        nt = np.arange(-tmax, tmax + 1)
        T = np.array(np.meshgrid(nt, nt, nt, indexing="ij")).reshape(3, -1).T * self.L
        self.lattice_np = T[(T ** 2).sum(1) <= rcut ** 2]
        g = self.g
        k = np.arange(g)
        k = np.where(k <= g // 2, k, k - g).astype(float)
        G1 = k * 2 * np.pi / self.L
        G2 = G1[:, None, None] ** 2 + G1[None, :, None] ** 2 + G1[None, None, :] ** 2
        with np.errstate(divide="ignore"):
            VG = 4 * np.pi / self.omega / G2
        VG[0, 0, 0] = 0.0
        self.VG_np = VG
        dev = self.device
        f64 = torch.float64
        self.alpha = torch.tensor(b.alpha, dtype=f64, device=dev)
        self.coef = torch.tensor(b.coef, dtype=f64, device=dev).contiguous()
        self.comp = torch.tensor(b.comp, dtype=torch.int32, device=dev)
        self.lattice = torch.tensor(self.lattice_np, dtype=f64, device=dev).contiguous()
        self.VG = torch.tensor(VG.ravel(), dtype=f64, device=dev)
        self.atom_of_ao = torch.arange(self.nao, device=dev) // self.nloc


class Engine:
    """Batched Gamma-point SCF on the GPU.  precision: 'fp64' or 'fp32' (the
    AO x grid contractions run on FP32 MFMA; everything else stays FP64)."""

    def __init__(self, cell, dft=False, xc="lda,vwn", smearing=False, smearing_method="fermi",
                 smearing_sigma=0.0, search_method="bisect", search_cycle=100, search_tol=1e-7,
                 tol=1e-7, max_cycle=100, diis=True, diis_space=8, diis_start_cycle=1,
                 diis_damp=0.0, precision="fp64"):
        self.c = cell
        self.dft = dft
        if dft:
            xname, cname = xc.split(",")
            if xname != "lda":
                raise ValueError("exchange functional not supported yet.")
            if cname in ("pz81", "pz"):
                raise NameError("name 'pw91' is not defined")   # hqc/tools/excor.py:116-117
            if cname not in XC_IDS:
                raise ValueError("correlation functional not supported yet.")
            self.xc_id = XC_IDS[cname]
        else:
            self.xc_id = 0
        n, nmo = cell.n, cell.nao
        assert nmo >= n // 2                                  # smearing.py:19
        self.smearing = smearing
        if smearing:
            assert smearing_method in ("fermi", "gauss")
            self.sigma = float(smearing_sigma)
            self.method = 0 if smearing_method == "fermi" else 1
            if self.method == 0:
                self.mu_shift = -self.sigma * math.log(2 * nmo / n - 1)     # smearing.py:66
            else:
                self.mu_shift = -self.sigma * float(
                    torch.special.erfinv(torch.tensor(1 - n / nmo, dtype=torch.float64)))
            self.search = 0 if search_method == "bisect" else 1
            self.search_cycle, self.search_tol = int(search_cycle), float(search_tol)
        else:
            self.sigma = 0.0
        self.tol, self.max_cycle = float(tol), int(max_cycle)
        self.diis, self.diis_space = diis, int(diis_space)
        self.diis_start_cycle, self.diis_damp = int(diis_start_cycle), float(diis_damp)
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.precision = precision
        self._work = {}

    # ------------------------------------------------------------------ utils
    def _workspace(self, B, prec):
        key = (B, prec)
        if key not in self._work:
            nbytes = K.vmat_workspace(B, self.c.nao_pad, self.c.ld, prec)
            self._work[key] = torch.empty(max(nbytes // 8, 1), dtype=torch.float64,
                                          device=self.c.device)
        return self._work[key]

    def _pad(self, M):
        c = self.c
        if c.nao_pad == c.nao:
            return M.contiguous()
        out = torch.zeros(M.shape[0], c.nao_pad, c.nao_pad, dtype=M.dtype, device=M.device)
        out[:, :c.nao, :c.nao] = M
        return out

    def vmat(self, phi, v):
        """dV * phi diag(v) phi^T for a batch -> (B, nao, nao)."""
        c = self.c
        B = phi.shape[0]
        prec = 64 if phi.dtype == torch.float64 else 32
        V = torch.empty(B, c.nao_pad, c.nao_pad, dtype=torch.float64, device=c.device)
        K.ao_vmat(phi, v, c.dV, V, self._workspace(B, prec))
        return V[:, :c.nao, :c.nao]

    def density(self, phi, D, X=None):
        c = self.c
        rho = torch.zeros(D.shape[0], c.ld, dtype=torch.float64, device=c.device)
        K.ao_density(phi, self._pad(D), rho, X)
        return rho

    def poisson(self, rho, save=False):
        """v_H on the grid (complex buffer, real part = v_H) and optionally rho~(G)."""
        c = self.c
        B = rho.shape[0]
        cbuf = torch.empty(B, c.g, c.g, c.g, dtype=torch.complex128, device=c.device)
        K.real_to_complex(rho, c.g, c.dV, cbuf)
        K.fft3d(cbuf, -1)
        saved = torch.empty_like(cbuf) if save else None
        K.poisson_mul(cbuf, c.VG, saved)
        K.fft3d(cbuf, +1)
        return cbuf, saved

    # ----------------------------------------------------------------- set-up
    def setup(self, xpb):
        """Per-geometry constants for xp (B, n, 3) in Bohr on the device."""
        c = self.c
        B = xpb.shape[0]
        dev, f64 = c.device, torch.float64
        st = {"xpb": xpb.contiguous(), "B": B}
        S = torch.empty(B, c.nao, c.nao, dtype=f64, device=dev)
        T = torch.empty_like(S)
        K.overlap_kinetic(st["xpb"], c.lattice, c.alpha, c.coef, c.comp, S, T)
        w, u = torch.linalg.eigh(S)                                     # solver.py:729-730
        st["S"], st["T"] = S, T
        st["v"] = u * w.rsqrt().unsqueeze(-2)
        P = torch.empty(B, c.n, 3, c.g, c.na, c.nl, dtype=f64, device=dev)
        K.gto_tables(st["xpb"], c.g, c.L, c.tmax_img, c.alpha, c.nl, P)
        phi = torch.zeros(B, c.nao_pad, c.ld, dtype=f64, device=dev)
        K.ao_values(P, c.coef, c.comp, c.nao_pad, c.ld, phi)
        st["phi"] = phi
        if self.precision == "fp32":
            phi32 = torch.zeros(B, c.nao_pad, c.ld, dtype=torch.float32, device=dev)
            K.ao_values(P, c.coef, c.comp, c.nao_pad, c.ld, phi32)
            st["phi_c"] = phi32
        else:
            st["phi_c"] = phi
        cb = torch.empty(B, c.g, c.g, c.g, dtype=torch.complex128, device=dev)
        K.vloc_G(st["xpb"], c.g, c.L, c.VG, cb)
        K.fft3d(cb, +1)
        vl = torch.zeros(B, c.ld, dtype=f64, device=dev)
        K.complex_real(cb, c.ld, vl)
        st["vloc"] = vl
        st["V"] = self.vmat(st["phi_c"], vl)
        st["Hcore"] = T + st["V"]
        if not self.dft:
            ones = torch.zeros(B, c.ld, dtype=f64, device=dev)
            ones[:, :c.N] = 1.0
            st["Sgrid"] = self.vmat(phi, ones)           # dV phi phi^T (exchange G=0 term)
        return st

    # ------------------------------------------------------------- occupations
    def occupations(self, w1):
        c = self.c
        B = w1.shape[0]
        if not self.smearing:
            occ = torch.zeros(B, c.nao, dtype=torch.float64, device=c.device)
            occ[:, :c.n // 2] = 2.0                                     # smearing.py:171
            return occ
        occ = torch.empty_like(w1)
        mu = torch.empty(B, dtype=torch.float64, device=c.device)
        K.occupations(w1.contiguous(), c.n, self.sigma, self.method, self.search,
                      self.search_cycle, self.search_tol, self.mu_shift, occ, mu)
        return occ

    def density_matrix(self, mo, w1):
        occ = self.occupations(w1)
        return (mo * occ.unsqueeze(-2)) @ mo.transpose(-1, -2), occ

    @staticmethod
    def entropy(occ, eps=1e-15):
        """hqc/pbc/solver.py:579-590."""
        h = occ / 2
        a = torch.where(h > eps, h * torch.log(torch.where(h > 0, h, torch.ones_like(h))), 0.0)
        b = torch.where(1 - h > eps, (1 - h) * torch.log(torch.where(1 - h > 0, 1 - h,
                                                                         torch.ones_like(h))), 0.0)
        return -2 * (a.sum(-1) + b.sum(-1))

    # ----------------------------------------------------------- two-electron
    def exchange(self, st, idx, mo, occ):
        """K (B, nao, nao) from pair-density FFTs (equals sum_rs D_rs (eris+eris0)[p,q,s,r])."""
        c = self.c
        phi = st["phi"][idx, :c.nao, :c.N]
        B = phi.shape[0]
        Kmat = torch.zeros(B, c.nao, c.nao, dtype=torch.float64, device=c.device)
        psi = mo.transpose(-1, -2) @ phi                                # (B, nmo, N)
        keep = occ > 1e-14
        g = c.g
        for b in range(B):
            for i in torch.nonzero(keep[b]).flatten().tolist():
                pair = (phi[b] * psi[b, i]).contiguous()                 # (nao, N)
                cb = pair.to(torch.complex128).reshape(c.nao, g, g, g).contiguous()
                cb *= c.dV
                K.fft3d(cb, -1)
                cb *= c.VG.reshape(1, g, g, g)
                K.fft3d(cb, +1)
                u = cb.real.reshape(c.nao, c.N)
                Kmat[b] += occ[b, i] * c.dV * (pair @ u.transpose(0, 1))
        D = (mo * occ.unsqueeze(-2)) @ mo.transpose(-1, -2)
        Sg = st["Sgrid"][idx]
        c0 = 4 * math.pi / c.L * EXCHANGE_G0_CONST
        return Kmat + c0 * (Sg @ D @ Sg)

    def fock2e(self, st, idx, D, mo=None, occ=None):
        """(J + V_xc or J - K/2, E_H + E_xc or E_H + E_x) for the geometries idx."""
        c = self.c
        if idx is not None and idx.numel() == st["B"]:
            idx = None                      # whole batch: no gather of phi
        phi_c = st["phi_c"][idx] if idx is not None else st["phi_c"]
        rho = self.density(phi_c, D)
        vh, _ = self.poisson(rho)
        B = D.shape[0]
        nblk = K.xc_blocks(c.g)
        v = torch.zeros(B, c.ld, dtype=torch.float64, device=c.device)
        part = torch.empty(B, nblk, 2, dtype=torch.float64, device=c.device)
        K.xc_eval(rho, vh, c.g, self.xc_id, v, part)
        E = part.sum(1)
        Exc, EH = E[:, 0] * c.dV, 0.5 * E[:, 1] * c.dV
        V2 = self.vmat(phi_c, v)
        if self.dft:
            return V2, EH + Exc
        Vx = -0.5 * self.exchange(st, idx if idx is not None else slice(None), mo, occ)
        Ex = 0.5 * torch.einsum("bpq,bqp->b", Vx, D)
        return V2 + Vx, EH + Ex

    # --------------------------------------------------------------------- SCF
    @staticmethod
    def _eigh_orth(F, v):
        w, cc = torch.linalg.eigh(v.transpose(-1, -2) @ F @ v)
        return w, v @ cc

    def scf(self, st):
        """hqc/pbc/scf.py: returns mo (B,nao,nmo), w1 (B,nmo), E (B) Hartree,
        converged (B) bool, n_iter (B)."""
        c = self.c
        B = st["B"]
        v, H, S = st["v"], st["Hcore"], st["S"]
        w1, mo = self._eigh_orth(H, v)                                   # solver.py:740-745
        D, occ = self.density_matrix(mo, w1)
        allidx = torch.arange(B, device=c.device)
        tol = self.tol

        def step(Fin, idx):
            w1, mo = self._eigh_orth(Fin, v[idx])
            D, occ = self.density_matrix(mo, w1)
            V2, E2 = self.fock2e(st, idx, D, mo, occ)
            Enew = torch.einsum("bpq,bqp->b", H[idx], D) + E2
            F = H[idx] + V2
            e = S[idx] @ D @ F - F @ D @ S[idx]
            return mo, w1, Enew, F, e

        if not self.diis:                                                 # scf.py:21-88
            E_old = torch.zeros(B, dtype=torch.float64, device=c.device)
            E = torch.ones_like(E_old)
            Dc, Oc = D.clone(), occ.clone()
            Mo, W1 = mo.clone(), w1.clone()
            nit = torch.zeros(B, dtype=torch.long, device=c.device)
            act = allidx
            loop = 0
            while act.numel() > 0:
                V2, E2 = self.fock2e(st, act, Dc[act], Mo[act], Oc[act])
                Enew = torch.einsum("bpq,bqp->b", H[act], Dc[act]) + E2
                w1n, mon = self._eigh_orth(H[act] + V2, v[act])
                Dn, On = self.density_matrix(mon, w1n)
                E_old[act], E[act] = E[act], Enew
                Dc[act], Oc[act], Mo[act], W1[act] = Dn, On, mon, w1n
                loop += 1
                nit[act] = loop
                keep = ((E[act] - E_old[act]).abs() > tol) & (loop < self.max_cycle)
                act = act[keep]
            return Mo, W1, E, nit != self.max_cycle, nit

        m = self.diis_space
        V2, E2 = self.fock2e(st, None, D, mo, occ)                        # scf.py:122-128
        F0 = H + V2
        e0 = S @ D @ F0 - F0 @ D @ S
        Fk = F0.unsqueeze(1).repeat(1, m, 1, 1)
        ek = e0.unsqueeze(1).repeat(1, m, 1, 1)
        E = torch.ones(B, dtype=torch.float64, device=c.device)
        Mo, W1 = mo, w1
        for _ in range(self.diis_start_cycle + m):                        # scf.py:130-172
            Mo, W1, Enew, F, e = step(Fk[:, -1], allidx)
            Fk = torch.cat([Fk[:, 1:], F.unsqueeze(1)], 1)
            ek = torch.cat([ek[:, 1:], e.unsqueeze(1)], 1)
            E = Enew
        loop = self.diis_start_cycle + m
        E_old = E - 1.0
        nit = torch.full((B,), loop, dtype=torch.long, device=c.device)
        Mo, W1 = Mo.clone(), W1.clone()

        def cont(idx):
            err = torch.linalg.norm(ek[idx, -1], dim=(-2, -1))
            return (((E[idx] - E_old[idx]).abs() > tol) | (err > 10 * tol)) & (loop < self.max_cycle)

        act = allidx[cont(allidx)]
        eye = torch.eye(m, dtype=torch.float64, device=c.device)
        while act.numel() > 0:                                            # scf.py:174-239
            e = ek[act]
            Bm = torch.einsum("bimn,bjmn->bij", e, e)
            eta = 1e-12 * (Bm.diagonal(dim1=-2, dim2=-1).sum(-1) / m + 1.0)
            Bm = Bm + eta[:, None, None] * eye
            nb = act.numel()
            h = torch.zeros(nb, m + 1, m + 1, dtype=torch.float64, device=c.device)
            h[:, :m, :m] = Bm
            h[:, :m, m] = -1.0
            h[:, m, :m] = -1.0
            rhs = torch.zeros(nb, m + 1, 1, dtype=torch.float64, device=c.device)
            rhs[:, m, 0] = -1.0
            ck = torch.linalg.solve(h, rhs)[:, :m, 0]
            Fg = torch.einsum("bk,bkpq->bpq", ck, Fk[act])
            Fg = (1 - self.diis_damp) * Fg + self.diis_damp * Fk[act, -1]
            mo_a, w1_a, Enew, F, e_new = step(Fg, act)
            Fk[act] = torch.cat([Fk[act, 1:], F.unsqueeze(1)], 1)
            ek[act] = torch.cat([ek[act, 1:], e_new.unsqueeze(1)], 1)
            E_old[act] = E[act]
            E[act] = Enew
            Mo[act], W1[act] = mo_a, w1_a
            loop += 1
            nit[act] = loop
            act = act[cont(act)]
        return Mo, W1, E, nit != self.max_cycle, nit

    # ---------------------------------------------------------------- results
    def solve(self, xpb):
        """Full solver output for a batch (B, n, 3) Bohr: dict of tensors (Hartree)."""
        st = self.setup(xpb)
        mo, w1, E, conv, nit = self.scf(st)
        D, occ = self.density_matrix(mo, w1)                              # solver.py:758-765
        idx = torch.arange(st["B"], device=self.c.device)
        V2, E2 = self.fock2e(st, None, D, mo, occ)
        Ki = torch.einsum("bpq,bqp->b", st["T"], D)
        Vep = torch.einsum("bpq,bqp->b", st["V"], D)
        Vee = E2
        Se = self.entropy(occ)
        st.update(mo=mo, w1=w1, D=D, occ=occ, E=E, conv=conv, nit=nit, Ki=Ki, Vep=Vep, Vee=Vee,
                  Se=Se, idx=idx)
        return st
