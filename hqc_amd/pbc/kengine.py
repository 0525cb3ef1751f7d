"""GPU engine for the periodic SCF at one k-point per geometry (complex Bloch AOs).

Reference: hf_kpt / dft_kpt (hqc/pbc/solver.py:769-837, 901-962) with
eval_overlap_kinetic_kpt (:257-305), eval_pbc_ao_kpt and the k tables
(hqc/pbc/gto.py:371-420), eval_vep_eris(_dft)_kpt (:481-560) restated on the
grid as at Gamma.  Each batch entry is one (geometry, k) pair, so a
Monkhorst-Pack mesh of one geometry is a batch of the same positions with
different k.

Layout: complex AO values are kept "stacked" as real rows (Re phi; Im phi) of
a (B, 2 nao_pad, ld) FP64 array, so every complex grid contraction is one real
MFMA contraction:
  V = dV phi^H diag(v) phi:  M = stacked vmat, Re V = M_rr + M_ii,
                             Im V = M_ri - M_ir;
  rho = sum_i f_i |psi_i|^2: orbital density with the real coefficient columns
                             (Re C; -Im C) and (Im C; Re C).
The SCF (DIIS with the reference's complex symmetric B, no conjugate) is the
Gamma engine's; eigh is torch's complex Hermitian solver applied to the
Hermitian part of v^H F v, as jnp.linalg.eigh does (symmetrize_input).
precision='mixed' takes the Gamma engine's defect correction with FP32 stacked rows
(complex64 grid work): V increments and the warm-up density on FP32 MFMA, FP64
anchors, accurate density, exchange and the complex128 eigensolve.
"""
import math

import torch

from .. import _lib as K
from .engine import EXCHANGE_G0_CONST, Engine, _round_up


class KEngine(Engine):
    """Batched SCF at k-points (one k per batch entry), complex Hermitian."""

    # the mixed path's low-precision operands are FP32 stacked rows (complex64 grid
    # work): 2 nao_pad rows exceed the split-FP16 and int8 kernels' 128, so V
    # increments and the warm-up density run on FP32 MFMA, the accurate density in FP64
    MIXED_SPLIT16 = False
    MIXED_Q8 = False
    MIXED_X32 = False              # the complex pair densities stay FP64
    mat_dtype = torch.complex128

    def __init__(self, cell, **kw):
        super().__init__(cell, **kw)
        self._kwork = {}

    # ------------------------------------------------------------- contractions
    def vmat_dv(self, phi, v, dV, bidx=None, vmax=None):
        """dV phi^H diag(v) phi (B, nao, nao) complex from stacked phi (B, 2 nao_pad, ld)
        (FP64; vmax is not used)."""
        c = self.c
        B = v.shape[0]
        n2 = 2 * c.nao_pad
        prec = 64 if phi.dtype == torch.float64 else 32
        need = K.vmat_workspace(B, n2, c.ld, prec)
        work = self._kwork.get(prec)
        if work is None or work.numel() * 8 < need:
            work = torch.empty(max((need + 7) // 8, 1), dtype=torch.float64, device=c.device)
            self._kwork[prec] = work
        M = torch.empty(B, n2, n2, dtype=torch.float64, device=c.device)
        K.ao_vmat(phi, v.contiguous(), dV, M, work, bidx)
        n, p = c.nao, c.nao_pad
        rr, ii = M[:, :n, :n], M[:, p:p + n, p:p + n]
        ri, ir = M[:, :n, p:p + n], M[:, p:p + n, :n]
        return torch.complex(rr + ii, ri - ir)

    def orb_density(self, phi, mo, occ, bidx=None, keep=None):
        """rho = sum_i f_i |psi_i|^2 over orbitals with f_i > OCC_CUT, psi_i = sum_p C_pi phi_p
        (keep: their largest count over the batch, if known)."""
        c = self.c
        B = mo.shape[0]
        if keep is None:
            keep = int((occ > self.OCC_CUT).sum(-1).max().item())
        keep = max(keep, 1)
        m_pad = _round_up(keep, 16)
        n, p = c.nao, c.nao_pad
        At = torch.zeros(B, 2 * p, 2 * m_pad, dtype=torch.float64, device=c.device)
        w = occ[:, None, :keep].clamp(min=0).sqrt()
        Cr, Ci = mo.real[:, :, :keep] * w, mo.imag[:, :, :keep] * w
        At[:, :n, :keep] = Cr
        At[:, p:p + n, :keep] = -Ci
        At[:, :n, m_pad:m_pad + keep] = Ci
        At[:, p:p + n, m_pad:m_pad + keep] = Cr
        rho = torch.zeros(B, c.ld, dtype=torch.float64, device=c.device)
        K.orb_density(phi, At, rho, bidx)
        return rho

    def stacked_dm(self, D):
        """Real symmetric M = [[Re D, Im D], [-Im D, Re D]] padded to 2 nao_pad, so that
        rho = rows^T M rows and X = M rows (Re X = top, -Im X = bottom half)."""
        c = self.c
        n, p = c.nao, c.nao_pad
        M = torch.zeros(D.shape[0], 2 * p, 2 * p, dtype=torch.float64, device=c.device)
        M[:, :n, :n] = D.real
        M[:, :n, p:p + n] = D.imag
        M[:, p:p + n, :n] = -D.imag
        M[:, p:p + n, p:p + n] = D.real
        return M

    # ------------------------------------------------------------------ set-up
    def setup(self, xpb, kpt=None, Lb=None, forces=False):
        if Lb is not None:
            raise NotImplementedError("k-point batches share one cell length")
        if kpt is None:
            raise ValueError("the k-point engine needs kpt (B, 3) in 1/Bohr")
        c = self.c
        B = xpb.shape[0]
        dev, f64, c128 = c.device, torch.float64, torch.complex128
        st = {"xpb": xpb.contiguous(), "kpt": kpt.contiguous(), "B": B, "hetero": False,
              "dV": c.dV}
        st["Lb"] = torch.full((B,), c.L, dtype=f64, device=dev)
        st["dVb"] = torch.full((B,), c.dV, dtype=f64, device=dev)
        S = torch.empty(B, c.nao, c.nao, dtype=c128, device=dev)
        T = torch.empty_like(S)
        dS = torch.empty(B, c.nao, c.nao, 3, dtype=c128, device=dev) if forces else None
        dT = torch.empty_like(dS) if forces else None
        K.overlap_kinetic_k(st["xpb"], st["kpt"], c.lattice, c.alpha, c.coef, c.comp, S, T,
                            dS, dT)
        st["S"], st["T"] = S, T
        if forces:
            st["dS"], st["dT"] = dS, dT
        self._orthogonalizer(st, S)                                     # solver.py:934-935
        P = torch.empty(B, c.n, 3, c.g, c.na, c.nl, dtype=c128, device=dev)
        K.gto_tables_k(st["xpb"], st["kpt"], c.g, c.L, c.tmax_img, c.alpha, c.nl, P)
        phi = torch.zeros(B, 2 * c.nao_pad, c.ld, dtype=f64, device=dev)
        K.ao_values_k(P, c.coef, c.comp, c.nao_pad, c.ld, phi)
        st["phi"] = phi
        # as in the Gamma engine: DFT keeps phi in geometry order, slot -> geometry in gidx
        st["gidx"] = torch.arange(B, dtype=torch.int32, device=dev)
        per = ["xpb", "kpt", "Lb", "dVb", "S", "T", "v", "vloc", "V", "Hcore"]
        per.append("gidx" if self.dft else "phi")
        if self.precision == "mixed":
            # FP32 stacked rows (complex64 AO values) and the FP64 anchors of V
            st["phi_lp"] = phi.float()
            st["va"] = torch.zeros(B, c.ld, dtype=f64, device=dev)
            st["Va"] = torch.zeros(B, c.nao, c.nao, dtype=c128, device=dev)
            st["vmax"] = torch.zeros(B, dtype=f64, device=dev)
            st["va_set"] = torch.zeros(B, dtype=torch.bool, device=dev)
            st["err"] = torch.full((B,), float("inf"), dtype=f64, device=dev)
            per += ["va", "Va", "vmax", "va_set", "err"]
            if not self.dft:
                per.append("phi_lp")
        if forces:
            per += ["dS", "dT"]
        if "lindep" in st:
            per += ["lindep", "ovlp_w", "ovlp_u"]
        cb = torch.empty(B, c.g, c.g, c.g, dtype=c128, device=dev)
        K.vloc_G(st["xpb"], c.g, c.L, c.VG, cb)
        K.fft3d(cb, +1)
        vl = torch.zeros(B, c.ld, dtype=f64, device=dev)
        K.complex_real(cb, c.ld, vl)
        st["vloc"] = vl
        st["V"] = self.vmat(phi, vl)
        st["Hcore"] = T + st["V"]
        if not self.dft:
            ones = torch.zeros(B, c.ld, dtype=f64, device=dev)
            ones[:, :c.N] = 1.0
            st["Sgrid"] = self.vmat(phi, ones)
            per.append("Sgrid")
        st["per"] = per
        return st

    # ----------------------------------------------------------------- algebra
    def density_matrix(self, mo, w1):
        occ = self.occupations(w1)
        return (mo * occ.unsqueeze(-2)) @ mo.mH, occ

    def _eigh_orth(self, F, v, Cw=None, shift=None, tol=None):
        """F' = v^H F v, eigh of its Hermitian part, mo = v C' (hqc/pbc/scf.py:68-72)."""
        Fp = v.mH @ (F @ v)
        if shift is not None:
            Fp = Fp + torch.diag_embed(shift).to(Fp.dtype)
        w, C = torch.linalg.eigh(0.5 * (Fp + Fp.mH))
        return w, v @ C, C

    # ---------------------------------------------------------------- exchange
    def exchange(self, st, na, mo, occ, low=None):
        """K_pq = sum_i f_i sum_G VG conj(A_pi) A_qi + c0 (S~ D S~)_pq with
        A_pi = dV FFT[phi_p conj(psi_i)] (hqc/pbc/solver.py:521-528, 604-614): complex
        pair densities, one FFT each, contracted as stacked real rows."""
        c = self.c
        dev = c.device
        n, p, N, g = c.nao, c.nao_pad, c.N, c.g
        phi = st["phi"][:na]
        phic = torch.complex(phi[:, :n, :N], phi[:, p:p + n, :N])          # (na, nao, N)
        psi = mo.transpose(-1, -2).contiguous() @ phic                      # (na, nmo, N)
        Kmat = torch.zeros(na, n, n, dtype=torch.complex128, device=dev)
        per_orb = n * N * 16 + 2 * p * c.ld * 8
        chunk = max(1, int(self.EXCHANGE_CHUNK_BYTES // per_orb))
        VGp = torch.zeros(c.ld, dtype=torch.float64, device=dev)
        VGp[:N] = c.VG
        keep = (occ > 1e-14).cpu()
        for b in range(na):
            idx = torch.nonzero(keep[b]).flatten().to(dev)
            w = occ[b, idx].sqrt() * c.dV
            for s0 in range(0, idx.numel(), chunk):
                ii = idx[s0:s0 + chunk]
                m = ii.numel()
                z = phic[b, None] * (psi[b, ii].conj() * w[s0:s0 + m, None])[:, None, :]
                K.fft3d(z.reshape(m * n, g, g, g), -1)
                z = z.reshape(m, n, N)
                M = torch.zeros(m, 2 * p, c.ld, dtype=torch.float64, device=dev)
                M[:, :n, :N] = z.real
                M[:, p:p + n, :N] = z.imag
                del z
                Kmat[b] += self.vmat_dv(M, VGp.expand(m, c.ld).contiguous(), 1.0).sum(0)
        D = (mo * occ.unsqueeze(-2)) @ mo.mH
        Sg = st["Sgrid"][:na]
        c0 = 4 * math.pi / c.L * EXCHANGE_G0_CONST
        return Kmat + c0 * (Sg @ D @ Sg)
