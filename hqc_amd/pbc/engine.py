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
from .. import tune
from ..basis import HBasis

RY = 2.0
EXCHANGE_G0_CONST = 0.22578495          # hqc/pbc/solver.py:404
GGA_NAMES = {"pbe", "gga_x_pbe", "gga_c_pbe"}    # PBE exchange + correlation (new)
XC_IDS = {"": 2, "vwn": 1, "hl71": 3, "hl": 3, "bh72": 4, "bh": 4, "gl77": 5, "gl": 5}


def _round_up(x, m):
    return (x + m - 1) // m * m


def _tr(A, B):
    """Batched Tr[A B] = sum_pq A_pq B_qp (elementwise; a batched einsum here
    lowers to a slow GEMM shape)."""
    return (A * B.transpose(-1, -2)).sum((-1, -2))


class CellConsts:
    """Cell-size dependent constants for one cell length L (Bohr) on a fixed grid."""

    def __init__(self, L, g, N, ld, rcut, device):
        f64 = torch.float64
        self.L = L
        self.omega = L ** 3
        self.dV = (L / g) ** 3
        self.tmax_img = int(24.0 // L)                   # gto.py:48 (rcut fixed at 24 there)
        tmax = rcut // L                                 # solver.py:168
        nt = np.arange(-tmax, tmax + 1)
        T = np.array(np.meshgrid(nt, nt, nt, indexing="ij")).reshape(3, -1).T * L
        self.lattice_np = T[(T ** 2).sum(1) <= rcut ** 2]
        k = np.arange(g)
        k = np.where(k <= g // 2, k, k - g).astype(float)
        G1 = k * 2 * np.pi / L
        G2 = G1[:, None, None] ** 2 + G1[None, :, None] ** 2 + G1[None, None, :] ** 2
        with np.errstate(divide="ignore"):
            VG = 4 * np.pi / self.omega / G2
        VG[0, 0, 0] = 0.0
        self.VG_np = VG
        self.lattice = torch.tensor(self.lattice_np, dtype=f64, device=device).contiguous()
        self.VG = torch.tensor(VG.ravel(), dtype=f64, device=device)
        # exchange weights on the half G-set H (centred order, columns < (N-1)/2) for the
        # [Re | Im] row blocks of the pair-density transforms: 2 VG on both halves
        nh = (N - 1) // 2
        cc = np.arange(nh)
        h = g // 2
        s1, s2, s3 = cc // (g * g) - h, (cc // g) % g - h, cc % g - h
        kk = ((s1 % g) * g + s2 % g) * g + s3 % g
        wh = np.zeros(ld)
        wh[:nh] = 2 * VG.ravel()[kk]
        wh[nh:2 * nh] = wh[:nh]
        self.VGh = torch.tensor(wh, dtype=f64, device=device)
        # the same weights for the transformed pair grids read whole as real rows
        # [Re Z, Im Z interleaved] (hqc_pair_pack_fft): VG on every G (0 at G = 0)
        self.ldz = _round_up(2 * N, 64)
        wz = np.zeros(self.ldz)
        wz[0:2 * N:2] = VG.ravel()
        wz[1:2 * N:2] = VG.ravel()
        self.VGz = torch.tensor(wz, dtype=f64, device=device)
        self.c0 = 4 * math.pi / L * EXCHANGE_G0_CONST   # eris0 factor (solver.py:431)


class Cell:
    """Static data of one (n, L, rs, basis, grid) family, host and device copies.

    Geometries of other cell lengths that land on the same odd grid (same n, g)
    can share the cell: consts(L) gives their L-dependent constants."""

    def __init__(self, n, L, rs, basis, rcut=24.0, grid_length=0.12, device="cuda"):
        self.device = torch.device(device)
        self.n, self.rs = n, rs
        self.rcut, self.grid_length = rcut, grid_length
        self.basis = HBasis(basis)
        b = self.basis
        self.L = L * rs                                    # Bohr (solver.py:154)
        self.g = self.grid_of(self.L)                      # odd (solver.py:159)
        self.N = self.g ** 3
        self.ld = _round_up(self.N, 64)
        self.nloc, self.na = b.nloc, b.na
        self.nao = n * b.nloc
        self.nao_pad = _round_up(self.nao, 16)
        self.nl = b.l_max + 1
        self._consts = {}
        cst = self.consts(self.L)
        self.omega, self.dV, self.tmax_img = cst.omega, cst.dV, cst.tmax_img
        self.lattice_np, self.lattice = cst.lattice_np, cst.lattice
        self.VG_np, self.VG, self.VGh = cst.VG_np, cst.VG, cst.VGh
        dev = self.device
        f64 = torch.float64
        self.alpha = torch.tensor(b.alpha, dtype=f64, device=dev)
        self.coef = torch.tensor(b.coef, dtype=f64, device=dev).contiguous()
        self.comp = torch.tensor(b.comp, dtype=torch.int32, device=dev)
        self.atom_of_ao = torch.arange(self.nao, device=dev) // self.nloc

    def grid_of(self, L_bohr):
        """Odd grid size of hqc/pbc/solver.py:159 (Python round, half to even)."""
        return round(L_bohr / self.grid_length / 2) * 2 + 1

    def consts(self, L_bohr):
        key = float(L_bohr)
        if key not in self._consts:
            if self.grid_of(key) != getattr(self, "g", self.grid_of(key)):
                raise ValueError(f"cell length {key} Bohr gives another grid than {self.g}^3")
            self._consts[key] = CellConsts(key, self.grid_of(key), self.grid_of(key) ** 3,
                                           _round_up(self.grid_of(key) ** 3, 64), self.rcut,
                                           self.device)
        return self._consts[key]


class Engine:
    """Batched Gamma-point SCF on the GPU.

    precision='fp64': every contraction in FP64 (MFMA f64).
    precision='mixed': defect correction.  rho(D) and V(v) are linear, so the
    engine keeps FP64 anchors (D_a, rho_a) and (v_a, V_a) and evaluates only the
    increments rho(D - D_a) and V(v - v_a) in low precision: V from FP16 (hi, lo)
    split products with FP32 accumulation on MFMA (FP32 MFMA for n_ao > 128), rho
    with FP32 AO values on FP32 MFMA.
    The increment error is proportional to the increment, which shrinks as the
    SCF converges; an FP64 re-anchor is forced whenever the estimated increment
    error exceeds MIXED_SAFETY times the current DIIS residual per matrix
    element, so the iteration converges to the FP64 fixed point.  While the DIIS
    residual is still large (>= MIXED_DIRECT_ERR, the fixed-point warm-up) rho
    and V are taken in FP32 outright.  The choice is per geometry: the kernels
    read a subset of the batch in place through a geometry index list, so each
    precision class costs one launch over its own geometries.  Eigensolves,
    DIIS, occupations, FFTs, xc and every energy sum stay FP64.

    Density increments are off by default (MIXED_RHO_INCREMENTS): the D form
    costs 2 n_ao^2 flops per grid point against 2 n_ao n_occ for the FP64
    orbital form, and with n_occ ~ n_ao / 2 the FP32 MFMA rate (2x FP64) does
    not pay that back, so past the warm-up rho is taken in the accurate orbital
    form: int8 digit products with exact integer accumulation (MIXED_Q8), or FP64."""

    # error of an FP32-accumulated V per element relative to max|v|: measured up to
    # 1.46e-6 (640 AOs, 65^3, the potential of the superposed AO densities), so 5e-6
    # keeps a 3x margin at the largest shape (tests/test_gpu_kernels.py
    # test_mixed_v_error_margin); the split-FP16 kernel stays below 1.5e-7
    MIXED_EPS = 5e-6
    MIXED_SAFETY = 1e-2     # potential increment: error per element <= 1e-2 residual / n_ao
    MIXED_RHO_SAFETY = 1.0  # density increment: |v| x relative rho error <= residual
    MIXED_DIRECT_ERR = 1e-1  # residual above which V is taken in low precision outright
    # the same for the density (orbital form): the split-FP16 density leaves a DIIS
    # residual floor of ~3e-6..1e-5 at 54 H dzv, so the switch to the accurate form
    # stays an order of magnitude above it (3e-6 stalled geometries at max_cycle;
    # 1e-3 runs within 1% of 1e-4 but put the bands of the geometries left at
    # max_cycle 5e-7 Ha off FP64 against 1e-7)
    MIXED_RHO_DIRECT_ERR = 1e-4
    MIXED_RHO_INCREMENTS = False  # FP32 D-form density increments (else FP64 orbital form)
    # low-precision V from FP16 (hi, lo) split products on MFMA (n_ao <= 128; FP32
    # MFMA beyond).  Its measured error is ~1.3e-6 of max|V|; the increment constant
    # is 3x that (54 H dzv, batch 256: 8e-6 anchors twice as often for 3% less
    # throughput, 1e-6, below the real error, stalls geometries and loses 18%)
    MIXED_SPLIT16 = True
    MIXED_EPS_SPLIT16 = 4e-6
    # past the warm-up the density comes from base-256 int8 digits of phi with exact
    # integer products (n_ao <= 128; FP64 vector FMAs beyond).  4 digits: phi is fixed
    # to 2^-30 of its per-point maximum (a basis perturbation that is the same on
    # every call).  Digit products up to level Q8_TOP = 4 (14 of the 16 pairs) add
    # ~2^-38 per AO of coefficient noise, ~1.5e-11 of rho: a Fock-matrix noise ~1e-11
    # Ha, three orders below the SCF tolerance, and ~5e-10 Ha in the energy, so no
    # residual floor appears and level 4 serves to convergence (measured at 512
    # geometries of 54 H: the same iteration counts and the same differences to FP64
    # and to the oracle as level 5, 18 pairs, ~2^-46, at 14/18 of its MFMA time).
    # While the DIIS residual is >= MIXED_Q8_COARSE_ERR the products stop at level
    # Q8_TOP_COARSE = 3 (10 pairs, ~2^-30 per AO: three orders below such a residual)
    MIXED_Q8 = True
    Q8_SLICES = 4
    Q8_TOP = 4
    Q8_TOP_COARSE = 3
    MIXED_Q8_COARSE_ERR = 1e-6
    # the warm-up density (DIIS residual >= MIXED_RHO_DIRECT_ERR) from the first 3
    # digits (2^-22 per value, a fixed basis with no residual floor; 0: the
    # split-FP16 form) and digit products to level 3 (9 products; ~2^-30 per AO of
    # coefficient noise, five orders below such a residual)
    Q8_WARMUP_DIGITS = 3
    Q8_WARMUP_TOP = 3
    mat_dtype = torch.float64        # n_ao x n_ao matrices (complex128 at a k-point)
    # Hartree-Fock in the mixed mode: the exchange pair densities, their transforms
    # (complex64) and the contraction in FP32 while the DIIS residual is >= this, in
    # FP64 after.  Restricted HF of random hydrogen cells often oscillates near the
    # stopping rule; 24 geometries of the uneven HF batch converged 13 (FP64), 14
    # (switch at 1e-2, 14% less time) and 10 (switch at 1e-4), so the FP32 exchange
    # is kept to the early iterations
    MIXED_X32 = True
    MIXED_X_DIRECT_ERR = 1e-2

    def __init__(self, cell, dft=False, xc="lda,vwn", smearing=False, smearing_method="fermi",
                 smearing_sigma=0.0, search_method="bisect", search_cycle=100, search_tol=1e-7,
                 tol=1e-7, max_cycle=100, diis=True, diis_space=8, diis_start_cycle=1,
                 diis_damp=0.0, precision="fp64", ovlp_cut=0.0, tuning=None):
        self.c = cell
        # launch parameters from the autotuner (hqc_amd.tune), set on the host thread
        # at the start of every solve; None keeps the library defaults
        self.tuning = dict(tuning) if tuning else None
        self.ovlp_cut = float(ovlp_cut)
        self.dft = dft
        self.gga = False
        if dft:
            xname, cname = xc.split(",")
            if xname in GGA_NAMES or cname in GGA_NAMES:
                if not (xname in GGA_NAMES and cname in GGA_NAMES):
                    raise ValueError("GGA needs PBE exchange and PBE correlation ('pbe,pbe').")
                self.gga = True
                cname = ""
            elif xname != "lda":
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
        if precision == "fp32":
            precision = "mixed"
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        # kernel launches: FP64 / FP32 densities and potential matrices
        self.stats = {"rho64": 0, "rhoq8": 0, "rhoq8_top4": 0, "rhoq8_top3": 0, "rho32": 0,
                      "v64": 0, "v32": 0,
                      "x64": 0, "x32": 0}
        self.sweeps = torch.zeros((), dtype=torch.long, device=cell.device)
        self.eigh_calls = 0
        self.active = []
        self._work = {}

    # ------------------------------------------------------------------ utils
    def _workspace(self, B, prec, ld=None):
        # one grow-only buffer per precision; calls of any batch size share it (ld: the
        # operand's row length, the grid's by default)
        ld = self.c.ld if ld is None else ld
        nbytes = (K.vmat_h2_workspace(B, self.c.nao_pad, ld) if prec == 16 else
                  K.vmat_workspace(B, self.c.nao_pad, ld, prec))
        buf = self._work.get(prec)
        if buf is None or buf.numel() * 8 < nbytes:
            buf = torch.empty(max((nbytes + 7) // 8, 1), dtype=torch.float64,
                              device=self.c.device)
            self._work[prec] = buf
        return buf

    def _pad(self, M):
        c = self.c
        if c.nao_pad == c.nao:
            return M.contiguous()
        out = torch.zeros(M.shape[0], c.nao_pad, c.nao_pad, dtype=M.dtype, device=M.device)
        out[:, :c.nao, :c.nao] = M
        return out

    def vmat(self, phi, v, dV=None, bidx=None, vmax=None):
        """dV * phi diag(v) phi^T for a batch -> (B, nao, nao); dV a float (None: the
        cell's) or a (B,) tensor of per-geometry volume elements; bidx (B,) int32
        selects the geometries of phi (None: phi holds exactly the batch); vmax (B,)
        max_r |v| if known (used by the split-product kernel)."""
        if dV is None:
            return self.vmat_dv(phi, v, self.c.dV, bidx, vmax)
        if torch.is_tensor(dV):
            return self.vmat_dv(phi, v * dV[:, None], 1.0, bidx,
                                None if vmax is None else vmax * dV)
        return self.vmat_dv(phi, v, dV, bidx, vmax)

    @staticmethod
    def _dv(st, na):
        """Volume element(s) of geometries [0, na): the cell's float, or (na,) per geometry."""
        return st["dVb"][:na] if st.get("hetero") else st["dV"]

    @staticmethod
    def _geo(st, na):
        return {k: st[k][:na] for k in ("Lb", "dVb", "VGb")} if st.get("hetero") else None

    def vmat_dv(self, phi, v, dV, bidx=None, vmax=None):
        c = self.c
        B = v.shape[0]
        V = torch.empty(B, c.nao_pad, c.nao_pad, dtype=torch.float64, device=c.device)
        if phi.dtype == torch.int32:                  # FP16 pairs: split-precision products
            K.ao_vmat_h2(phi, v.contiguous(), self._phi_scale, dV, V, self._workspace(B, 16),
                         bidx, None if vmax is None else vmax.contiguous())
        else:
            prec = 64 if phi.dtype == torch.float64 else 32
            K.ao_vmat(phi, v.contiguous(), dV, V, self._workspace(B, prec, phi.shape[-1]), bidx)
        return V[:, :c.nao, :c.nao]

    def density(self, phi, D, X=None):
        c = self.c
        rho = torch.zeros(D.shape[0], c.ld, dtype=torch.float64, device=c.device)
        K.ao_density(phi, self._pad(D), rho, X)
        return rho

    OCC_CUT = 1e-15   # orbitals below this occupation contribute < 1e-15 relative to rho

    def _at(self, mo, occ, keep=None):
        """At = C sqrt(f) (B, nao_pad, m_pad) over the orbitals with f > OCC_CUT (keep:
        their largest count over the batch, if known)."""
        c = self.c
        B = mo.shape[0]
        if keep is None:
            keep = int((occ > self.OCC_CUT).sum(-1).max().item())
        keep = max(keep, 1)
        m_pad = _round_up(keep, 16)
        At = torch.zeros(B, c.nao_pad, m_pad, dtype=torch.float64, device=c.device)
        At[:, :c.nao, :keep] = mo[:, :, :keep] * occ[:, None, :keep].clamp(min=0).sqrt()
        return At

    def orb_density(self, phi, mo, occ, bidx=None, keep=None):
        """rho = sum_i f_i psi_i^2 over the orbitals with f_i > OCC_CUT (equal to the
        D form with D = C f C^T up to that cut); bidx as in vmat, keep as in _at."""
        c = self.c
        At = self._at(mo, occ, keep)
        rho = torch.zeros(mo.shape[0], c.ld, dtype=torch.float64, device=c.device)
        K.orb_density(phi, At, rho, bidx,
                      self._phi_scale if phi.dtype == torch.int32 else None)
        return rho

    def orb_density_q8(self, st, mo, occ, bidx=None, keep=None, nuse=None, top=5):
        """orb_density from the int8 digits of phi (st['phi_q8'], st['phi_e8']); nuse
        of the stored digits (all by default), digit products up to level top."""
        c = self.c
        At = self._at(mo, occ, keep)
        rho = torch.empty(mo.shape[0], c.ld, dtype=torch.float64, device=c.device)
        K.orb_density_q8(st["phi_q8"], st["phi_e8"], At, rho, self.Q8_SLICES, bidx, nuse, top)
        return rho

    def poisson(self, rho, save=False, geo=None, paired=False):
        """v_H on the grid (complex buffer, real part = v_H) and optionally
        rho~(G) = dV FFT(rho).  geo: per-geometry (Lb, dVb, VGb) of a batch of
        different cell sizes (None: the cell's constants).  paired (only without
        save and geo): geometries 2p, 2p+1 share one complex grid as Re and Im, so a
        batch takes half the transforms; v_H of geometry b is then Re or Im of grid b/2."""
        c = self.c
        B = rho.shape[0]
        if paired:
            assert not save and geo is None
            cbuf = torch.empty((B + 1) // 2, c.g, c.g, c.g, dtype=torch.complex128,
                               device=c.device)
            K.real_pair_to_complex(rho, c.g, c.dV, cbuf)
            K.fft3d(cbuf, -1)
            K.poisson_mul(cbuf, c.VG)
            K.fft3d(cbuf, +1)
            return cbuf, None
        cbuf = torch.empty(B, c.g, c.g, c.g, dtype=torch.complex128, device=c.device)
        if geo is None:
            K.real_to_complex(rho, c.g, c.dV, cbuf)
        else:
            K.real_to_complex((rho * geo["dVb"][:, None]).contiguous(), c.g, 1.0, cbuf)
        K.fft3d(cbuf, -1)
        saved = torch.empty_like(cbuf) if save else None
        if geo is None:
            K.poisson_mul(cbuf, c.VG, saved)
        else:
            K.poisson_mul(cbuf, geo["VGb"], saved)
        K.fft3d(cbuf, +1)
        return cbuf, saved

    def xc_potential(self, rho, save_rg=False, geo=None):
        """(v = v_H + v_xc (B, ld), per-geometry sums (B, 2) of the xc energy density and
        rho v_H, rho~(G) = dV FFT(rho) if save_rg or GGA).

        GGA (PBE): grad rho = IFFT(i G rho~) / Omega on the odd grid, then
        v_xc = de/drho - div(2 de/dsigma grad rho) with the same spectral
        derivative, the exact derivative of the discrete functional."""
        c = self.c
        B, g = rho.shape[0], c.g
        paired = not (self.gga or save_rg) and geo is None and B > 1
        vh, rg = self.poisson(rho, save=self.gga or save_rg, geo=geo, paired=paired)
        v = torch.zeros(B, c.ld, dtype=torch.float64, device=c.device)
        if not self.gga:
            part = torch.empty(B, K.xc_blocks(B, g), 2, dtype=torch.float64, device=c.device)
            K.xc_eval(rho, vh, g, self.xc_id, v, part, vh_paired=paired)
            return v, part.sum(1), rg
        grad = torch.empty(B, 3, g, g, g, dtype=torch.complex128, device=c.device)
        if geo is None:
            K.grad_G(rg, g, c.L, 1.0 / c.omega, grad)
        else:                                   # rho~ / Omega_b = FFT(rho) / N
            K.grad_G((rg / geo["dVb"][:, None, None, None]).contiguous(), g, geo["Lb"],
                     1.0 / c.N, grad)
        K.fft3d(grad.view(3 * B, g, g, g), +1)
        w = torch.empty_like(grad)
        part = torch.empty(B, K.gga_blocks(g), 2, dtype=torch.float64, device=c.device)
        K.xc_gga_eval(rho, grad, vh, g, v, w, part)
        del grad
        K.fft3d(w.view(3 * B, g, g, g), -1)
        K.div_G(w, g, c.L if geo is None else geo["Lb"], 1.0 / c.N, vh)
        K.fft3d(vh, +1)
        v[:, :c.N] -= vh.real.reshape(B, c.N)
        return v, part.sum(1), rg

    # ----------------------------------------------------------------- set-up
    def _cell_groups(self, B, Lb):
        """[(CellConsts, index tensor or None for the whole batch)] per distinct cell length."""
        c = self.c
        if Lb is None:
            return [(c.consts(c.L), None)]
        vals = Lb.cpu().numpy()
        groups = []
        for L in np.unique(vals):
            idx = torch.tensor(np.nonzero(vals == L)[0], device=c.device)
            groups.append((c.consts(float(L)), None if len(idx) == B else idx))
        return groups

    @staticmethod
    def _sub(t, idx):
        return t if idx is None else t[idx].contiguous()

    @staticmethod
    def _put(dst, idx, val):
        if idx is None:
            if val is not dst:
                dst.copy_(val)
        else:
            dst[idx] = val

    LINDEP_SHIFT = 1e4     # Hartree: energy given to overlap directions removed by ovlp_cut

    def _orthogonalizer(self, st, S):
        """v = U s^-1/2 (hqc/pbc/solver.py:729-730, no eigenvalue cut).  With
        ovlp_cut > 0 (an option beyond the reference, for nearly linearly dependent
        bases such as dense dzvp) overlap eigenvalues <= ovlp_cut get a zero column
        in v and the eigenvalue LINDEP_SHIFT in F' = v^H F v, so they become empty
        high-lying bands and every shape stays n_ao."""
        w, u = torch.linalg.eigh(S)
        if self.ovlp_cut > 0.0:
            keep = w > self.ovlp_cut
            st["v"] = u * torch.where(keep, w.clamp(min=self.ovlp_cut).rsqrt(), 0.0).unsqueeze(-2)
            st["lindep"] = (~keep).to(torch.float64) * self.LINDEP_SHIFT
            st["ovlp_w"], st["ovlp_u"] = w, u           # for the forces of the cut basis
        else:
            st["v"] = u * w.rsqrt().unsqueeze(-2)

    def setup(self, xpb, kpt=None, Lb=None, forces=False):
        """Per-geometry constants for xp (B, n, 3) in Bohr on the device.  Every
        tensor whose leading axis is the geometry is listed in st['per'] so the
        SCF can keep the active geometries as a contiguous prefix.

        Lb (B,) Bohr gives each geometry its own cell length (all on this cell's
        grid); the L-dependent set-up kernels then run once per distinct length.
        forces=True also keeps the S/T centre derivatives (st['dS'], st['dT']) from
        the same lattice sums, for the force pass."""
        c = self.c
        B = xpb.shape[0]
        dev, f64 = c.device, torch.float64
        hetero = Lb is not None
        if hetero:
            Lb = torch.as_tensor(Lb, dtype=f64).to(dev).reshape(B).contiguous()
        groups = self._cell_groups(B, Lb)
        st = {"xpb": xpb.contiguous(), "B": B, "hetero": hetero, "dV": c.dV}
        st["Lb"] = Lb if hetero else torch.full((B,), c.L, dtype=f64, device=dev)
        st["dVb"] = (st["Lb"] / c.g) ** 3
        # AO values stay in geometry order while the SCF compacts the slots (gidx maps
        # slot -> geometry) for DFT; the exchange path moves them with the slots
        inplace = self.dft
        per = ["xpb", "Lb", "dVb", "S", "T", "v", "vloc", "V", "Hcore"]
        phi_keys = ["phi"]
        st["gidx"] = torch.arange(B, dtype=torch.int32, device=dev)
        if inplace:
            per.append("gidx")
        if hetero:
            st["VGb"] = torch.empty(B, c.N, dtype=f64, device=dev)
            for cst, idx in groups:
                self._put(st["VGb"], idx, cst.VG.expand(B if idx is None else len(idx), c.N))
            per.append("VGb")
        S = torch.empty(B, c.nao, c.nao, dtype=f64, device=dev)
        T = torch.empty_like(S)
        dS = torch.empty(B, c.nao, c.nao, 3, dtype=f64, device=dev) if forces else None
        dT = torch.empty_like(dS) if forces else None
        def ao_buffer(nb, dtype):
            # AO values are written for the real AOs at the real grid points only
            buf = torch.empty(nb, c.nao_pad, c.ld, dtype=dtype, device=dev)
            buf[:, c.nao:, :] = 0
            buf[:, :c.nao, c.N:] = 0
            return buf
        phi = ao_buffer(B, f64)
        mixed = self.precision == "mixed"
        use_h2 = mixed and self.MIXED_SPLIT16 and c.nao_pad <= 128
        # FP32 AO values: the FP32 kernels' operand, unless FP16 pairs serve them all
        need32 = mixed and (not use_h2 or self.MIXED_RHO_INCREMENTS)
        phi32 = ao_buffer(B, torch.float32) if need32 else None
        # int8 digits and FP16 pairs cut in one read of phi: the AO-value pass gives
        # max |phi| per geometry first (the pairs' scale)
        q8 = mixed and self.MIXED_Q8 and c.nao_pad <= K.Q8_MAX_NAO_PAD
        amax_g = torch.empty(B, dtype=f64, device=dev) if q8 and use_h2 else None
        vl = torch.zeros(B, c.ld, dtype=f64, device=dev)
        for cst, idx in groups:
            x = self._sub(st["xpb"], idx)
            nb = x.shape[0]
            Sg = S if idx is None else torch.empty(nb, c.nao, c.nao, dtype=f64, device=dev)
            Tg = T if idx is None else torch.empty_like(Sg)
            dSg = dTg = None
            if forces:
                dSg = dS if idx is None else torch.empty(nb, c.nao, c.nao, 3, dtype=f64,
                                                         device=dev)
                dTg = dT if idx is None else torch.empty_like(dSg)
            K.overlap_kinetic(x, cst.lattice, c.alpha, c.coef, c.comp, Sg, Tg, dSg, dTg,
                              s_only=c.basis.l_max == 0)
            self._put(S, idx, Sg)
            self._put(T, idx, Tg)
            if forces:
                self._put(dS, idx, dSg)
                self._put(dT, idx, dTg)
            P = torch.empty(nb, c.n, 3, c.g, c.na, c.nl, dtype=f64, device=dev)
            K.gto_tables(x, c.g, cst.L, cst.tmax_img, c.alpha, c.nl, P)
            pg = phi if idx is None else ao_buffer(nb, f64)
            ag = None
            if amax_g is not None:
                ag = amax_g if idx is None else torch.empty(nb, dtype=f64, device=dev)
            K.ao_values(P, c.coef, c.comp, c.nao_pad, c.ld, pg, amax=ag)
            self._put(phi, idx, pg)
            if ag is not None:
                self._put(amax_g, idx, ag)
            if need32:
                pg32 = phi32 if idx is None else ao_buffer(nb, torch.float32)
                K.ao_values(P, c.coef, c.comp, c.nao_pad, c.ld, pg32)
                self._put(phi32, idx, pg32)
            cb = torch.empty(nb, c.g, c.g, c.g, dtype=torch.complex128, device=dev)
            K.vloc_G(x, c.g, cst.L, cst.VG, cb)
            K.fft3d(cb, +1)
            vg = vl if idx is None else torch.zeros(nb, c.ld, dtype=f64, device=dev)
            K.complex_real(cb, c.ld, vg)
            self._put(vl, idx, vg)
        st["S"], st["T"] = S, T
        if forces:
            st["dS"], st["dT"] = dS, dT
            per += ["dS", "dT"]
        self._orthogonalizer(st, S)
        if "lindep" in st:
            per += ["lindep", "ovlp_w", "ovlp_u"]
        st["phi"] = phi
        dVs = self._dv(st, B)
        if mixed:
            if need32:
                st["phi_lp"] = phi32
                phi_keys.append("phi_lp")
            if use_h2:
                # FP16 (hi, lo) pairs of s*phi, s = 2^k with |s phi| <= 2^14
                am = amax_g.max() if amax_g is not None else \
                    torch.linalg.vector_norm(phi, float("inf"))
                amax = float(am.item())
                self._phi_scale = 2.0 ** math.floor(math.log2(2.0 ** 14 / max(amax, 1e-300)))
                st["phi_h2"] = torch.empty(B, c.nao_pad, c.ld, dtype=torch.int32, device=dev)
                phi_keys.append("phi_h2")
            if q8:
                # int8 digits of phi for the accurate density (exact integer products)
                ns = self.Q8_SLICES
                st["phi_q8"] = torch.empty(B, c.ld * ns * c.nao_pad, dtype=torch.int8, device=dev)
                st["phi_e8"] = torch.empty(B, c.ld, dtype=torch.int16, device=dev)
                if use_h2:
                    K.phi_slice_i8_h2(phi, ns, self._phi_scale, st["phi_q8"], st["phi_e8"],
                                      st["phi_h2"])
                else:
                    K.phi_slice_i8(phi, ns, st["phi_q8"], st["phi_e8"])
                phi_keys += ["phi_q8", "phi_e8"]
            elif use_h2:
                K.phi_split_f16(phi, self._phi_scale, st["phi_h2"])
            if self.MIXED_RHO_INCREMENTS:
                st["phisq"] = (phi[:, :c.nao, :c.N] ** 2).sum(1)         # sum_p phi_p(r)^2
                st["Da"] = torch.zeros(B, c.nao, c.nao, dtype=f64, device=dev)
                st["rhoa"] = torch.zeros(B, c.ld, dtype=f64, device=dev)
                st["kappa"] = torch.zeros(B, dtype=f64, device=dev)      # min rho_a/phisq
                per += ["phisq", "Da", "rhoa", "kappa"]
            st["va"] = torch.zeros(B, c.ld, dtype=f64, device=dev)
            st["Va"] = torch.zeros(B, c.nao, c.nao, dtype=f64, device=dev)
            st["vmax"] = torch.zeros(B, dtype=f64, device=dev)
            st["va_set"] = torch.zeros(B, dtype=torch.bool, device=dev)
            st["err"] = torch.full((B,), float("inf"), dtype=f64, device=dev)
            per += ["va", "Va", "vmax", "va_set", "err"]
        st["vloc"] = vl
        st["V"] = self.vmat(phi, vl, dVs)
        st["Hcore"] = T + st["V"]
        if not self.dft:
            ones = torch.zeros(B, c.ld, dtype=f64, device=dev)
            ones[:, :c.N] = 1.0
            st["Sgrid"] = self.vmat(phi, ones, dVs)      # dV phi phi^T (exchange G=0 term)
            st["VGh"] = torch.empty(B, c.ld, dtype=f64, device=dev)
            st["c0b"] = torch.empty(B, dtype=f64, device=dev)
            fused = c.g <= K.PAIR_FFT_MAX_G
            if fused:
                ldz = groups[0][0].ldz
                st["VGz"] = torch.empty(B, ldz, dtype=f64, device=dev)
            for cst, idx in groups:
                nb = B if idx is None else len(idx)
                self._put(st["VGh"], idx, cst.VGh.expand(nb, c.ld))
                if fused:
                    self._put(st["VGz"], idx, cst.VGz.expand(nb, ldz))
                self._put(st["c0b"], idx, torch.full((nb,), cst.c0, dtype=f64, device=dev))
            per += ["Sgrid", "VGh", "c0b"] + (["VGz"] if fused else [])
        if not inplace:
            per += phi_keys
        st["per"] = per
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
        return (mo * occ.unsqueeze(-2)) @ mo.transpose(-1, -2).contiguous(), occ

    @staticmethod
    def entropy(occ, eps=1e-15):
        """hqc/pbc/solver.py:579-590."""
        h = occ / 2
        a = torch.where(h > eps, h * torch.log(torch.where(h > 0, h, torch.ones_like(h))), 0.0)
        b = torch.where(1 - h > eps, (1 - h) * torch.log(torch.where(1 - h > 0, 1 - h,
                                                                         torch.ones_like(h))), 0.0)
        return -2 * (a.sum(-1) + b.sum(-1))

    # ----------------------------------------------------------- two-electron
    EXCHANGE_CHUNK_BYTES = 4 << 30      # pair grids + row blocks per launch group

    def exchange(self, st, na, mo, occ, low=None):
        """K (B, nao, nao) for geometries [0, na): equals sum_rs D_rs (eris+eris0)[p,q,s,r]
        (hqc/pbc/solver.py:430-431, 604-614).

        With A_i = sqrt(f_i) dV FFT(phi * psi_i), K = sum_i sum_G VG Re[A_i A_i^H]
        over G != 0, plus the G=0 term c0 S~ D S~ with the literal 0.22578495.
        Two occupied orbitals a, b share one complex FFT Z = A_a + i A_b.  On grids whose
        y-z plane fits the LDS (the solver's) the pair densities are formed on the
        transform's first pass (hqc_pair_pack_fft) into real rows [Re Z | Im Z], and
        K = sum over pairs of Z diag(VG) Z^T on every G: the terms odd in G cancel over
        the symmetric G-set, leaving VG (Re[A_a A_a^H] + Re[A_b A_b^H]).  Elsewhere
        pack, transform and the split kernel give the [Re | Im] blocks of A_a and A_b on
        half the G-set, contracted with 2 VG.  Either contraction is the AO x grid MFMA
        kernel, batched over all pairs of all geometries.  low (host list of bools
        over [0, na), or None): those
        geometries take the pair densities, their transforms and the contraction in
        FP32 (complex64 grids, FP32 MFMA), the others in FP64."""
        c = self.c
        dev = c.device
        keep = (occ > 1e-14).cpu()
        rows, nmax = [], 0
        for b in range(na):
            idx = torch.nonzero(keep[b]).flatten().tolist()
            if idx:
                nmax = max(nmax, idx[-1] + 1)
            for k in range(0, len(idx), 2):
                rows.append((b, idx[k], idx[k + 1] if k + 1 < len(idx) else -1))
        Kmat = torch.zeros(na, c.nao, c.nao, dtype=torch.float64, device=dev)
        if rows:
            dVs = self._dv(st, na)
            dVs = dVs[:, None] if torch.is_tensor(dVs) else dVs
            wgt = occ[:, :nmax].clamp_min(0).sqrt() * dVs * keep[:na, :nmax].to(dev)
            psi = (mo[:, :, :nmax] * wgt.unsqueeze(-2)).transpose(-1, -2).contiguous() \
                @ st["phi"][:na, :c.nao, :]                                  # (na, nmax, ld)
            VGh = st["VGh"]
            phi = st["phi"][:na]
            for f32 in (False, True):
                sel = [r for r in rows if bool(low[r[0]]) == f32] if low is not None else \
                    (rows if not f32 else [])
                if not sel:
                    continue
                table = torch.tensor(sel, dtype=torch.int32)
                zb, mb = (8, 4) if f32 else (16, 8)
                cplx = torch.complex64 if f32 else torch.complex128
                if "VGz" in st:
                    # pair densities formed on the transform's first pass, written as
                    # real rows [Re Z | Im Z] of the padded operand layout: K = sum over
                    # pairs of Z diag(VG) Z^T on every G (no split; see hqc_pair_pack_fft)
                    ldz = st["VGz"].shape[1]
                    per_pair = c.nao_pad * ldz * zb // 2 + c.nao_pad * (c.nao_pad + 1) * mb
                    chunk = max(1, min(len(sel), int(self.EXCHANGE_CHUNK_BYTES // per_pair),
                                       K.PAIR_FFT_MAX_GRIDS // c.nao))
                    # the transform writes every grid whole; only the padding rows and
                    # the columns past 2N are zeroed, once
                    zbuf = torch.empty(chunk, c.nao_pad, ldz // 2, device=dev, dtype=cplx)
                    zbuf[:, c.nao:].zero_()
                    zbuf[:, :c.nao, c.N:].zero_()
                    for s0 in range(0, len(sel), chunk):
                        t = table[s0:s0 + chunk]
                        mp = t.shape[0]
                        z = zbuf[:mp]
                        K.pair_pack(phi, psi, t, c.nao, c.g, z, fft_sign=-1)
                        zr = torch.view_as_real(z).view(mp, c.nao_pad, ldz)
                        part = self.vmat_dv(zr, st["VGz"][t[:, 0].long().to(dev)], 1.0)
                        geo, cnt = torch.unique_consecutive(t[:, 0], return_counts=True)
                        sums = torch.stack([q.sum(0) for q in torch.split(part, cnt.tolist())])
                        Kmat[geo.long().to(dev)] += sums
                    del zbuf
                else:
                    per_pair = c.nao * c.N * zb + 2 * c.nao_pad * (c.ld + c.nao_pad + 1) * mb
                    chunk = max(1, min(len(sel), int(self.EXCHANGE_CHUNK_BYTES // per_pair)))
                    for s0 in range(0, len(sel), chunk):
                        t = table[s0:s0 + chunk]
                        mp = t.shape[0]
                        z = torch.empty(mp, c.nao, c.N, device=dev, dtype=cplx)
                        K.pair_pack(phi, psi, t, c.nao, c.g, z)
                        K.fft3d(z.view(mp * c.nao, c.g, c.g, c.g), -1)
                        M = torch.empty(2 * mp, c.nao_pad, c.ld, device=dev,
                                        dtype=torch.float32 if f32 else torch.float64)
                        K.pair_split(z, c.nao_pad, c.ld, M)
                        del z
                        gid = t[:, 0].long().repeat_interleave(2).to(dev)
                        part = self.vmat_dv(M, VGh[gid], 1.0)
                        # rows of one geometry are contiguous: a fixed-order sum per
                        # geometry (index_add_ would accumulate with atomics, in no
                        # fixed order)
                        geo, cnt = torch.unique_consecutive(t[:, 0], return_counts=True)
                        sums = torch.stack([q.sum(0) for q in
                                            torch.split(part, (2 * cnt).tolist())])
                        Kmat[geo.long().to(dev)] += sums
                self.stats["x32" if f32 else "x64"] += 1
        D = (mo * occ.unsqueeze(-2)) @ mo.transpose(-1, -2).contiguous()
        Sg = st["Sgrid"][:na]
        return Kmat + st["c0b"][:na, None, None] * (Sg @ D @ Sg)

    @staticmethod
    def _gix(st, na, idx=None):
        """int32 geometry index (into the AO-value arrays) of the slots [0, na), or of
        the subset idx of them."""
        g = st["gidx"][:na]
        return g if idx is None else g[idx.long()]

    @staticmethod
    def _split(mask, extra=()):
        """((index, count) of the True entries, (index, count) of the False entries,
        host values of the integer tensors in extra) from one host round trip; an
        index is int32 on the device, or None when it covers the whole mask."""
        n = mask.numel()
        h = torch.cat([mask.reshape(-1).to(torch.int64)] +
                      [t.reshape(-1).to(torch.int64) for t in extra]).tolist()
        yes = [i for i in range(n) if h[i]]
        no = [i for i in range(n) if not h[i]]

        def up(lst):
            if len(lst) == n:
                return None
            return torch.tensor(lst, dtype=torch.int32).pin_memory().to(mask.device,
                                                                         non_blocking=True)
        return (up(yes), len(yes)), (up(no), len(no)), h[n:]

    @staticmethod
    def _classes(cls, ncls, extra=()):
        """([(index, count) of the entries of class c] for c < ncls, host values of the
        integer tensors in extra) from one host round trip; an index is int32 on the
        device, or None when it covers all of cls."""
        n = cls.numel()
        h = torch.cat([cls.reshape(-1).to(torch.int64)] +
                      [t.reshape(-1).to(torch.int64) for t in extra]).tolist()
        out = []
        for k in range(ncls):
            lst = [i for i in range(n) if h[i] == k]
            out.append((None if len(lst) == n else
                        torch.tensor(lst, dtype=torch.int32).pin_memory().to(
                            cls.device, non_blocking=True), len(lst)))
        return out, h[n:]

    @staticmethod
    def _subset(mask):
        """(index tensor int32 or None when the mask covers the batch, count); one
        host round trip (the size of nonzero's result)."""
        idx = torch.nonzero(mask).flatten()
        n = idx.numel()
        if n == mask.numel():
            return None, n
        return idx.to(torch.int32), n

    def _mixed_density(self, st, na, D, mo, occ):
        """rho for geometries [0, na) in the mixed mode, chosen per geometry: the
        low-precision orbital form outright while its DIIS residual is large
        (err >= MIXED_RHO_DIRECT_ERR), FP32 increments rho_a + rho(D - D_a) where the
        increment bound holds (off by default), the accurate orbital form (int8 digit
        products, or FP64; a new anchor) elsewhere.  Each class is one launch over its
        geometries, read in place from the AO-value arrays through a geometry index list."""
        c = self.c
        sel = lambda key: st[key][:na]
        err = sel("err")
        direct = err >= self.MIXED_RHO_DIRECT_ERR
        if "rhoa" in st:
            dD = D - sel("Da")
            # bound on the relative density error of the FP32 increment
            rel = self.MIXED_EPS * torch.linalg.norm(dD, dim=(-2, -1)) / sel("kappa")
            inc = ~direct & (rel <= 1e-2) & (rel * sel("vmax") <= self.MIXED_RHO_SAFETY * err)
        else:
            inc = None
        rho = torch.empty(na, c.ld, dtype=torch.float64, device=c.device)
        if inc is None:
            # the classes (low precision; accurate with digit products to level 4, to
            # level 5) and the occupied-orbital count in one round trip
            kc = (occ > self.OCC_CUT).sum(-1).amax()
            coarse = err >= self.MIXED_Q8_COARSE_ERR if "phi_q8" in st else \
                torch.zeros_like(direct)
            cls = torch.where(direct, 0, torch.where(coarse, 1, 2))
            ((idx, n), sub_c, sub_f), (keep,) = self._classes(cls, 3, (kc,))
            accurate = ((sub_c, self.Q8_TOP_COARSE), (sub_f, self.Q8_TOP))
        else:
            full = ~(direct | inc)
            idx, n = self._subset(direct)
            keep = None
            accurate = ((self._subset(full), 5),)
        if n:
            # orbital form sum_i f_i psi_i^2: non-negative in any precision, unlike the
            # D form, whose FP32 rounding can turn a tiny density negative
            mo_d = mo if idx is None else mo[idx.long()]
            occ_d = occ if idx is None else occ[idx.long()]
            if "phi_q8" in st and self.Q8_WARMUP_DIGITS:
                r = self.orb_density_q8(st, mo_d, occ_d, self._gix(st, na, idx), keep,
                                        self.Q8_WARMUP_DIGITS, top=self.Q8_WARMUP_TOP)
            else:
                lp = "phi_h2" if "phi_h2" in st else "phi_lp"
                r = self.orb_density(st[lp], mo_d, occ_d, self._gix(st, na, idx), keep)
            self._scatter(rho, idx, r)
            self.stats["rho32"] += 1
        idx, n = self._subset(inc) if inc is not None else (None, 0)
        if n:
            il = None if idx is None else idx.long()
            r = torch.zeros(n, c.ld, dtype=torch.float64, device=c.device)
            K.ao_density(st["phi_lp"], self._pad(dD if il is None else dD[il]), r,
                         bidx=self._gix(st, na, idx))
            r += sel("rhoa") if il is None else st["rhoa"][il]
            self._scatter(rho, idx, r)
            self.stats["rho32"] += 1
        for (idx, n), top in accurate:
            if not n:
                continue
            il = torch.arange(na, device=c.device) if idx is None else idx.long()
            gi = self._gix(st, na, idx)
            if "phi_q8" in st:
                r64 = self.orb_density_q8(st, mo[il], occ[il], gi, keep, top=top)
                self.stats[{3: "rhoq8_top3", 4: "rhoq8_top4"}.get(top, "rhoq8")] += 1
            else:
                r64 = self.orb_density(st["phi"], mo[il], occ[il], gi, keep)
                self.stats["rho64"] += 1
            self._scatter(rho, idx, r64)
            if "rhoa" in st:                                 # anchor of the increments
                st["Da"][il] = D[il]
                st["rhoa"][il] = r64
                st["kappa"][il] = (r64[:, :c.N] / st["phisq"][il]).amin(-1)
        return rho

    @staticmethod
    def _scatter(dst, idx, val):
        if idx is None:
            dst.copy_(val)
        else:
            dst[idx.long()] = val

    def _mixed_vmat(self, st, na, v):
        """V(v) for geometries [0, na) in the mixed mode, per geometry like the density:
        low-precision V(v) outright in the large-residual phase, V_a + V(v - v_a) in low
        precision where the increment bound holds (both in one launch of the FP16
        split-product kernel, or FP32), FP64 (and a new anchor) elsewhere."""
        c = self.c
        sel = lambda key: st[key][:na]
        err = sel("err")
        direct = err >= self.MIXED_DIRECT_ERR
        # max |v|, max |v - v_a| and the operand v (direct) or v - v_a, in one pass
        v = v.contiguous()
        vv = torch.empty_like(v)
        dvmax = torch.empty(na, dtype=torch.float64, device=c.device)
        K.mixed_v_prep(v, sel("va"), direct.contiguous(), vv, st["vmax"][:na], dvmax)
        tol_e = self.MIXED_SAFETY * err / c.nao                 # per-element error budget
        lp = "phi_h2" if "phi_h2" in st else "phi_lp"
        eps = self.MIXED_EPS_SPLIT16 if lp == "phi_h2" else self.MIXED_EPS
        inc = ~direct & sel("va_set") & (eps * dvmax <= tol_e)
        low = direct | inc
        dVs = self._dv(st, na)
        V2 = torch.empty(na, c.nao, c.nao, dtype=self.mat_dtype, device=c.device)
        (idx, n), high_sub, _ = self._split(low)
        if n:
            il = None if idx is None else idx.long()
            vvm = torch.where(direct, st["vmax"][:na], dvmax)
            d = dVs if not torch.is_tensor(dVs) or il is None else dVs[il]
            V32 = self.vmat(st[lp], vv if il is None else vv[il], d, self._gix(st, na, idx),
                            vvm if il is None else vvm[il])
            V32 = V32 + torch.where(inc[:, None, None], sel("Va"), 0.0)[
                slice(None) if il is None else il]
            self._scatter(V2, idx, V32)
            self.stats["v32"] += 1
        idx, n = high_sub
        if n:
            il = torch.arange(na, device=c.device) if idx is None else idx.long()
            d = dVs if not torch.is_tensor(dVs) else dVs[il]
            V64 = self.vmat(st["phi"], v[il], d, self._gix(st, na, idx))
            self._scatter(V2, idx, V64)
            st["va"][il] = v[il]
            st["Va"][il] = V64
            st["va_set"][il] = True
            self.stats["v64"] += 1
        return V2

    def _final_grid_pass(self, st, B, D):
        """E_H + E_xc of the final density for a DFT solve followed by forces: one FP64
        D-form pass gives rho and the force adjoint's X = D phi together, and the
        potential, rho~(G) and X are kept in st['force_grid'] for the force pass (the
        same density as the orbital form, summed in another order)."""
        c = self.c
        X = torch.empty(B, c.nao_pad, c.ld, dtype=torch.float64, device=c.device)
        rho = torch.empty(B, c.ld, dtype=torch.float64, device=c.device)
        K.ao_density(st["phi"], self._pad(D), rho, X, bidx=st["gidx"])
        self.stats["rho64"] += 1
        v, E, rg = self.xc_potential(rho, save_rg=True, geo=self._geo(st, B))
        st["force_grid"] = (X, v, rg)
        dVs = self._dv(st, B)
        Exc, EH = E[:, 0] * dVs, 0.5 * E[:, 1] * dVs
        return EH + Exc

    def fock2e(self, st, na, D, mo, occ, full=False, need_v=True):
        """(J + V_xc or J - K/2, E_H + E_xc or E_H + E_x) for geometries [0, na).
        need_v=False (FP64 DFT only): the energies alone, V is None."""
        c = self.c
        mixed = self.precision == "mixed" and not full
        dVs = self._dv(st, na)
        geo = self._geo(st, na)
        if not mixed:
            gi = self._gix(st, na)
            rho = self.orb_density(st["phi"], mo, occ, gi)
            self.stats["rho64"] += 1
            v, E, _ = self.xc_potential(rho, geo=geo)
            V2 = None
            if need_v or not self.dft:
                V2 = self.vmat(st["phi"], v, dVs, gi)
                self.stats["v64"] += 1
        else:
            rho = self._mixed_density(st, na, D, mo, occ)
            v, E, _ = self.xc_potential(rho, geo=geo)
            V2 = self._mixed_vmat(st, na, v)
        Exc, EH = E[:, 0] * dVs, 0.5 * E[:, 1] * dVs
        if self.dft:
            return V2, EH + Exc
        low = None
        if mixed and self.MIXED_X32:
            low = (st["err"][:na] >= self.MIXED_X_DIRECT_ERR).tolist()
        Vx = -0.5 * self.exchange(st, na, mo, occ, low)
        Ex = 0.5 * _tr(Vx, D).real
        return V2 + Vx, EH + Ex

    # --------------------------------------------------------------------- SCF
    EIG_TOL_MIN, EIG_TOL_MAX = 1e-14, 1e-8   # Jacobi stopping range (relative off-diagonal)
    EIG_TOL_PER_RESIDUAL = 1e-4              # tol = this x the smallest DIIS residual

    def _eig_tol(self, res):
        """Jacobi tolerance for the next Roothaan step: the eigenvector error it leaves
        (~ tol |F| / gap) stays far below the current DIIS residual; 1e-14 when the
        residual is unknown."""
        if res is None:
            return self.EIG_TOL_MIN
        return min(self.EIG_TOL_MAX, max(self.EIG_TOL_MIN, self.EIG_TOL_PER_RESIDUAL * res))

    def _eigh_orth(self, F, v, Cw=None, shift=None, tol=None):
        """Roothaan step (hqc/pbc/scf.py:68-72): F' = v^T F v, eigh, mo = v C'.

        n_ao <= 110 uses the LDS Jacobi kernel, warm-started in the previous
        eigenbasis Cw when given (A = Cw^T F' Cw is then nearly diagonal and a
        few sweeps suffice); larger problems use torch.linalg.eigh.  Returns
        (w, mo, C') with C' orthonormal eigenvectors of F'."""
        # transposed operands are made contiguous: rocBLAS picks a far slower
        # kernel for batched 108x108 products with a transposed A
        Fp = v.transpose(-1, -2).contiguous() @ (F @ v)
        if shift is not None:
            Fp = Fp + torch.diag_embed(shift)
        if self.c.nao <= K.EIGH_JACOBI_MAXN:
            sw = torch.zeros(F.shape[0], dtype=torch.int32, device=F.device)
            if Cw is not None:
                w, U = K.eigh_jacobi(Cw.transpose(-1, -2).contiguous() @ (Fp @ Cw), sweeps=sw,
                                     tol=tol or self.EIG_TOL_MIN)
                C = Cw @ U
            else:
                w, C = K.eigh_jacobi(Fp, sweeps=sw, tol=tol or self.EIG_TOL_MIN)
            self.sweeps += sw.sum()
            self.eigh_calls += F.shape[0]
        else:
            w, C = torch.linalg.eigh(Fp)
        return w, v @ C, C

    @staticmethod
    def _errvec(st, na, S, D, F):
        """DIIS error SDF - FDS (hqc/pbc/solver.py:696-700).  With an overlap cut it is
        taken in the orthogonalised basis, v^H (SDF - FDS) v, whose zero is the
        stationarity condition of the truncated problem."""
        e = S @ D @ F - F @ D @ S
        if "lindep" in st:
            v = st["v"][:na]
            e = v.mH @ e @ v
        return e

    @staticmethod
    def _compact(st, keys, loc, order, keep, na):
        """Move the geometries flagged in keep (bool over [0, na), a device tensor or
        a host list) to a prefix by swapping them with the inactive ones inside it;
        returns the new prefix."""
        if torch.is_tensor(keep):
            keep = keep.tolist()
        new_na = sum(keep)
        if new_na == na:
            return na
        holes = [i for i in range(new_na) if not keep[i]]
        movers = [i for i in range(new_na, na) if keep[i]]
        if holes:
            dev = order.device
            pairs = torch.tensor([holes, movers], dtype=torch.long).pin_memory().to(
                dev, non_blocking=True)
            ts = [st[k] for k in keys] + list(loc.values()) + [order]
            ts = [t for t in ts if t.numel()]
            flat = [t for t in ts if t.is_contiguous()]
            assert len({t.data_ptr() for t in flat}) == len(flat), "aliased per-geometry state"
            K.swap_rows(flat, pairs)                 # one launch for every contiguous tensor
            rest = [t for t in ts if not t.is_contiguous()]
            if rest:
                ia, ib = pairs.reshape(-1), pairs.flip(0).reshape(-1)
                for t in rest:
                    t[ia] = t[ib]
        return new_na

    def scf(self, st):
        """hqc/pbc/scf.py on the batch.  Returns tensors in SLOT order (the SCF
        permutes geometries so the active ones form a prefix) plus `order`:
        slot i holds geometry order[i].  mo (B,nao,nmo), w1 (B,nmo), E (B)
        Hartree, converged (B), n_iter (B)."""
        c = self.c
        B = st["B"]
        dev = c.device
        lindep = st.get("lindep")
        sh = (lambda na: None) if lindep is None else (lambda na: st["lindep"][:na])
        w1, mo, Cw = self._eigh_orth(st["Hcore"], st["v"], shift=sh(B))  # solver.py:740-745
        D, occ = self.density_matrix(mo, w1)
        tol = self.tol
        order = torch.arange(B, device=dev)
        keys = st["per"]

        if not self.diis:                                                 # scf.py:21-88
            loc = {"E_old": torch.zeros(B, dtype=torch.float64, device=dev),
                   "E": torch.ones(B, dtype=torch.float64, device=dev),
                   "Dc": D.clone(), "Oc": occ.clone(), "Mo": mo.clone(), "W1": w1.clone(),
                   "Cw": Cw.clone(), "nit": torch.zeros(B, dtype=torch.long, device=dev)}
            na, loop = B, 0
            while na > 0:
                H = st["Hcore"][:na]
                Dc = loc["Dc"][:na]
                V2, E2 = self.fock2e(st, na, Dc, loc["Mo"][:na], loc["Oc"][:na], full=True)
                Enew = _tr(H, Dc).real + E2
                w1n, mon, cwn = self._eigh_orth(H + V2, st["v"][:na], loc["Cw"][:na], sh(na))
                loc["Cw"][:na] = cwn
                Dn, On = self.density_matrix(mon, w1n)
                loc["E_old"][:na] = loc["E"][:na]
                loc["E"][:na] = Enew
                loc["Dc"][:na], loc["Oc"][:na] = Dn, On
                loc["Mo"][:na], loc["W1"][:na] = mon, w1n
                loop += 1
                loc["nit"][:na] = loop
                keep = ((loc["E"][:na] - loc["E_old"][:na]).abs() > tol) & (loop < self.max_cycle)
                na = self._compact(st, keys, loc, order, keep, na)
            return loc["Mo"], loc["W1"], loc["E"], loc["nit"] != self.max_cycle, loc["nit"], order

        def step(Fin, na, tol):
            w1, mo, cw = self._eigh_orth(Fin, st["v"][:na], loc["Cw"][:na], sh(na), tol)
            loc["Cw"][:na] = cw
            D, occ = self.density_matrix(mo, w1)
            V2, E2 = self.fock2e(st, na, D, mo, occ)
            H, S = st["Hcore"][:na], st["S"][:na]
            Enew = _tr(H, D).real + E2
            F = H + V2
            e = self._errvec(st, na, S, D, F)
            if self.precision == "mixed":
                st["err"][:na] = torch.linalg.norm(e, dim=(-2, -1))
            return mo, w1, Enew, F, e

        # the DIIS window is a ring: the newest (F, e) overwrites the oldest slot, so
        # nothing is shifted; the extrapolation is invariant under a reordering of the
        # window (the bordered Pulay system permutes with it)
        last = [self.diis_space - 1]

        def push(loc, na, F, e):
            """Newest (F, e) into the ring and row/column of B_ij = <e_i, e_j>."""
            m = self.diis_space
            s = (last[0] + 1) % m
            loc["Fk"][:na, s] = F
            loc["ek"][:na, s] = e
            # elementwise product + sum: a batched GEMV with K = n_ao^2 gets a slow
            # library kernel here
            d = (loc["ek"][:na] * e.unsqueeze(1)).sum((-1, -2))
            loc["Bh"][:na, s, :] = d
            loc["Bh"][:na, :, s] = d
            last[0] = s

        m = self.diis_space
        V2, E2 = self.fock2e(st, B, D, mo, occ)                           # scf.py:122-128
        H, S = st["Hcore"], st["S"]
        F0 = H + V2
        e0 = self._errvec(st, B, S, D, F0)
        loc = {"Fk": F0.unsqueeze(1).repeat(1, m, 1, 1).contiguous(),
               "ek": e0.unsqueeze(1).repeat(1, m, 1, 1).contiguous(),
               "Bh": (e0 * e0).sum((-1, -2))[:, None, None].repeat(1, m, m).contiguous(),
               "E": torch.ones(B, dtype=torch.float64, device=dev),
               "E_old": torch.zeros(B, dtype=torch.float64, device=dev),
               "Mo": mo.clone(), "W1": w1.clone(), "Cw": Cw.clone(),
               "nit": torch.zeros(B, dtype=torch.long, device=dev)}
        na = B
        for _ in range(self.diis_start_cycle + m):                        # scf.py:130-172
            mo_, w1_, Enew, F, e = step(loc["Fk"][:, last[0]], na,
                                        self.EIG_TOL_MIN if lindep is not None else
                                        self.EIG_TOL_MAX)
            push(loc, na, F, e)
            loc["Mo"][:], loc["W1"][:] = mo_, w1_
            loc["E_old"][:] = loc["E"]
            loc["E"][:] = Enew
        loop = self.diis_start_cycle + m
        loc["E_old"][:] = loc["E"] - 1.0
        loc["nit"][:] = loop

        res = [None]

        def cont(na):
            """Host list of the geometries to keep; one round trip for the residuals
            and energy changes together."""
            err = torch.linalg.norm(loc["ek"][:na, last[0]], dim=(-2, -1))
            dE = (loc["E"][:na] - loc["E_old"][:na]).abs()
            e_h, d_h = torch.stack([err, dE]).tolist() if na else ([], [])
            res[0] = min(e_h) if na else None
            more = loop < self.max_cycle
            return [more and (d > tol or e > 10 * tol) for e, d in zip(e_h, d_h)]

        na = self._compact(st, keys, loc, order, cont(na), na)
        self.active = []          # geometries in each DIIS iteration (host ints, for reports)
        # complex (k-point) error vectors give a complex symmetric B, as 'imn,jmn->ij'
        # does in the reference (no conjugate)
        eye = torch.eye(m, dtype=loc["Bh"].dtype, device=dev)
        while na > 0:                                                     # scf.py:174-239
            Bm = loc["Bh"][:na]
            eta = 1e-12 * (Bm.diagonal(dim1=-2, dim2=-1).sum(-1) / m + 1.0)
            h = torch.zeros(na, m + 1, m + 1, dtype=Bm.dtype, device=dev)
            h[:, :m, :m] = Bm + eta[:, None, None] * eye
            h[:, :m, m] = -1.0
            h[:, m, :m] = -1.0
            rhs = torch.zeros(na, m + 1, 1, dtype=Bm.dtype, device=dev)
            rhs[:, m, 0] = -1.0
            ck = torch.linalg.solve_ex(h, rhs)[0][:, :m, 0]      # no host check of info
            Fk = loc["Fk"][:na]
            Fg = (ck[:, :, None, None] * Fk).sum(1)
            if self.diis_damp:
                Fg = (1 - self.diis_damp) * Fg + self.diis_damp * Fk[:, last[0]]
            # the overlap-cut shift dominates |F'|, so a relative tolerance would be loose
            etol = self.EIG_TOL_MIN if lindep is not None else self._eig_tol(res[0])
            mo_a, w1_a, Enew, F, e_new = step(Fg, na, etol)
            push(loc, na, F, e_new)
            loc["E_old"][:na] = loc["E"][:na]
            loc["E"][:na] = Enew
            loc["Mo"][:na], loc["W1"][:na] = mo_a, w1_a
            loop += 1
            loc["nit"][:na] = loop
            self.active.append(na)
            na = self._compact(st, keys, loc, order, cont(na), na)
        return loc["Mo"], loc["W1"], loc["E"], loc["nit"] != self.max_cycle, loc["nit"], order

    # ---------------------------------------------------------------- results
    def solve(self, xpb, kpt=None, Lb=None, forces=False):
        """Full solver output for a batch (B, n, 3) Bohr.  forces=True also keeps what
        the force pass reuses: the S/T centre derivatives and, for a DFT solve at
        Gamma, st['force_grid'] = (X = D phi, v, rho~(G)) of the final density, which
        compute_forces takes (and so releases).

        Per-geometry device state in st stays in slot order (st['order'][i] is
        the geometry held by slot i); the user-facing results mo, D, w1, E, conv,
        nit, Ki, Vep, Vee, Se, occ are returned in input order."""
        tune.apply(self.tuning)
        st = self.setup(xpb, kpt, Lb, forces)
        mo, w1, E, conv, nit, order = self.scf(st)
        B = st["B"]
        D, occ = self.density_matrix(mo, w1)                              # solver.py:758-765
        # the Fock matrix of the final density is needed for HF (exchange enters the
        # residual checks) and for the forces of a cut basis; DFT energies need rho only
        need_v = not self.dft or "lindep" in st
        if forces and not need_v and st.get("kpt") is None:
            V2, E2 = None, self._final_grid_pass(st, B, D)
        else:
            V2, E2 = self.fock2e(st, B, D, mo, occ, full=True, need_v=need_v)
        Ki = _tr(st["T"], D).real
        Vep = _tr(st["V"], D).real
        Se = self.entropy(occ)
        F = st["Hcore"] + V2 if V2 is not None else None       # Fock matrix of the final D
        slot = dict(mo=mo, w1=w1, D=D, occ=occ, E=E, F=F)
        st["slot"] = slot
        st["order"] = order

        def unperm(t):
            out = torch.empty_like(t)
            out[order] = t
            return out

        st.update({k: unperm(t) for k, t in dict(mo=mo, w1=w1, D=D, occ=occ, E=E, conv=conv,
                                                 nit=nit, Ki=Ki, Vep=Vep, Vee=E2,
                                                 Se=Se, F=F).items() if t is not None})
        st["unperm"] = unperm
        return st
