"""make_pbc_gto and make_slater: drop-ins for hqc.pbc.gto.make_pbc_gto's point
evaluation (hqc/pbc/gto.py:6-431) and hqc.pbc.slater.make_slater
(hqc/pbc/slater.py:7-104), the VMC consumers of the SCF orbitals.

AO values at electron positions come from one kernel (hqc_ao_points: the 1-D
image sums over -rcut//L..rcut//L, with the Bloch phase at a k-point); the
Slater matrices are their product with the occupied MO columns.  Every call
also takes batches: xe (W, n, 3) for W walkers (xp (n, 3) shared or (W, n, 3)),
giving slater_up/dn of shape (W, n/2, n/2).
"""
import numpy as np
import torch

from .. import _lib as K
from ..basis import HBasis


def _t(x, device, dtype=torch.float64):
    return torch.as_tensor(np.asarray(x) if not torch.is_tensor(x) else x,
                           dtype=dtype).to(device)


class _PointAO:
    def __init__(self, basis, L_bohr, rcut, device):
        b = HBasis(basis)
        self.L, self.tmax = float(L_bohr), int(rcut // L_bohr)     # gto.py:48-49
        self.nloc, self.nl = b.nloc, b.l_max + 1
        self.alpha = torch.tensor(b.alpha, dtype=torch.float64, device=device)
        self.coef = torch.tensor(b.coef, dtype=torch.float64, device=device).contiguous()
        self.comp = torch.tensor(b.comp, dtype=torch.int32, device=device)
        self.device = device

    def __call__(self, xp, xe, kpt=None):
        """xp (..., n, 3), xe (..., ne, 3) Bohr (same leading shape, or xp unbatched);
        returns (..., ne, n_ao), complex128 with kpt (..., 3)."""
        xp, xe = _t(xp, self.device), _t(xe, self.device)
        lead = xe.shape[:-2]
        xe2 = xe.reshape(-1, xe.shape[-2], 3)
        W = xe2.shape[0]
        xp2 = xp.reshape(-1, xp.shape[-2], 3)
        if xp2.shape[0] == 1 and W > 1:
            xp2 = xp2.expand(W, -1, -1)
        xp2 = xp2.contiguous()
        n = xp2.shape[1]
        k2 = None
        if kpt is not None:
            k2 = _t(kpt, self.device).reshape(-1, 3)
            k2 = (k2.expand(W, 3) if k2.shape[0] == 1 else k2).contiguous()
        out = torch.empty(W, xe2.shape[1], n * self.nloc, device=self.device,
                          dtype=torch.float64 if k2 is None else torch.complex128)
        K.ao_points(xp2, xe2.contiguous(), k2, self.L, self.tmax, self.alpha, self.nl,
                    self.coef, self.comp, out)
        return out.reshape(*lead, xe.shape[-2], n * self.nloc)


def make_pbc_gto(basis, L, rcut=24, gamma=True, device="cuda"):
    """eval_pbc_ao(xp, xe) -> (n_ao,) at one electron (or (..., n_ao) for xe (..., 3)),
    or eval_pbc_ao_kpt(xp, xe, kpt) complex; L, xp, xe in Bohr (hqc/pbc/gto.py:6-21)."""
    ev = _PointAO(basis, L, rcut, device)

    def eval_pbc_ao(xp, xe, kpt=None):
        return ev(xp, _t(xe, device).unsqueeze(-2), kpt)[..., 0, :]

    if gamma:
        return lambda xp, xe: eval_pbc_ao(xp, xe)
    return lambda xp, xe, kpt: eval_pbc_ao(xp, xe, kpt)


def make_slater(n, L, rs, basis, rcut=24, groundstate=True, gamma=True, device="cuda"):
    """Same signature and call forms as hqc/pbc/slater.py:7-36:
        gamma, groundstate:     slater(xp, xe, mo_coeff)
        k-point, groundstate:   slater(xp, xe, mo_coeff, kpt)
        gamma, excited:         slater(xp, xe, mo_coeff, state_idx)
        k-point, excited:       slater(xp, xe, mo_coeff, kpt, state_idx)
    xp, xe in rs units; returns (slater_up (n/2, n/2), slater_dn (n/2, n/2)),
    or (W, n/2, n/2) each for xe (W, n, 3)."""
    assert n % 2 == 0
    ev = _PointAO(basis, L * rs, rcut, device)
    h = n // 2
    ground = torch.cat([torch.arange(h), torch.arange(h)])

    def slater(xp, xe, mo_coeff, kpt, state_idx):
        xp = _t(xp, device)
        assert xp.shape[-2] == n
        ao = ev(xp * rs, _t(xe, device) * rs, kpt)                 # (..., n, n_ao)
        mo = _t(mo_coeff, device, torch.complex128 if kpt is not None else torch.float64)
        idx = torch.as_tensor(state_idx).to(mo.device).long()
        up = ao[..., :h, :] @ mo[..., idx[:h]]
        dn = ao[..., h:, :] @ mo[..., idx[h:]]
        return up, dn

    if groundstate:
        if gamma:
            return lambda xp, xe, mo_coeff: slater(xp, xe, mo_coeff, None, ground)
        return lambda xp, xe, mo_coeff, kpt: slater(xp, xe, mo_coeff, kpt, ground)
    if gamma:
        return lambda xp, xe, mo_coeff, state_idx: slater(xp, xe, mo_coeff, None, state_idx)
    return slater
