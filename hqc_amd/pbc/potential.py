"""Ewald sums on the GPU (hqc/pbc/potential.py:6-130).

potential_energy_pp(xp, L, rs, kappa=10, Gmax=15) -> V_pp in Ry, same units
and conventions as the reference (xp and L in rs units, minimum image in
fractional coordinates).  `ewald_pp` additionally returns -dE/dR in Ha/Bohr.
potential_energy(x, kappa, G, L, rs) -> (v_pp, v_ep, v_ee) for batches of
proton + electron configurations (the VMC observables path).
"""
import math

import numpy as np
import torch

from .. import _lib as K

_GCACHE = {}


def kpoints(dim, Gmax, device):
    """Integer vectors with 0 < |G|^2 <= Gmax^2 (potential.py:6-16)."""
    key = (dim, Gmax, str(device))
    if key not in _GCACHE:
        r = np.arange(-Gmax, Gmax + 1)
        G = np.array(np.meshgrid(*([r] * dim), indexing="ij")).reshape(dim, -1).T
        G2 = (G ** 2).sum(-1)
        G = G[(G2 <= Gmax ** 2) & (G2 > 0)].astype(float)
        _GCACHE[key] = torch.tensor(G, dtype=torch.float64, device=device).contiguous()
    return _GCACHE[key]


def _weights(G, kappa):
    Gn2 = (G * G).sum(-1)
    return torch.exp(-math.pi ** 2 * Gn2 / kappa ** 2) / (math.pi * Gn2)


def madelung(kappa, G):
    """potential.py:18-33 (dim 3)."""
    return _weights(G, kappa).sum() - math.pi / kappa ** 2 - 2 * kappa / math.sqrt(math.pi)


def ewald_pp(xp, L, rs, kappa=10, Gmax=15, forces=False):
    """xp (B, n, 3) rs units on the GPU -> E (B) Ry [, F (B, n, 3) Ha/Bohr].

    The reference's pair sum (potential.py:35-67): per pair erfc(kappa d)/d + sum_G
    w_G cos(2 pi G.r_ij) - pi/kappa^2 in fractional coordinates.  The real-space
    terms are taken per pair; the G sums summed over the pairs come from structure
    factors (hqc_ewald_recip), n N_G phases per geometry instead of n^2/2 N_G."""
    B, n, _ = xp.shape
    dev = xp.device
    x = xp - L * torch.floor(xp / L)
    frac = (x / L).contiguous()
    i, j = torch.triu_indices(n, n, offset=1, device=dev)
    rij = frac[:, i] - frac[:, j]
    rij = rij - torch.round(rij)
    G = kpoints(3, Gmax, dev)
    gk = _weights(G, kappa).contiguous()
    Erec, grec = K.ewald_recip(frac, G, gk)
    d = torch.linalg.vector_norm(rij, dim=-1)
    er = torch.special.erfc(kappa * d)
    g0 = -math.pi / kappa ** 2
    psum = (er / d).sum(-1) + Erec + i.numel() * g0
    E = 2 * psum / L / rs + n * madelung(kappa, G) / L / rs
    if not forces:
        return E
    Lb = L * rs
    dsr = -(er / (d * d) + 2.0 * kappa / math.sqrt(math.pi) * torch.exp(-(kappa * d) ** 2) / d)
    grad = (dsr / d).unsqueeze(-1) * rij / Lb ** 2                 # d(pair)/d r_ij
    # pair terms into an antisymmetric (n, n) array, then a fixed-order row sum
    # (index_add_ would accumulate with atomics, in no fixed order)
    Mp = torch.zeros(xp.shape[0], n, n, 3, dtype=grad.dtype, device=dev)
    Mp[:, i, j] = -grad
    Mp[:, j, i] = grad
    F = -grec / Lb ** 2 + Mp.sum(2)
    return E, F


def potential_energy_pp(xp, L, rs, kappa=10, Gmax=15):
    """Reference signature: one geometry (n, 3) -> scalar tensor (Ry)."""
    xp = torch.as_tensor(xp, dtype=torch.float64, device="cuda")
    single = xp.dim() == 2
    E = ewald_pp(xp.unsqueeze(0) if single else xp, L, rs, kappa, Gmax)
    return E[0] if single else E


def potential_energy(x, kappa, G, L, rs):
    """hqc/pbc/potential.py:70-101, batched like its jax.vmap: x (W, n, 3) (or (n, 3))
    in rs units, first n/2 rows protons (Z = +1), last n/2 electrons (Z = -1).
    Returns (v_pp, v_ep, v_ee), each 2 rs / L times the psi pair sums of like
    protons, unlike pairs (negated) and like electrons ("Ry/rs^2" as there)."""
    x = torch.as_tensor(np.asarray(x) if not torch.is_tensor(x) else x,
                        dtype=torch.float64).to("cuda")
    single = x.dim() == 2
    if single:
        x = x.unsqueeze(0)
    W, n, _ = x.shape
    dev = x.device
    x = x - L * torch.floor(x / L)
    i, j = torch.triu_indices(n, n, offset=1, device=dev)
    rij = (x[:, i] - x[:, j]) / L
    rij = (rij - torch.round(rij)).contiguous()
    Gd = torch.as_tensor(np.asarray(G) if not torch.is_tensor(G) else G,
                         dtype=torch.float64).to(dev).contiguous()
    gk = _weights(Gd, kappa).contiguous()
    out = torch.empty(W, rij.shape[1], 4, dtype=torch.float64, device=dev)
    K.ewald_pairs(rij, Gd, gk, float(kappa), -math.pi / kappa ** 2, out)
    v = out[..., 0]
    Z = torch.cat([torch.ones(n // 2), -torch.ones(n // 2)]).to(dev)
    tc = Z[i] + Z[j]
    f = 2 * rs / L
    vpp = f * (v * (tc == 2)).sum(-1)
    vep = -f * (v * (tc == 0)).sum(-1)
    vee = f * (v * (tc == -2)).sum(-1)
    if single:
        return vpp[0], vep[0], vee[0]
    return vpp, vep, vee
