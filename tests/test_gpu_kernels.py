"""Kernel-level parity on the GPU: each C-ABI entry point against the oracle
(or a plain torch FP64 reference for the pure linear-algebra kernels)."""
import math

import numpy as np
import pytest
import torch

from oracle import cpu_ref as R
from q8_digits import q8_density_digits, q8_digits, q8_exponent

pytestmark = pytest.mark.gpu

DEV = "cuda"


@pytest.fixture(scope="module")
def K():
    from hqc_amd import _lib
    _lib.lib()
    return _lib


def _sys(n=4, rs=1.5, basis="gth-dzvp", grid_length=0.3, seed=0):
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    xp = np.random.default_rng(seed).uniform(0, L, (n, 3))
    return R.System(n, L, rs, basis, grid_length=grid_length), xp * rs, L


@pytest.mark.parametrize("nao_pad,g", [(16, 7), (48, 13), (112, 21), (176, 9), (256, 13),
                                       (640, 7)])
def test_density_and_vmat_f64(K, nao_pad, g):
    print(f"[density/vmat f64] nao_pad={nao_pad} g={g}", flush=True)
    torch.manual_seed(0)
    N = g ** 3
    ld = (N + 63) // 64 * 64
    phi = torch.zeros(2, nao_pad, ld, dtype=torch.float64, device=DEV)
    phi[:, :, :N] = torch.randn(2, nao_pad, N, dtype=torch.float64, device=DEV)
    A = torch.randn(2, nao_pad, nao_pad, dtype=torch.float64, device=DEV)
    D = (A + A.transpose(1, 2)).contiguous()
    rho = torch.zeros(2, ld, dtype=torch.float64, device=DEV)
    X = torch.zeros_like(phi)
    K.ao_density(phi, D, rho, X)
    Xref = D @ phi
    rref = (phi * Xref).sum(1)
    torch.cuda.synchronize()
    assert torch.allclose(X, Xref, rtol=1e-12, atol=1e-10)
    assert torch.allclose(rho, rref, rtol=1e-12, atol=1e-9)
    v = torch.zeros(2, ld, dtype=torch.float64, device=DEV)
    v[:, :N] = torch.randn(2, N, dtype=torch.float64, device=DEV)
    V = torch.empty(2, nao_pad, nao_pad, dtype=torch.float64, device=DEV)
    work = torch.empty(K.vmat_workspace(2, nao_pad, ld, 64) // 8 + 1, dtype=torch.float64,
                       device=DEV)
    K.ao_vmat(phi, v, 0.5, V, work)
    Vref = 0.5 * (phi * v[:, None, :]) @ phi.transpose(1, 2)
    torch.cuda.synchronize()
    assert torch.allclose(V, Vref, rtol=1e-11, atol=1e-9 * math.sqrt(N))
    assert torch.equal(V, V.transpose(1, 2))


@pytest.mark.parametrize("nao_pad,g", [(48, 13), (112, 21), (224, 13), (272, 11)])
def test_density_and_vmat_f32(K, nao_pad, g):
    print(f"[density/vmat f32] nao_pad={nao_pad} g={g}", flush=True)
    torch.manual_seed(1)
    N = g ** 3
    ld = (N + 63) // 64 * 64
    phi = torch.zeros(1, nao_pad, ld, dtype=torch.float64, device=DEV)
    phi[:, :, :N] = torch.rand(1, nao_pad, N, dtype=torch.float64, device=DEV)
    A = torch.rand(1, nao_pad, nao_pad, dtype=torch.float64, device=DEV)
    D = (A + A.transpose(1, 2)).contiguous()
    phi32 = phi.float().contiguous()
    rho = torch.zeros(1, ld, dtype=torch.float64, device=DEV)
    K.ao_density(phi32, D, rho)
    rref = (phi * (D @ phi)).sum(1)
    assert torch.allclose(rho, rref, rtol=2e-5)
    v = torch.zeros(1, ld, dtype=torch.float64, device=DEV)
    v[:, :N] = torch.rand(1, N, dtype=torch.float64, device=DEV)
    V = torch.empty(1, nao_pad, nao_pad, dtype=torch.float64, device=DEV)
    work = torch.empty(K.vmat_workspace(1, nao_pad, ld, 32) // 8 + 1, dtype=torch.float64,
                       device=DEV)
    K.ao_vmat(phi32, v, 1.0, V, work)
    Vref = (phi * v[:, None, :]) @ phi.transpose(1, 2)
    assert torch.allclose(V, Vref, rtol=2e-5)


@pytest.mark.parametrize("nao_pad,g,B", [(16, 5, 1), (48, 13, 3), (112, 21, 2), (128, 9, 2)])
def test_vmat_split16(K, nao_pad, g, B):
    """Split-precision V (FP16 hi/lo products, FP32 accumulation) against FP64:
    error within 2^-19 of sum |phi_p phi_q v| per element (operands carry 22 bits,
    products accumulate in FP32 over a block's chunks), for signed v of very
    different size per geometry, and through a geometry index list."""
    print(f"[vmat split16] nao_pad={nao_pad} g={g} B={B}", flush=True)
    torch.manual_seed(3)
    N = g ** 3
    ld = (N + 63) // 64 * 64
    phi = torch.zeros(B, nao_pad, ld, dtype=torch.float64, device=DEV)
    phi[:, :, :N] = torch.randn(B, nao_pad, N, dtype=torch.float64, device=DEV)
    v = torch.zeros(B, ld, dtype=torch.float64, device=DEV)
    v[:, :N] = torch.randn(B, N, dtype=torch.float64, device=DEV)
    v *= torch.logspace(0, -9, B, dtype=torch.float64, device=DEV)[:, None]
    scale = 2.0 ** math.floor(math.log2(2.0 ** 14 / phi.abs().max().item()))
    ph = torch.empty(B, nao_pad, ld, dtype=torch.int32, device=DEV)
    K.phi_split_f16(phi, scale, ph)
    work = torch.empty(K.vmat_h2_workspace(B, nao_pad, ld) // 8 + 1, dtype=torch.float64,
                       device=DEV)
    V = torch.empty(B, nao_pad, nao_pad, dtype=torch.float64, device=DEV)
    K.ao_vmat_h2(ph, v, scale, 0.5, V, work)
    Vref = 0.5 * (phi * v[:, None, :]) @ phi.transpose(1, 2)
    bound = 0.5 * (phi.abs() * v.abs()[:, None, :]) @ phi.abs().transpose(1, 2)
    torch.cuda.synchronize()
    r = ((V - Vref).abs() / bound.clamp(min=1e-300)).max().item()
    print(f"  max |V - Vref| / sum|phi phi v| = {r:.3e} (2^{math.log2(max(r, 1e-300)):.1f})",
          flush=True)
    assert ((V - Vref).abs() <= 2.0 ** -19 * bound + 1e-300).all()
    assert torch.equal(V, V.transpose(1, 2))
    # the caller's max |v| gives the same scales, hence the same result
    V3 = torch.empty_like(V)
    K.ao_vmat_h2(ph, v, scale, 0.5, V3, work, vmax=v.abs().amax(-1))
    assert torch.equal(V3, V)
    # geometry index list: entries read phi geometries in reverse order
    bidx = torch.arange(B - 1, -1, -1, dtype=torch.int32, device=DEV)
    V2 = torch.empty_like(V)
    K.ao_vmat_h2(ph, v, scale, 0.5, V2, work, bidx)
    Vref2 = 0.5 * (phi.flip(0) * v[:, None, :]) @ phi.flip(0).transpose(1, 2)
    bound2 = 0.5 * (phi.flip(0).abs() * v.abs()[:, None, :]) @ phi.flip(0).abs().transpose(1, 2)
    assert ((V2 - Vref2).abs() <= 2.0 ** -19 * bound2 + 1e-300).all()


def test_mixed_v_prep(K):
    """The fused operand pass of the mixed V step equals the separate torch
    operations bit for bit (max |v|, max |v - v_a|, select), NaN included."""
    torch.manual_seed(9)
    B, ld = 5, 4160
    v = torch.randn(B, ld, dtype=torch.float64, device=DEV) * 10.0
    va = v + 1e-6 * torch.randn(B, ld, dtype=torch.float64, device=DEV)
    va[3] = 0.0
    v[4, 77] = float("nan")
    direct = torch.tensor([True, False, True, False, False], device=DEV)
    vv = torch.empty_like(v)
    vmax = torch.empty(B, dtype=torch.float64, device=DEV)
    dvmax = torch.empty_like(vmax)
    K.mixed_v_prep(v, va, direct, vv, vmax, dvmax)
    dv = v - va
    torch.cuda.synchronize()
    assert torch.equal(vv[:4], torch.where(direct[:, None], v, dv)[:4])
    assert torch.equal(vmax[:4], v.abs().amax(-1)[:4])
    assert torch.equal(dvmax[:4], dv.abs().amax(-1)[:4])
    assert torch.isnan(vmax[4]) and torch.isnan(dvmax[4])


@pytest.mark.parametrize("nao_pad,m,dt", [(48, 20, torch.float64), (112, 60, torch.float64),
                                           (112, 60, torch.float32), (112, 100, torch.float32),
                                           (64, 140, torch.float64), (224, 100, torch.float64),
                                           (640, 64, torch.float64)])
def test_orb_density(K, nao_pad, m, dt):
    """rho from At = C sqrt(f) equals the D-form density with D = C f C^T."""
    torch.manual_seed(2)
    g = 11
    N = g ** 3
    ld = (N + 63) // 64 * 64
    phi = torch.zeros(2, nao_pad, ld, dtype=torch.float64, device=DEV)
    phi[:, :, :N] = torch.randn(2, nao_pad, N, dtype=torch.float64, device=DEV)
    C = torch.randn(2, nao_pad, m, dtype=torch.float64, device=DEV)
    f = torch.rand(2, m, dtype=torch.float64, device=DEV) * 2
    m_pad = (m + 15) // 16 * 16
    At = torch.zeros(2, nao_pad, m_pad, dtype=torch.float64, device=DEV)
    At[:, :, :m] = C * f.sqrt()[:, None, :]
    rho = torch.zeros(2, ld, dtype=torch.float64, device=DEV)
    K.orb_density(phi.to(dt).contiguous(), At, rho)
    D = (C * f[:, None, :]) @ C.transpose(1, 2)
    rref = (phi * (D @ phi)).sum(1)
    tol = 1e-11 if dt == torch.float64 else 1e-4
    assert torch.allclose(rho, rref, rtol=tol, atol=tol * rref.abs().max().item())


@pytest.mark.parametrize("nao_pad,m", [(16, 8), (48, 20), (112, 60), (112, 100), (64, 140)])
def test_orb_density_split16(K, nao_pad, m):
    """Orbital density from FP16 (hi, lo) pairs with split coefficients: within
    2^-18 sum_i (sum_k |At_ki phi_k|)^2 of the FP64 value (psi carries ~2^-22)."""
    print(f"[orb density split16] nao_pad={nao_pad} m={m}", flush=True)
    torch.manual_seed(5)
    g = 11
    N = g ** 3
    ld = (N + 63) // 64 * 64
    B = 2
    phi = torch.zeros(B, nao_pad, ld, dtype=torch.float64, device=DEV)
    phi[:, :, :N] = torch.randn(B, nao_pad, N, dtype=torch.float64, device=DEV)
    m_pad = (m + 15) // 16 * 16
    At = torch.zeros(B, nao_pad, m_pad, dtype=torch.float64, device=DEV)
    At[:, :, :m] = torch.randn(B, nao_pad, m, dtype=torch.float64, device=DEV)
    At[1] *= 1e-3                                   # a different coefficient scale
    scale = 2.0 ** math.floor(math.log2(2.0 ** 14 / phi.abs().max().item()))
    ph = torch.empty(B, nao_pad, ld, dtype=torch.int32, device=DEV)
    K.phi_split_f16(phi, scale, ph)
    rho = torch.zeros(B, ld, dtype=torch.float64, device=DEV)
    K.orb_density(ph, At, rho, phi_scale=scale)
    psi = At.transpose(1, 2) @ phi
    rref = (psi ** 2).sum(1)
    bound = ((At.abs().transpose(1, 2) @ phi.abs()) ** 2).sum(1)
    torch.cuda.synchronize()
    assert ((rho - rref).abs() <= 2.0 ** -18 * bound + 1e-300).all()
    # through a geometry index list
    bidx = torch.tensor([1, 0], dtype=torch.int32, device=DEV)
    rho2 = torch.zeros_like(rho)
    K.orb_density(ph, At, rho2, bidx, phi_scale=scale)
    rref2 = ((At.transpose(1, 2) @ phi.flip(0)) ** 2).sum(1)
    bound2 = ((At.abs().transpose(1, 2) @ phi.flip(0).abs()) ** 2).sum(1)
    assert ((rho2 - rref2).abs() <= 2.0 ** -18 * bound2 + 1e-300).all()


@pytest.mark.parametrize("nao_pad,m,ns", [(16, 8, 4), (48, 20, 3), (112, 60, 4), (112, 100, 5),
                                          (128, 140, 4)])
def test_orb_density_q8(K, nao_pad, m, ns):
    """Orbital density from base-256 int8 digits (exact integer products), with
    digit products to level 5 and to level 4: equal to the host restatement of the
    digit algorithm to FP64 rounding, and within the digit error bound of the FP64
    density,
        |rho~ - rho| <= sum_i (2|psi_i| + d_i) d_i,
        d_i(r) = nao_pad max_k|At_ki| max_k|phi_k(r)| (2^(-6-8(ns-1)) + ns 2^(-3-8 top))."""
    print(f"[orb density q8] nao_pad={nao_pad} m={m} ns={ns}", flush=True)
    torch.manual_seed(7)
    g = 11
    N = g ** 3
    ld = (N + 63) // 64 * 64
    B = 2
    phi = torch.zeros(B, nao_pad, ld, dtype=torch.float64, device=DEV)
    phi[:, :, :N] = torch.randn(B, nao_pad, N, dtype=torch.float64, device=DEV)
    # per-point magnitudes over 30 binary orders, some points all zero
    phi[:, :, :N] *= torch.exp2(-torch.randint(0, 30, (B, 1, N), device=DEV).double())
    phi[:, :, 5:9] = 0.0
    # values whose digit expansion carries (a remainder of exactly 1/2 after an even
    # digit): at point 20 of geometry 0 the maximum is 1 (y_0 = 32 phi), AOs 1-7 hold
    # (n + 1/2) / 32 with n even (the tie at digit 0), AO 8 one at digit 1
    phi[0, :, 20] = 0.25
    phi[0, 0, 20] = 1.0
    phi[0, 1:8, 20] = (16.5 + 2.0 * torch.arange(7, device=DEV)) / 32
    phi[0, 8, 20] = (10 + 64.5 / 256) / 32
    m_pad = (m + 15) // 16 * 16
    At = torch.zeros(B, nao_pad, m_pad, dtype=torch.float64, device=DEV)
    At[:, :, :m] = torch.randn(B, nao_pad, m, dtype=torch.float64, device=DEV)
    At[:, :, :m] *= torch.exp(-0.2 * torch.arange(m, device=DEV, dtype=torch.float64))
    At[1] *= 1e-3
    At[0, :, m // 2] = 0.0                              # an empty orbital column
    q = torch.empty(B, ld * ns * nao_pad, dtype=torch.int8, device=DEV)
    pe = torch.empty(B, ld, dtype=torch.int16, device=DEV)
    amax = torch.full((B,), float("nan"), dtype=torch.float64, device=DEV)
    K.phi_slice_i8(phi, ns, q, pe, amax)
    torch.cuda.synchronize()
    assert torch.equal(amax, phi.abs().amax(dim=(1, 2)))       # max |phi| per geometry
    ph, A = phi.cpu().numpy(), At.cpu().numpy()
    # the stored digits themselves: [tile][digit][AO block][point][AO]
    er = q8_exponent(np.abs(ph).max(1))
    xd = q8_digits(ph, er[:, None, :], ns)
    qq = q.cpu().numpy().reshape(B, ld // 16, ns, nao_pad // 16, 16, 16)
    want = np.stack(xd, 0).reshape(ns, B, nao_pad // 16, 16, ld // 16, 16)
    assert np.array_equal(qq, want.transpose(1, 4, 0, 2, 5, 3).astype(np.int8))
    assert np.array_equal(pe.cpu().numpy(), er)
    psi = np.matmul(A.transpose(0, 2, 1), ph)                    # (B, m, ld)
    rref = (psi ** 2).sum(1)
    for top in (5, 4):
        rho = torch.full((B, ld), float("nan"), dtype=torch.float64, device=DEV)
        K.orb_density_q8(q, pe, At, rho, ns, top=top)
        assert torch.isfinite(rho).all()
        emul = q8_density_digits(ph, A, ns, top)
        r = rho.cpu().numpy()
        assert np.allclose(r, emul, rtol=1e-13, atol=1e-300), top
        d = (nao_pad * np.abs(A).max(1)[:, :, None] * np.abs(ph).max(1)[:, None, :]
             * (2.0 ** (-6 - 8 * (ns - 1)) + ns * 2.0 ** (-3 - 8 * top)))
        bound = ((2 * np.abs(psi) + d) * d).sum(1)
        assert (np.abs(r - rref) <= bound).all(), top
        # through a geometry index list: slot 0 reads geometry 1 and slot 1 geometry 0
        bidx = torch.tensor([1, 0], dtype=torch.int32, device=DEV)
        rho2 = torch.empty_like(rho)
        K.orb_density_q8(q, pe, At, rho2, ns, bidx, top=top)
        emul2 = q8_density_digits(ph[::-1].copy(), A, ns, top)
        assert np.allclose(rho2.cpu().numpy(), emul2, rtol=1e-13, atol=1e-300), top
        # the first 3 of the stored digits (the prefix of the ns-digit expansion)
        if ns > 3:
            rho3 = torch.empty_like(rho)
            K.orb_density_q8(q, pe, At, rho3, ns, nuse=3, top=top)
            assert np.allclose(rho3.cpu().numpy(), q8_density_digits(ph, A, 3, top, ns),
                               rtol=1e-13, atol=1e-300), top
    # level 3 (9 digit products) from the first 3 stored digits, the early SCF
    # iterations' form: the host restatement and the error bound with 3 digits
    rho3 = torch.full((B, ld), float("nan"), dtype=torch.float64, device=DEV)
    K.orb_density_q8(q, pe, At, rho3, ns, nuse=3, top=3)
    r3 = rho3.cpu().numpy()
    assert np.allclose(r3, q8_density_digits(ph, A, 3, 3, ns), rtol=1e-13, atol=1e-300)
    d = (nao_pad * np.abs(A).max(1)[:, :, None] * np.abs(ph).max(1)[:, None, :]
         * (2.0 ** (-6 - 8 * 2) + 3 * 2.0 ** (-3 - 8 * 3)))
    assert (np.abs(r3 - rref) <= ((2 * np.abs(psi) + d) * d).sum(1)).all()
    if ns > 3:
        # level 3 from 4 digits (10 products): the coarse accurate form
        K.orb_density_q8(q, pe, At, rho3, ns, nuse=4, top=3)
        r4 = rho3.cpu().numpy()
        assert np.allclose(r4, q8_density_digits(ph, A, 4, 3, ns), rtol=1e-13, atol=1e-300)
        d = (nao_pad * np.abs(A).max(1)[:, :, None] * np.abs(ph).max(1)[:, None, :]
             * (2.0 ** (-6 - 8 * 3) + 4 * 2.0 ** (-3 - 8 * 3)))
        assert (np.abs(r4 - rref) <= ((2 * np.abs(psi) + d) * d).sum(1)).all()
    if ns > 4:
        with pytest.raises(RuntimeError):               # level 3 takes 3 or 4 phi digits
            K.orb_density_q8(q, pe, At, rho3, ns, nuse=5, top=3)


@pytest.mark.parametrize("g", [1, 2, 5, 7, 13, 17, 24, 25, 27, 31, 33, 48, 49, 51, 55, 57, 61,
                               64, 65, 69, 71, 97, 127, 128])
def test_fft3d(K, g):
    print(f"[fft3d] g={g}", flush=True)
    torch.manual_seed(g)
    x = torch.randn(2, g, g, g, dtype=torch.complex128, device=DEV)
    y = x.clone()
    K.fft3d(y, -1)
    ref = torch.fft.fftn(x, dim=(1, 2, 3))
    assert torch.allclose(y, ref, rtol=1e-12, atol=1e-11 * g ** 1.5)
    K.fft3d(y, +1)
    assert torch.allclose(y / g ** 3, x, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("g", [7, 25, 49, 61, 65, 71, 97, 128])
def test_fft3d_f32(K, g):
    """complex64 transforms (plane and line kernels, prime lengths on the direct pass)
    against torch's complex128 FFT of the same inputs: FP32 rounding only, ~1e-7 of the
    typical output g^1.5."""
    print(f"[fft3d f32] g={g}", flush=True)
    torch.manual_seed(g)
    x = torch.randn(3, g, g, g, dtype=torch.complex64, device=DEV)
    y = x.clone()
    K.fft3d(y, -1)
    ref = torch.fft.fftn(x.to(torch.complex128), dim=(1, 2, 3))
    err = (y.to(torch.complex128) - ref).abs().max().item()
    assert err < 2e-6 * g ** 1.5, err
    K.fft3d(y, +1)
    back = (y.to(torch.complex128) / g ** 3 - x.to(torch.complex128)).abs().max().item()
    assert back < 2e-6, back


@pytest.mark.parametrize("g", [43, 49, 51])
def test_fft3d_plane_thread_counts_agree(K, g):
    """Planes up to 51^2 run with 256 threads once there are >= 512 planes (12
    grids here), with 1024 threads otherwise (2 grids): the same outputs bit for
    bit, and both equal to torch's FFT."""
    print(f"[fft3d plane threads] g={g}", flush=True)
    torch.manual_seed(100 + g)
    x = torch.randn(12, g, g, g, dtype=torch.complex128, device=DEV)
    y = x.clone()
    K.fft3d(y, -1)                                   # 12 g planes >= 512: 256 threads
    y2 = x[:2].clone()
    K.fft3d(y2, -1)                                  # 2 g planes < 512: 1024 threads
    ref = torch.fft.fftn(x, dim=(1, 2, 3))
    torch.cuda.synchronize()
    assert 12 * g >= 512 and 2 * g < 512
    assert torch.equal(y[:2], y2)
    assert torch.allclose(y, ref, rtol=1e-12, atol=1e-11 * g ** 1.5)


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
@pytest.mark.parametrize("g", [26, 34, 45, 49, 57, 61, 64, 65, 97])
def test_fft3d_schedules_agree(K, g, dtype):
    """In-place LDS passes (1 or 2 work items per thread; 256, 512 or 1024 threads
    per plane) and the two-buffer passes form every output with the same
    operations: all schedules give the same bits.  Sizes cover radix 2/3/4/5/7/13,
    the generic radix (34 = 2*17, 57 = 3*19), the direct prime pass (61, 97), and
    grids whose in-place planes fit where two-buffer planes do not (97)."""
    from hqc_amd import tune
    print(f"[fft3d schedules] g={g} {dtype}", flush=True)
    torch.manual_seed(300 + g)
    x = torch.randn(3, g, g, g, dtype=dtype, device=DEV)
    try:
        tune.apply({"fft_two_buffer": 1})
        ref = x.clone()
        K.fft3d(ref, -1)
        for nt in (0, 256, 512, 1024):
            for mi in (0, 1, 2):
                tune.apply({"fft_plane_threads": nt, "fft_ip_items": mi})
                y = x.clone()
                K.fft3d(y, -1)
                torch.cuda.synchronize()
                assert torch.equal(y, ref), (nt, mi)
    finally:
        tune.apply({})
    exact = torch.fft.fftn(x.to(torch.complex128), dim=(1, 2, 3))
    tol = 1e-11 if dtype == torch.complex128 else 2e-6
    assert (ref.to(torch.complex128) - exact).abs().max().item() < tol * g ** 1.5


def test_gto_tables_and_values(K):
    s, xpb, L = _sys()
    print(f"[gto] g={s.g}", flush=True)
    b = s.basis
    P_ref = R.gto_tables(s, xpb, deriv=True)          # nl = l_max + 3
    nl = P_ref.shape[-1]
    xp_t = torch.tensor(xpb[None], device=DEV)
    alpha = torch.tensor(b.alpha, device=DEV)
    P = torch.empty(1, s.n, 3, s.g, len(b.alpha), nl, dtype=torch.float64, device=DEV)
    K.gto_tables(xp_t, s.g, s.L, int(24 // s.L), alpha, nl, P)
    assert np.allclose(P[0].cpu().numpy(), P_ref, rtol=1e-13, atol=1e-14)
    nao_pad = (s.n_ao + 15) // 16 * 16
    ld = (s.N + 63) // 64 * 64
    phi = torch.zeros(1, nao_pad, ld, dtype=torch.float64, device=DEV)
    coef = torch.tensor(b.coef, device=DEV).contiguous()
    comp = torch.tensor(b.comp, dtype=torch.int32, device=DEV)
    K.ao_values(P[..., :b.l_max + 1].contiguous(), coef, comp, nao_pad, ld, phi)
    ref = R.ao_values(s, xpb)
    assert np.allclose(phi[0, :s.n_ao, :s.N].cpu().numpy(), ref, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("basis,grid_length", [("gth-szv", 0.3), ("gth-dzv", 0.2),
                                               ("gth-dzvp", 0.3), ("gth-dzvp", 0.052)])
def test_ao_values_line_form(K, basis, grid_length):
    """AO values along grid lines (lane = z index) against the pointwise form, FP64
    and FP32, for s-only tables (one power) and p functions, on grids of one and
    two lane chunks (25..65+ points per line), at three geometries."""
    s, xpb, L = _sys(n=3, basis=basis, grid_length=grid_length)
    print(f"[ao values lines] {basis} g={s.g}", flush=True)
    b = s.basis
    nl = b.l_max + 1
    rng = np.random.default_rng(s.g)
    xp = torch.tensor(np.stack([xpb, rng.uniform(0, L, xpb.shape), rng.uniform(0, L, xpb.shape)]),
                      device=DEV)
    P = torch.empty(3, s.n, 3, s.g, len(b.alpha), nl, dtype=torch.float64, device=DEV)
    K.gto_tables(xp, s.g, s.L, int(24 // s.L), torch.tensor(b.alpha, device=DEV), nl, P)
    nao_pad = (s.n_ao + 15) // 16 * 16
    ld = (s.N + 255) // 256 * 256
    coef = torch.tensor(b.coef, device=DEV).contiguous()
    comp = torch.tensor(b.comp, dtype=torch.int32, device=DEV)
    for dt in (torch.float64, torch.float32):
        out = []
        for pointwise in (0, 1):
            K.set_tuning("ao_pointwise", pointwise)
            phi = torch.full((3, nao_pad, ld), 7.0, dtype=dt, device=DEV)
            K.ao_values(P, coef, comp, nao_pad, ld, phi)
            out.append(phi)
        K.set_tuning("ao_pointwise", 0)
        lines, pw = out[0][:, :s.n_ao, :s.N], out[1][:, :s.n_ao, :s.N]
        tol = 1e-14 if dt == torch.float64 else 1e-6
        assert (lines - pw).abs().max().item() <= tol * pw.abs().max().item()
        # rows and columns past the AOs and the grid are left as they were
        assert bool((out[0][:, s.n_ao:] == 7).all()) and bool((out[0][:, :, s.N:] == 7).all())
    # max |phi| per geometry from the AO-value pass, exactly, both forms
    for pointwise in (0, 1):
        K.set_tuning("ao_pointwise", pointwise)
        phi = torch.zeros(3, nao_pad, ld, dtype=torch.float64, device=DEV)
        amax = torch.full((3,), 5.0, dtype=torch.float64, device=DEV)
        K.ao_values(P, coef, comp, nao_pad, ld, phi, amax=amax)
        assert torch.equal(amax, phi.abs().amax((1, 2))), pointwise
    K.set_tuning("ao_pointwise", 0)


@pytest.mark.parametrize("ns", [3, 4, 5])
def test_phi_slice_with_pairs(K, ns):
    """hqc_phi_slice_i8_h2 writes the digits and exponents of hqc_phi_slice_i8 and
    the FP16 pairs of hqc_phi_split_f16 from one read of phi, bit for bit (values
    of both signs over eight binades, a zero row, a column of one large value)."""
    print(f"[slice + pairs] ns={ns}", flush=True)
    gen = torch.Generator().manual_seed(ns)
    G, nao_pad, ld = 3, 112, 128 * 5
    phi = (torch.rand(G, nao_pad, ld, generator=gen, dtype=torch.float64) - 0.5) * \
        torch.exp2(-8 * torch.rand(G, nao_pad, ld, generator=gen, dtype=torch.float64))
    phi[:, 7] = 0.0
    phi[1, :, 33] = 0.0
    phi[1, 5, 33] = 3.75
    phi = phi.to(DEV)
    scale = 2.0 ** 12
    q = torch.empty(G, ld * ns * nao_pad, dtype=torch.int8, device=DEV)
    pe = torch.empty(G, ld, dtype=torch.int16, device=DEV)
    K.phi_slice_i8(phi, ns, q, pe)
    h2 = torch.empty(G, nao_pad, ld, dtype=torch.int32, device=DEV)
    K.phi_split_f16(phi, scale, h2)
    q2, pe2 = torch.full_like(q, 9), torch.full_like(pe, 9)
    h22 = torch.full_like(h2, 9)
    K.phi_slice_i8_h2(phi, ns, scale, q2, pe2, h22)
    assert torch.equal(q, q2) and torch.equal(pe, pe2) and torch.equal(h2, h22)


def test_overlap_kinetic(K):
    for basis in ("gth-szv", "gth-dzv", "gth-dzvp"):
        s, xpb, L = _sys(n=3, basis=basis)
        print(f"[S/T] {basis}", flush=True)
        b = s.basis
        S_ref, T_ref, dS_ref, dT_ref = R.overlap_kinetic(s, xpb, deriv=True)
        nao = s.n_ao
        S = torch.empty(1, nao, nao, dtype=torch.float64, device=DEV)
        T = torch.empty_like(S)
        dS = torch.empty(1, nao, nao, 3, dtype=torch.float64, device=DEV)
        dT = torch.empty_like(dS)
        K.overlap_kinetic(torch.tensor(xpb[None], device=DEV),
                          torch.tensor(s.lattice, device=DEV).contiguous(),
                          torch.tensor(b.alpha, device=DEV), torch.tensor(b.coef, device=DEV),
                          torch.tensor(b.comp, dtype=torch.int32, device=DEV), S, T, dS, dT)
        for a, r in ((S, S_ref), (T, T_ref), (dS, dS_ref), (dT, dT_ref)):
            assert np.allclose(a[0].cpu().numpy(), r, rtol=1e-12, atol=1e-13)


def test_overlap_kinetic_batch(K):
    """Two geometries in one launch (each thread writes its pair and the mirrored
    one): both match the oracle, including the centre derivatives."""
    s, xa, L = _sys(n=3, basis="gth-dzvp", seed=1)
    _, xb, _ = _sys(n=3, basis="gth-dzvp", seed=2)
    b = s.basis
    nao = s.n_ao
    S = torch.empty(2, nao, nao, dtype=torch.float64, device=DEV)
    T = torch.empty_like(S)
    dS = torch.empty(2, nao, nao, 3, dtype=torch.float64, device=DEV)
    dT = torch.empty_like(dS)
    K.overlap_kinetic(torch.tensor(np.stack([xa, xb]), device=DEV),
                      torch.tensor(s.lattice, device=DEV).contiguous(),
                      torch.tensor(b.alpha, device=DEV), torch.tensor(b.coef, device=DEV),
                      torch.tensor(b.comp, dtype=torch.int32, device=DEV), S, T, dS, dT)
    for i, xp in enumerate((xa, xb)):
        refs = R.overlap_kinetic(s, xp, deriv=True)
        for a, r in zip((S, T, dS, dT), refs):
            assert np.allclose(a[i].cpu().numpy(), r, rtol=1e-12, atol=1e-13)


def test_kpoint_tables_values_and_integrals(K):
    """Bloch-phase tables, stacked complex AO values and complex S/T (+ derivatives)
    against the oracle at a general k-point."""
    s, xpb, L = _sys(n=3, basis="gth-dzvp")
    b = s.basis
    kpt = np.array([0.31, -0.57, 0.12])
    print(f"[k tables] g={s.g}", flush=True)
    P_ref = R.gto_tables(s, xpb, kpt, deriv=True)
    nl = P_ref.shape[-1]
    xp_t = torch.tensor(xpb[None], device=DEV)
    k_t = torch.tensor(kpt[None], device=DEV)
    alpha = torch.tensor(b.alpha, device=DEV)
    P = torch.empty(1, s.n, 3, s.g, len(b.alpha), nl, dtype=torch.complex128, device=DEV)
    K.gto_tables_k(xp_t, k_t, s.g, s.L, int(24 // s.L), alpha, nl, P)
    assert np.allclose(P[0].cpu().numpy(), P_ref, rtol=1e-13, atol=1e-14)
    nao_pad = (s.n_ao + 15) // 16 * 16
    ld = (s.N + 63) // 64 * 64
    phi = torch.zeros(1, 2 * nao_pad, ld, dtype=torch.float64, device=DEV)
    coef = torch.tensor(b.coef, device=DEV).contiguous()
    comp = torch.tensor(b.comp, dtype=torch.int32, device=DEV)
    K.ao_values_k(P[..., :b.l_max + 1].contiguous(), coef, comp, nao_pad, ld, phi)
    ref = R.ao_values(s, xpb, kpt)
    got = phi[0, :s.n_ao, :s.N].cpu().numpy() + 1j * phi[0, nao_pad:nao_pad + s.n_ao, :s.N].cpu().numpy()
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-14)
    S_ref, T_ref, dS_ref, dT_ref = R.overlap_kinetic(s, xpb, kpt, deriv=True)
    nao = s.n_ao
    S = torch.empty(1, nao, nao, dtype=torch.complex128, device=DEV)
    T = torch.empty_like(S)
    dS = torch.empty(1, nao, nao, 3, dtype=torch.complex128, device=DEV)
    dT = torch.empty_like(dS)
    K.overlap_kinetic_k(xp_t, k_t, torch.tensor(s.lattice, device=DEV).contiguous(), alpha,
                        torch.tensor(b.coef, device=DEV), comp, S, T, dS, dT)
    for a, r in ((S, S_ref), (T, T_ref), (dS, dS_ref), (dT, dT_ref)):
        assert np.allclose(a[0].cpu().numpy(), r, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("method,search", [("fermi", "bisect"), ("fermi", "newton"),
                                           ("gauss", "bisect"), ("gauss", "newton")])
def test_occupations(K, method, search):
    rng = np.random.default_rng(3)
    n, nmo = 16, 40
    w = np.sort(rng.normal(0.0, 0.5, (5, nmo)), axis=1)
    occ_fn = R.make_occupation(n, nmo, True, method, 0.05, search, 100, 1e-7)
    wt = torch.tensor(w, device=DEV)
    occ = torch.empty_like(wt)
    mu = torch.empty(5, dtype=torch.float64, device=DEV)
    if method == "fermi":
        shift = -0.05 * math.log(2 * nmo / n - 1)
    else:
        import scipy.special
        shift = -0.05 * scipy.special.erfinv(1 - n / nmo)
    K.occupations(wt, n, 0.05, 0 if method == "fermi" else 1, 0 if search == "bisect" else 1,
                  100, 1e-7, shift, occ, mu)
    for i in range(5):
        assert np.allclose(occ[i].cpu().numpy(), occ_fn(w[i]), rtol=1e-11, atol=1e-13)
        assert abs(mu[i].item() - occ_fn.mu(w[i])) < 1e-11


def test_bisection_return_rule_known_answers(K):
    """The kernel's mu on the hand-worked spectra of the oracle's pin
    (tests/test_oracle_pins.py test_bisection_return_rule): 4 electrons in 4
    orbitals, sigma = 0.01, mu_shift = 0; expected mu for search_cycle 100 / 3 / 2."""
    base = [0.0, 1.0, 2.0, 3.0]
    skew = [0.0, 1.0, 2.0, 4.0]
    cases = [(1.0, base, 100, 1.0), (-5.0, base, 100, -3.5), (-1.0, base, 100, 0.0),
             (-2.0, base, 100, -0.5), (1.0, skew, 100, 2.0), (1.0, skew, 3, 4.25),
             (-5.0, skew, 100, -3.5), (-5.0, skew, 3, -7.25), (-5.0, skew, 2, -6.5)]
    for a, w, cyc, want in cases:
        wt = torch.tensor([[a + x for x in w]], dtype=torch.float64, device=DEV)
        occ = torch.empty_like(wt)
        mu = torch.empty(1, dtype=torch.float64, device=DEV)
        K.occupations(wt, 4, 0.01, 0, 0, cyc, 1e-7, 0.0, occ, mu)
        assert mu.item() == want, (a, w, cyc, mu.item(), want)


@pytest.mark.parametrize("xc", ["lda,vwn", "lda,", "lda,hl", "lda,bh", "lda,gl"])
def test_xc_eval(K, xc):
    g = 9
    N = g ** 3
    ld = (N + 63) // 64 * 64
    rng = np.random.default_rng(4)
    rho_np = rng.uniform(1e-4, 2.0, N)
    vh_np = rng.normal(size=N)
    rho = torch.zeros(1, ld, dtype=torch.float64, device=DEV)
    rho[0, :N] = torch.tensor(rho_np, device=DEV)
    vh = torch.zeros(1, g, g, g, dtype=torch.complex128, device=DEV)
    vh.view(1, -1).real[0] = torch.tensor(vh_np, device=DEV)
    xc_id = {"lda,vwn": 1, "lda,": 2, "lda,hl": 3, "lda,bh": 4, "lda,gl": 5}[xc]
    v = torch.zeros(1, ld, dtype=torch.float64, device=DEV)
    part = torch.empty(1, K.xc_blocks(1, g), 2, dtype=torch.float64, device=DEV)
    K.xc_eval(rho, vh, g, xc_id, v, part)
    eps, vxc = R.xc_eps_v(rho_np, xc)
    assert np.allclose(v[0, :N].cpu().numpy(), vh_np + vxc, rtol=1e-13, atol=1e-13)
    sums = part.sum(1)[0].cpu().numpy()
    assert np.isclose(sums[0], np.sum(rho_np * eps), rtol=1e-12)
    assert np.isclose(sums[1], np.sum(rho_np * vh_np), rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("B", [2, 3])
def test_paired_poisson_matches_single(K, B):
    """Two geometries per complex FFT (Re, Im) give the same v_H + v_xc and energy
    sums as one transform per geometry (the Coulomb kernel is real and even on the
    odd grid), also for an odd batch."""
    from hqc_amd.pbc import make_pes
    n = 4
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    eng = make_pes(n, L, 1.5, "gth-dzv", grid_length=0.3, dft=True).solver.engine
    c = eng.c
    torch.manual_seed(B)
    rho = torch.zeros(B, c.ld, dtype=torch.float64, device=DEV)
    rho[:, :c.N] = torch.rand(B, c.N, dtype=torch.float64, device=DEV) + 1e-3
    v1, E1, _ = eng.xc_potential(rho)                  # paired
    v2, E2, _ = eng.xc_potential(rho, save_rg=True)    # one grid per geometry
    assert torch.allclose(v1, v2, rtol=1e-12, atol=1e-12 * v2.abs().max().item())
    assert torch.allclose(E1, E2, rtol=1e-12, atol=1e-12)


def test_pbe_points(K):
    """Dual-number PBE (e, de/drho, de/dsigma) against the oracle's complex-step values."""
    rng = np.random.default_rng(6)
    rho = rng.uniform(1e-4, 3.0, 500)
    sigma = rng.uniform(0, 5.0, 500) * rho ** (8 / 3)
    rho[:3] = [0.0, 1e-13, 0.5]
    sigma[2] = 0.0
    out = torch.empty(500, 3, dtype=torch.float64, device=DEV)
    K.pbe_points(torch.tensor(rho, device=DEV), torch.tensor(sigma, device=DEV), out)
    ref = np.stack(R.pbe_derivs(rho, sigma), axis=1)
    assert np.allclose(out.cpu().numpy(), ref, rtol=1e-12, atol=1e-15)


def test_gga_potential(K):
    """Engine GGA route (spectral gradient, PBE, divergence) against the oracle's
    discrete functional: v = v_H + v_xc and E_xc."""
    from hqc_amd.pbc.engine import Cell, Engine
    n, rs, gl = 3, 1.4, 0.25
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    s = R.System(n, L, rs, "gth-dzv", grid_length=gl)
    xp = np.random.default_rng(2).uniform(0, s.L, (n, 3))
    phi = R.ao_values(s, xp)
    rho_np = (phi ** 2).sum(0) * 0.5
    eng = Engine(Cell(n, L, rs, "gth-dzv", grid_length=gl), dft=True, xc="pbe,pbe")
    c = eng.c
    print(f"[gga potential] g={c.g}", flush=True)
    rho = torch.zeros(1, c.ld, dtype=torch.float64, device=DEV)
    rho[0, :c.N] = torch.tensor(rho_np, device=DEV)
    v, E, _ = eng.xc_potential(rho)
    Exc, vxc = R.xc_gga(s, rho_np)
    vh, _ = R.poisson(s, rho_np)
    ref = vh + vxc
    assert np.allclose(v[0, :c.N].cpu().numpy(), ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
    assert abs(E[0, 0].item() * c.dV - Exc) < 1e-11 * abs(Exc)


def test_vloc_and_poisson(K):
    s, xpb, L = _sys(n=5, basis="gth-szv", grid_length=0.25)
    print(f"[vloc] g={s.g}", flush=True)
    g = s.g
    VG = torch.tensor(s.VG.ravel(), device=DEV)
    out = torch.empty(1, g, g, g, dtype=torch.complex128, device=DEV)
    K.vloc_G(torch.tensor(xpb[None], device=DEV), g, s.L, VG, out)
    ref = R.vloc_G(s, xpb)
    assert np.allclose(out[0].cpu().numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    K.fft3d(out, +1)
    ld = (s.N + 63) // 64 * 64
    vl = torch.zeros(1, ld, dtype=torch.float64, device=DEV)
    K.complex_real(out, ld, vl)
    assert np.allclose(vl[0, :s.N].cpu().numpy(), R.vloc_real(s, xpb), rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("g,n", [(27, 40), (49, 54), (65, 17)])
def test_structure_factor_kernels(K, g, n):
    """v_loc(G) and the local-potential force sums from per-atom 1-D phase tables
    (several atom chunks, several geometries) against the direct sums
        v_loc(G) = -VG sum_A exp(-i G.R_A),
        dE[A, c] = sum_G G_c VG Im(rho~(G) exp(i G.R_A))."""
    print(f"[structure factors] g={g} n={n}", flush=True)
    rng = np.random.default_rng(g + n)
    B, L = 3, 7.3
    xp = rng.uniform(0, L, (B, n, 3))
    k = np.fft.fftfreq(g, 1.0 / g)                    # 0..g/2, then -g/2..-1
    G1 = 2 * np.pi * k / L
    Gx, Gy, Gz = np.meshgrid(G1, G1, G1, indexing="ij")
    G2 = Gx ** 2 + Gy ** 2 + Gz ** 2
    VG = np.where(G2 > 0, 4 * np.pi / np.where(G2 > 0, G2, 1.0) / L ** 3, 0.0)
    out = torch.empty(B, g, g, g, dtype=torch.complex128, device=DEV)
    xt = torch.tensor(xp, device=DEV)
    K.vloc_G(xt, g, L, torch.tensor(VG.ravel(), device=DEV), out)
    rg = rng.standard_normal((B, g, g, g)) + 1j * rng.standard_normal((B, g, g, g))
    dE = torch.empty(B, n, 3, dtype=torch.float64, device=DEV)
    K.vloc_force(xt, g, L, torch.tensor(VG.ravel(), device=DEV),
                 torch.tensor(rg.reshape(B, -1), device=DEV), dE)
    for b in range(B):
        ph = np.exp(-1j * (np.multiply.outer(Gx, xp[b, :, 0]) + np.multiply.outer(Gy, xp[b, :, 1])
                           + np.multiply.outer(Gz, xp[b, :, 2])))      # (g, g, g, n)
        ref = -VG * ph.sum(-1)
        assert np.allclose(out[b].cpu().numpy(), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
        t = (VG[..., None] * rg[b][..., None] * np.conj(ph)).imag        # (g, g, g, n)
        refF = np.stack([(Gc[..., None] * t).sum((0, 1, 2)) for Gc in (Gx, Gy, Gz)], -1)
        assert np.allclose(dE[b].cpu().numpy(), refF, rtol=0, atol=1e-11 * np.abs(refF).max())


def test_ewald(K):
    from hqc_amd.pbc.potential import ewald_pp
    n, rs = 14, 1.31
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    xp = np.random.default_rng(5).uniform(0, L, (2, n, 3))
    E, F = ewald_pp(torch.tensor(xp, device=DEV), L, rs, forces=True)
    for b in range(2):
        Er, Fr = R.ewald_pp(xp[b], L, rs, forces=True)
        assert abs(E[b].item() - Er) < 1e-11
        assert np.allclose(F[b].cpu().numpy(), Fr, atol=1e-11)


def test_ewald_recip_matches_pair_kernel(K):
    """Structure-factor reciprocal sums equal the pair kernel's G sums summed over
    the pairs, and their gradient the pair gradients gathered per atom."""
    from hqc_amd.pbc.potential import _weights, kpoints
    n = 54
    rng = np.random.default_rng(8)
    frac = torch.tensor(rng.uniform(0, 1, (3, n, 3)), device=DEV)
    G = kpoints(3, 15, DEV)
    gk = _weights(G, 10.0).contiguous()
    E, grad = K.ewald_recip(frac, G, gk)
    i, j = torch.triu_indices(n, n, offset=1, device=DEV)
    rij = frac[:, i] - frac[:, j]
    rij = (rij - torch.round(rij)).contiguous()
    out = torch.empty(3, rij.shape[1], 4, dtype=torch.float64, device=DEV)
    # kappa large: the real-space terms vanish, g0 = 0 leaves the G sums alone
    K.ewald_pairs(rij, G, gk, 1e3, 0.0, out)
    Ep = out[..., 0].sum(-1)
    gp = torch.zeros(3, n, 3, dtype=torch.float64, device=DEV)
    gp.index_add_(1, i, out[..., 1:])
    gp.index_add_(1, j, -out[..., 1:])
    assert torch.allclose(E, Ep, rtol=1e-12, atol=1e-10)
    assert torch.allclose(grad, gp, rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("n", [2, 7, 8, 33, 107, 108])
def test_eigh_jacobi(K, n):
    """Jacobi eigensolver vs torch.linalg.eigh: eigenvalues, residual, orthogonality;
    also warm-started in an approximate eigenbasis (as the SCF uses it)."""
    torch.manual_seed(n)
    A = torch.randn(3, n, n, dtype=torch.float64, device=DEV)
    A = A + A.transpose(1, 2)
    A[0] = torch.diag(torch.arange(n, dtype=torch.float64, device=DEV) % 3)  # degenerate
    sw = torch.zeros(3, dtype=torch.int32, device=DEV)
    w, V = K.eigh_jacobi(A, sweeps=sw)
    wr = torch.linalg.eigvalsh(A)
    scale = A.abs().max().item()
    assert torch.allclose(w, wr, atol=1e-12 * scale)
    I = torch.eye(n, dtype=torch.float64, device=DEV)
    assert torch.allclose(V.transpose(1, 2) @ V, I.expand(3, n, n), atol=1e-12)
    assert torch.allclose(A @ V, V * w[:, None, :], atol=1e-11 * scale)
    # warm start: rotate by a nearby orthogonal matrix, few sweeps needed
    _, C = torch.linalg.eigh(A + 1e-3 * torch.randn_like(A).transpose(1, 2).add(A * 0))
    Aw = C.transpose(1, 2) @ A @ C
    w2, U = K.eigh_jacobi(Aw, sweeps=sw)
    assert torch.allclose(w2, wr, atol=1e-12 * scale)
    assert torch.allclose(A @ (C @ U), (C @ U) * w2[:, None, :], atol=1e-11 * scale)
    # the schedule computed on the device is the host table's, bit for bit
    w3, V3 = K.eigh_jacobi(A, pairs=K.jacobi_pairs(n, A.device))
    assert torch.equal(w3, w) and torch.equal(V3, V)


@pytest.mark.parametrize("gamma", [True, False])
def test_slater_matrices(K, gamma):
    """make_slater / make_pbc_gto against the oracle, single and batched walkers."""
    from hqc_amd.pbc import make_pbc_gto, make_slater
    n, rs, basis = 6, 1.3, "gth-dzvp"
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    rng = np.random.default_rng(7)
    xp = rng.uniform(0, L, (n, 3))
    xe = rng.uniform(0, L, (5, n, 3))
    nao = n * 5
    kpt = None if gamma else np.array([0.25, -0.1, 0.4])
    mo = rng.normal(size=(nao, nao)) if gamma else \
        rng.normal(size=(nao, nao)) + 1j * rng.normal(size=(nao, nao))
    print(f"[slater] gamma={gamma}", flush=True)
    sl = make_slater(n, L, rs, basis, gamma=gamma)
    args = (xp, xe, mo) if gamma else (xp, xe, mo, kpt)
    up, dn = sl(*args)
    assert up.shape == (5, n // 2, n // 2)
    for w in range(5):
        ru, rd = R.slater(n, L, rs, basis, xp, xe[w], mo, kpt=kpt)
        assert np.allclose(up[w].cpu().numpy(), ru, rtol=1e-12, atol=1e-13)
        assert np.allclose(dn[w].cpu().numpy(), rd, rtol=1e-12, atol=1e-13)
    # excited-state form with explicit state indices
    idx = np.array([0, 2, 4, 1, 3, 5])
    sl2 = make_slater(n, L, rs, basis, groundstate=False, gamma=gamma)
    up2, _ = sl2(xp, xe[0], mo, idx) if gamma else sl2(xp, xe[0], mo, kpt, idx)
    ru2, _ = R.slater(n, L, rs, basis, xp, xe[0], mo, kpt=kpt, state_idx=idx)
    assert np.allclose(up2.cpu().numpy(), ru2, rtol=1e-12, atol=1e-13)
    gto = make_pbc_gto(basis, L * rs, gamma=gamma)
    one = gto(xp * rs, xe[0, 0] * rs) if gamma else gto(xp * rs, xe[0, 0] * rs, kpt)
    ref = R.ao_points(R.AtomBasis(basis), L * rs, xp * rs, xe[0, :1] * rs, kpt=kpt)[0]
    assert one.shape == (nao,) and np.allclose(one.cpu().numpy(), ref, rtol=1e-12, atol=1e-14)


def test_potential_energy_protons_electrons(K):
    from hqc_amd.pbc import potential_energy
    rng = np.random.default_rng(9)
    n, rs = 10, 1.4
    L = (4 / 3 * np.pi * (n // 2)) ** (1 / 3)
    x = rng.uniform(-L, 2 * L, (3, n, 3))
    G = R.ewald_kpoints(15)
    vpp, vep, vee = potential_energy(x, 10.0, G, L, rs)
    for w in range(3):
        ref = R.potential_energy(x[w], 10.0, G, L, rs)
        for a, r in zip((vpp[w], vep[w], vee[w]), ref):
            assert abs(a.item() - r) < 1e-11 * max(1.0, abs(r))


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_geometry_index_lists(K, dt):
    """Density, orbital density and potential matrix on a subset of the batch read in
    place through bidx equal the same kernels on a gathered copy."""
    torch.manual_seed(5)
    B, nao_pad, g = 5, 48, 9
    N = g ** 3
    ld = (N + 63) // 64 * 64
    phi = torch.zeros(B, nao_pad, ld, dtype=torch.float64, device=DEV)
    phi[:, :, :N] = torch.randn(B, nao_pad, N, dtype=torch.float64, device=DEV)
    phi = phi.to(dt).contiguous()
    idx = torch.tensor([3, 0, 4], dtype=torch.int32, device=DEV)
    sub = phi[idx.long()].contiguous()
    A = torch.randn(3, nao_pad, nao_pad, dtype=torch.float64, device=DEV)
    D = (A + A.transpose(1, 2)).contiguous()
    r1 = torch.zeros(3, ld, dtype=torch.float64, device=DEV)
    r2 = torch.zeros_like(r1)
    K.ao_density(phi, D, r1, bidx=idx)
    K.ao_density(sub, D, r2)
    assert torch.equal(r1, r2)
    At = torch.randn(3, nao_pad, 32, dtype=torch.float64, device=DEV)
    K.orb_density(phi, At, r1, bidx=idx)
    K.orb_density(sub, At, r2)
    assert torch.equal(r1, r2)
    v = torch.zeros(3, ld, dtype=torch.float64, device=DEV)
    v[:, :N] = torch.randn(3, N, dtype=torch.float64, device=DEV)
    V1 = torch.empty(3, nao_pad, nao_pad, dtype=torch.float64, device=DEV)
    V2 = torch.empty_like(V1)
    prec = 64 if dt == torch.float64 else 32
    work = torch.empty(K.vmat_workspace(3, nao_pad, ld, prec) // 8 + 1, dtype=torch.float64,
                       device=DEV)
    K.ao_vmat(phi, v, 0.7, V1, work, bidx=idx)
    K.ao_vmat(sub, v, 0.7, V2, work)
    assert torch.equal(V1, V2)


def test_autotune_changes_split_not_results(K):
    """The autotuner at a small shape returns one setting per tunable kernel; the
    potential matrix under any tile split equals the default one to rounding (partial
    sums grouped differently), and the table is restored afterwards."""
    from hqc_amd import tune
    from hqc_amd.pbc import make_pes
    n = 16
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    eng = make_pes(n, L, 1.25, "gth-dzv", grid_length=0.3, dft=True,
                   precision="mixed").solver.engine
    rec = tune.autotune(eng, 3, reps=1)
    assert {"vh_blocks", "vmat_parts", "vmat_tpw", "q8_blocks", "q8_min_tiles",
            "fft_plane_threads"} <= set(rec["choice"])
    assert all(K.get_tuning(k) == 0 for k in K.TUNE_KEYS)
    c = eng.c
    torch.manual_seed(4)
    phi = torch.zeros(3, c.nao_pad, c.ld, dtype=torch.float64, device=DEV)
    phi[:, :c.nao, :c.N] = torch.rand(3, c.nao, c.N, dtype=torch.float64, device=DEV)
    v = torch.zeros(3, c.ld, dtype=torch.float64, device=DEV)
    v[:, :c.N] = torch.rand(3, c.N, dtype=torch.float64, device=DEV)
    V0 = eng.vmat(phi, v)
    for tpw, parts in ((1, 2), (3, 64), (8, 5)):
        tune.apply({"vmat_tpw": tpw, "vmat_parts": parts})
        V1 = eng.vmat(phi, v)
        torch.cuda.synchronize()
        assert (V1 - V0).abs().max().item() <= 1e-13 * V0.abs().max().item()
    tune.apply({})


@pytest.mark.parametrize("basis", ["gth-szv", "gth-dzv"])
def test_overlap_kinetic_s_only_matches_generic(K, basis):
    """The s-only lattice-sum kernel (one thread per atom pair) against the generic
    one: the same S, T and centre derivatives to rounding (the same sums; the compiler
    fuses some products into the accumulation there, one rounding fewer)."""
    from hqc_amd.pbc.engine import Cell
    n = 9
    L = (4 / 3 * np.pi * n) ** (1 / 3)
    cell = Cell(n, L, 1.3, basis, grid_length=0.3)
    xp = torch.tensor(np.random.default_rng(2).uniform(0, L * 1.3, (5, n, 3)), device=DEV)
    out = []
    for s_only in (False, True):
        S = torch.empty(5, cell.nao, cell.nao, dtype=torch.float64, device=DEV)
        T = torch.empty_like(S)
        dS = torch.empty(5, cell.nao, cell.nao, 3, dtype=torch.float64, device=DEV)
        dT = torch.empty_like(dS)
        K.overlap_kinetic(xp, cell.lattice, cell.alpha, cell.coef, cell.comp, S, T, dS, dT,
                          s_only=s_only)
        S2 = torch.empty_like(S)
        T2 = torch.empty_like(S)
        K.overlap_kinetic(xp, cell.lattice, cell.alpha, cell.coef, cell.comp, S2, T2,
                          s_only=s_only)
        out.append((S, T, dS, dT, S2, T2))
    torch.cuda.synchronize()
    for a, b in zip(*out):
        d = (a - b).abs().max().item()
        assert d <= 1e-15 * max(a.abs().max().item(), 1.0), d


MIXED_V_SHAPES = [
    # n, rs, basis, grid_length: n_ao_pad and grid spread around the workloads
    (8, 1.25, "gth-szv", 0.3),       # 16 AOs, 25^3
    (16, 1.25, "gth-dzv", 0.16),     # 32 AOs, 33^3
    (54, 1.25, "gth-dzv", 0.16),     # 112 AOs, 49^3 (the benchmark)
    (64, 1.5, "gth-dzv", 0.16),      # 128 AOs, 61^3 (the HF batch's largest)
    (25, 1.25, "gth-dzvp", 0.16),    # 125 -> 128 AOs with p functions, 41^3
    (128, 1.25, "gth-dzvp", 0.16),   # 640 AOs, 65^3: FP32 MFMA, not split-FP16
]


@pytest.mark.parametrize("shape", MIXED_V_SHAPES, ids=[f"{s[0]}H-{s[2]}" for s in MIXED_V_SHAPES])
def test_mixed_v_error_margin(K, shape):
    """The mixed path takes V(dv) in low precision while eps * max|dv| stays within its
    error budget (engine.py _mixed_vmat), so eps must bound the per-element error of
    the low-precision V relative to max|v|.  On the real AO values of each shape and
    both a smooth potential (v_H + v_xc of the superposed-atom density) and a random
    one, the measured max|V_low - V_64| / max|v| stays at least 3x below the engine's
    eps (MIXED_EPS_SPLIT16 for the FP16 split kernel, MIXED_EPS for FP32)."""
    from hqc_amd.pbc.engine import Cell, Engine
    n, rs, basis, gl = shape
    print(f"[mixed V margin] {n} H {basis} gl {gl}", flush=True)
    L = (4 / 3 * math.pi * n) ** (1 / 3)
    cell = Cell(n, L, rs, basis, grid_length=gl, device=DEV)
    eng = Engine(cell, dft=True, precision="mixed")
    xp = torch.tensor(np.random.default_rng(n).uniform(0, L, (1, n, 3)) * rs, device=DEV)
    st = eng.setup(xp)
    lp = "phi_h2" if "phi_h2" in st else "phi_lp"
    eps = eng.MIXED_EPS_SPLIT16 if lp == "phi_h2" else eng.MIXED_EPS
    # smooth: the potential of the sum of the AO densities; random: uniform in [-1, 1]
    rho = (st["phi"][:, :cell.nao, :] ** 2).sum(1)
    v_s, _, _ = eng.xc_potential(rho)
    gen = torch.Generator(device="cpu").manual_seed(n)
    v_r = torch.zeros(1, cell.ld, dtype=torch.float64)
    v_r[:, :cell.N] = torch.rand(1, cell.N, generator=gen, dtype=torch.float64) * 2 - 1
    for name, v in (("smooth", v_s), ("random", v_r.to(DEV))):
        vmax = v.abs().amax(1)
        V64 = eng.vmat(st["phi"], v)
        Vlo = eng.vmat(st[lp], v, vmax=vmax)
        ratio = ((Vlo - V64).abs().max() / vmax.max()).item()
        print(f"  {name}: {lp} max|dV|/max|v| = {ratio:.2e} (eps {eps:.0e})", flush=True)
        assert ratio * 3 <= eps, (name, ratio, eps)


@pytest.mark.parametrize("g", [1, 3, 5, 9, 25])
@pytest.mark.parametrize("dt", [torch.complex128, torch.complex64])
def test_pair_split(K, g, dt):
    """Half-spectrum split of transformed pair grids Z = FFT(a + i b) into the rows
    [Re A_a(H) | Im A_a(H)], [Re A_b(H) | Im A_b(H)] with A_a(G) = (Z(G) + conj Z(-G))/2,
    A_b(G) = (Z(G) - conj Z(-G))/(2i) over the centred half G-set H: every element
    of M written (NaN-filled beforehand), padding rows and the columns past N - 1
    zero, including grids whose zero tail is wider than the half-set (g = 1, 3)."""
    print(f"[pair split] g={g} {dt}", flush=True)
    torch.manual_seed(g)
    mp, nao, nao_pad = 3, 5, 16
    N = g ** 3
    ld = (N + 63) // 64 * 64
    z = torch.randn(mp, nao, N, dtype=dt, device=DEV)
    rt = torch.float64 if dt == torch.complex128 else torch.float32
    M = torch.full((2 * mp, nao_pad, ld), float("nan"), dtype=rt, device=DEV)
    K.pair_split(z, nao_pad, ld, M)
    nh, h = (N - 1) // 2, g // 2
    c = torch.arange(nh, device=DEV)
    s1, s2, s3 = c // (g * g) - h, (c // g) % g - h, c % g - h
    k = ((s1 % g) * g + s2 % g) * g + s3 % g
    kn = (((-s1) % g) * g + (-s2) % g) * g + (-s3) % g
    zd = z.to(torch.complex128)
    zp, zn = zd[..., k], zd[..., kn]
    ref = torch.zeros(2 * mp, nao_pad, ld, dtype=torch.float64, device=DEV)
    ref[0::2, :nao, :nh] = 0.5 * (zp.real + zn.real)
    ref[0::2, :nao, nh:2 * nh] = 0.5 * (zp.imag - zn.imag)
    ref[1::2, :nao, :nh] = 0.5 * (zp.imag + zn.imag)
    ref[1::2, :nao, nh:2 * nh] = 0.5 * (zn.real - zp.real)
    torch.cuda.synchronize()
    assert not torch.isnan(M).any()
    assert torch.equal(M, ref.to(rt))


@pytest.mark.parametrize("g", [1, 3, 13, 25, 49, 61])
@pytest.mark.parametrize("dt", [torch.complex128, torch.complex64])
def test_pair_pack_fft_equals_pack_then_fft(K, g, dt):
    """The pair densities formed on the transform's first pass (hqc_pair_pack_fft)
    give the grids of hqc_pair_pack followed by the 3-D FFT bit for bit, for rows of
    two geometries, lone orbitals (j = -1) and both transform signs; the product with
    the pair table is also checked against torch's FFT of the pair densities."""
    print(f"[pair pack fft] g={g} {dt}", flush=True)
    torch.manual_seed(100 + g)
    B, nao, nao_pad, nmo = 2, 5, 16, 6
    N = g ** 3
    ld = (N + 63) // 64 * 64
    phi = torch.zeros(B, nao_pad, ld, dtype=torch.float64, device=DEV)
    phi[:, :nao, :N] = torch.randn(B, nao, N, dtype=torch.float64, device=DEV)
    psi = torch.zeros(B, nmo, ld, dtype=torch.float64, device=DEV)
    psi[:, :, :N] = torch.randn(B, nmo, N, dtype=torch.float64, device=DEV)
    table = torch.tensor([[0, 0, 1], [1, 2, 5], [0, 3, -1], [1, 4, 0]], dtype=torch.int32)
    mp = table.shape[0]
    for sign in (-1, 1):
        z1 = torch.empty(mp, nao, N, dtype=dt, device=DEV)
        K.pair_pack(phi, psi, table, nao, g, z1)
        K.fft3d(z1.view(mp * nao, g, g, g), sign)
        z2 = torch.full((mp, nao, N), float("nan"), dtype=dt, device=DEV)
        K.pair_pack(phi, psi, table, nao, g, z2, fft_sign=sign)
        torch.cuda.synchronize()
        assert torch.equal(z1, z2)
    # the forward transform of phi_p (psi_i + i psi_j) by torch (rocFFT)
    ref = torch.empty(mp, nao, N, dtype=torch.complex128, device=DEV)
    for m, (b, i, j) in enumerate(table.tolist()):
        pj = psi[b, j, :N] if j >= 0 else torch.zeros_like(psi[b, i, :N])
        dens = phi[b, :nao, :N] * torch.complex(psi[b, i, :N], pj)
        ref[m] = torch.fft.fftn(dens.view(nao, g, g, g), dim=(1, 2, 3)).reshape(nao, N)
    z3 = torch.empty(mp, nao, N, dtype=dt, device=DEV)
    K.pair_pack(phi, psi, table, nao, g, z3, fft_sign=-1)
    tol = (1e-12 if dt == torch.complex128 else 2e-5) * max(g, 2) ** 1.5
    assert torch.allclose(z3.to(torch.complex128), ref, rtol=0, atol=tol * ref.abs().max().item())


def test_pair_pack_fft_padded_rows(K):
    """The fused pack + transform into the padded row layout of the exchange
    contraction: grids at z[m, p, :g^3] equal pack + FFT, the padding rows and
    columns are left as they were (zero)."""
    print("[pair pack fft padded]", flush=True)
    torch.manual_seed(5)
    g, B, nao, nao_pad, nmo = 9, 2, 5, 16, 4
    N = g ** 3
    ld = (N + 63) // 64 * 64
    ldz = (2 * N + 63) // 64 * 64
    phi = torch.zeros(B, nao_pad, ld, dtype=torch.float64, device=DEV)
    phi[:, :nao, :N] = torch.randn(B, nao, N, dtype=torch.float64, device=DEV)
    psi = torch.zeros(B, nmo, ld, dtype=torch.float64, device=DEV)
    psi[:, :, :N] = torch.randn(B, nmo, N, dtype=torch.float64, device=DEV)
    table = torch.tensor([[1, 0, 3], [0, 2, -1], [1, 1, 2]], dtype=torch.int32)
    mp = table.shape[0]
    for dt in (torch.complex128, torch.complex64):
        z1 = torch.empty(mp, nao, N, dtype=dt, device=DEV)
        K.pair_pack(phi, psi, table, nao, g, z1)
        K.fft3d(z1.view(mp * nao, g, g, g), -1)
        zp = torch.zeros(mp, nao_pad, ldz // 2, dtype=dt, device=DEV)
        K.pair_pack(phi, psi, table, nao, g, zp, fft_sign=-1)
        torch.cuda.synchronize()
        assert torch.equal(zp[:, :nao, :N], z1)
        assert (zp[:, nao:] == 0).all() and (zp[:, :, N:] == 0).all()


@pytest.mark.parametrize("n,basis", [(4, "gth-dzv"), (6, "gth-szv")])
def test_exchange_without_split_equals_split(K, n, basis):
    """K from the transformed pair grids read whole (weights VG on every G, no split)
    equals K from the half-spectrum split of the same grids (weights 2 VG on half the
    G-set): the terms odd in G cancel over the symmetric G-set.  FP64 to 1e-12 of
    max|K|, and the FP32 (complex64) forms to 2e-6."""
    from hqc_amd.pbc.engine import Cell, Engine
    print(f"[exchange without split] {n} H {basis}", flush=True)
    rs, gl = 1.3, 0.3
    L = (4 / 3 * math.pi * n) ** (1 / 3)
    eng = Engine(Cell(n, L, rs, basis, grid_length=gl, device=DEV), dft=False)
    xp = torch.tensor(np.random.default_rng(n).uniform(0, L, (2, n, 3)) * rs, device=DEV)
    st = eng.setup(xp)
    assert "VGz" in st
    w, mo, _ = eng._eigh_orth(st["Hcore"], st["v"])
    D, occ = eng.density_matrix(mo, w)
    st_split = dict(st)
    del st_split["VGz"]
    for low in (None, [True, True]):
        Kf = eng.exchange(st, 2, mo, occ, low)
        Ks = eng.exchange(st_split, 2, mo, occ, low)
        torch.cuda.synchronize()
        tol = 1e-12 if low is None else 2e-6
        assert (Kf - Ks).abs().max().item() <= tol * Ks.abs().max().item()


def test_swap_rows_matches_gather_scatter(K):
    """hqc_swap_rows against torch's t[ia] = t[ib] with ia = holes + movers and
    ib = movers + holes, the compaction step of the batched SCF, over tensors of
    16-, 4- and 1-byte aligned rows and more tensors than one launch takes."""
    gen = torch.Generator().manual_seed(3)
    B = 37
    shapes = [((B, 5, 7), torch.float64), ((B,), torch.float64), ((B, 3), torch.int32),
              ((B, 5), torch.int8), ((B, 2, 9), torch.complex128), ((B,), torch.bool),
              ((B, 11), torch.int16), ((B,), torch.long)]
    ts = []
    for i in range(70):
        sh, dt = shapes[i % len(shapes)]
        x = torch.randint(-100, 100, sh, generator=gen)
        ts.append(x.to(dt).to(DEV) if dt != torch.bool else (x > 0).to(DEV))
    keep = (torch.rand(B, generator=gen) > 0.45).tolist()
    na = sum(keep)
    holes = [i for i in range(na) if not keep[i]]
    movers = [i for i in range(na, B) if keep[i]]
    assert holes and len(holes) == len(movers)
    ref = [t.clone() for t in ts]
    ia = torch.tensor(holes + movers, device=DEV)
    ib = torch.tensor(movers + holes, device=DEV)
    for t in ref:
        t[ia] = t[ib]
    pairs = torch.tensor([holes, movers], dtype=torch.long, device=DEV)
    K.swap_rows(ts, pairs)
    torch.cuda.synchronize()
    for t, r in zip(ts, ref):
        assert torch.equal(t, r)
    K.swap_rows(ts, pairs)                         # a swap undone by the same swap
    K.swap_rows(ts, pairs[:, :0])                  # no pairs: nothing launched
    torch.cuda.synchronize()
    for t, r in zip(ts, ref):
        t2 = r.clone()
        t2[ia] = t2[ib]
        assert torch.equal(t, t2)


@pytest.mark.parametrize("g,n,basis,zx", [(25, 3, "gth-dzv", False), (49, 2, "gth-dzvp", True),
                                          (65, 2, "gth-dzvp", False), (33, 2, "gth-szv", True)])
def test_grid_forces_line_form(K, g, n, basis, zx):
    """The grid adjoint along grid lines (one workgroup per AO, lane = z index) and
    the pointwise form against the direct sum
        F[b,A,c] = dV sum_r sum_{p in A} Z_p(r) d phi_p/d r_c,  Z = 2 v X + Zx,
    grad phi from the kernels' own table rule, on random tables and operands
    (65 points per line: two lane chunks; p functions in gth-dzvp)."""
    from hqc_amd.basis import HBasis
    print(f"[grid forces] g={g} n={n} {basis} zx={zx}", flush=True)
    hb = HBasis(basis)
    alpha, coef, comp = hb.alpha, np.asarray(hb.coef), np.asarray(hb.comp)
    na, nloc = len(alpha), len(comp)
    nl, B, dV = 3, 2, 0.37
    N = g ** 3
    ld = (N + 255) // 256 * 256
    nao_pad = (n * nloc + 15) // 16 * 16
    rng = np.random.default_rng(g + n)
    P = rng.standard_normal((B, n, 3, g, na, nl))
    X = rng.standard_normal((B, nao_pad, ld))
    v = rng.standard_normal((B, ld))
    Zx = rng.standard_normal((B, nao_pad, ld)) if zx else None
    t = lambda a: None if a is None else torch.tensor(a, device=DEV)
    args = (t(P), t(alpha), t(coef), torch.tensor(comp, dtype=torch.int32, device=DEV),
            nao_pad, ld, t(X), t(v), t(Zx), dV)
    nblk = K.force_blocks(g)
    out = {}
    for pointwise in (0, 1):
        K.set_tuning("force_pointwise", pointwise)
        part = torch.full((B, n, nblk, 3), float("nan"), dtype=torch.float64, device=DEV)
        K.grid_forces(*args, part)
        out[pointwise] = part.sum(2).cpu().numpy()
    K.set_tuning("force_pointwise", 0)
    ref = np.zeros((B, n, 3))
    for b in range(B):
        for A in range(n):
            for o in range(nloc):
                Z = 2 * v[b, :N] * X[b, A * nloc + o, :N]
                if zx:
                    Z = Z + Zx[b, A * nloc + o, :N]
                Z = Z.reshape(g, g, g)
                for a in range(na):
                    w = coef[a, o]
                    if w == 0:
                        continue
                    val, der = [], []
                    for d in range(3):
                        pw = int(comp[o] == d)
                        tt = P[b, A, d, :, a]
                        val.append(tt[:, pw])
                        der.append(-2 * alpha[a] * tt[:, pw + 1] + (tt[:, 0] if pw else 0))
                    for c in range(3):
                        f = [der[d] if d == c else val[d] for d in range(3)]
                        ref[b, A, c] += w * np.einsum("i,j,k,ijk->", f[0], f[1], f[2], Z) * dV
    scale = np.abs(ref).max()
    for pointwise in (0, 1):
        assert np.isfinite(out[pointwise]).all()
        assert np.abs(out[pointwise] - ref).max() <= 1e-12 * scale, pointwise
