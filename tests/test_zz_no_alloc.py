"""The allocation-free potential-matrix entry point (hqc_ao_vmat_f64_noalloc)
against hqc_ao_vmat_f64 with a caller workspace, called through the C-ABI:
two calls in a row, the second on a smaller input (the scratch array then
holds the first call's partial tiles), plus the size limit."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _inputs(B, nao, nao_pad, N, ld, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    phi = torch.zeros(B, nao_pad, ld, dtype=torch.float64)
    phi[:, :nao, :N] = torch.rand(B, nao, N, generator=g, dtype=torch.float64) - 0.5
    v = torch.zeros(B, ld, dtype=torch.float64)
    v[:, :N] = torch.rand(B, N, generator=g, dtype=torch.float64) - 0.5
    return phi.cuda(), v.cuda()


def _with_workspace(phi, v, dV, bidx=None):
    from hqc_amd import _lib as K
    B, nao_pad, ld = v.shape[0], phi.shape[1], phi.shape[2]
    work = torch.empty(K.vmat_workspace(B, nao_pad, ld, 64) // 8 + 1, dtype=torch.float64,
                       device="cuda")
    V = torch.empty(B, nao_pad, nao_pad, dtype=torch.float64, device="cuda")
    K.ao_vmat(phi, v, dV, V, work, bidx)
    return V


def _noalloc(phi, v, dV, bidx=None):
    from hqc_amd import _lib as K
    B, nao_pad = v.shape[0], phi.shape[1]
    V = torch.full((B, nao_pad, nao_pad), float("nan"), dtype=torch.float64, device="cuda")
    K.ao_vmat_noalloc(phi, v, dV, V, bidx)
    return V


def test_noalloc_vmat_matches_workspace_form_twice():
    print("[noalloc] 54 H dzv shape, 3 then 2 geometries", flush=True)
    nao, nao_pad, N = 108, 112, 25 ** 3
    ld = (N + 63) // 64 * 64
    phi, v = _inputs(3, nao, nao_pad, N, ld, 1)
    dV = 0.0123
    ref1 = _with_workspace(phi, v, dV)
    got1 = _noalloc(phi, v, dV)
    torch.cuda.synchronize()
    assert torch.equal(got1, ref1)
    # second call: fewer geometries, a geometry index list, another shape of grid
    N2 = 17 ** 3
    ld2 = (N2 + 63) // 64 * 64
    phi2, v2 = _inputs(4, 20, 32, N2, ld2, 2)
    bidx = torch.tensor([3, 1], dtype=torch.int32, device="cuda")
    v2 = v2[:2].contiguous()
    ref2 = _with_workspace(phi2, v2, 0.5, bidx)
    got2 = _noalloc(phi2, v2, 0.5, bidx)
    torch.cuda.synchronize()
    assert torch.equal(got2, ref2)
    # and against a plain FP64 product
    p = phi2[bidx.long()]
    plain = 0.5 * ((p * v2[:, None, :]) @ p.transpose(-1, -2).contiguous())
    assert (got2 - plain).abs().max().item() < 1e-12 * plain.abs().max().item()


def test_noalloc_vmat_large_batch_and_limit():
    """A batch whose partial tiles exceed the array runs with fewer grid chunks
    per geometry (the same sum to rounding); a batch that cannot fit even one
    chunk per geometry is refused before any launch."""
    from hqc_amd import _lib as K
    print("[noalloc] 64 geometries of 112 AOs; limit", flush=True)
    nao, nao_pad, N = 100, 112, 13 ** 3
    ld = (N + 63) // 64 * 64
    phi, v = _inputs(64, nao, nao_pad, N, ld, 3)
    ref = _with_workspace(phi, v, 1.0)
    got = _noalloc(phi, v, 1.0)
    torch.cuda.synchronize()
    assert (got - ref).abs().max().item() <= 1e-13 * ref.abs().max().item()
    cap = K.lib().hqc_ao_vmat_scratch_bytes()
    ntiles = (nao_pad // 16) * (nao_pad // 16 + 1) // 2
    B_big = cap // (ntiles * 256 * 8) + 1
    v_big = torch.zeros(B_big, ld, dtype=torch.float64, device="cuda")
    V_big = torch.empty(B_big, nao_pad, nao_pad, dtype=torch.float64, device="cuda")
    bidx = torch.zeros(B_big, dtype=torch.int32, device="cuda")
    with pytest.raises(ValueError, match="scratch"):
        K.ao_vmat_noalloc(phi, v_big, 1.0, V_big, bidx)
