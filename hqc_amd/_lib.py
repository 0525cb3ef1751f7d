"""ctypes binding of libhqc_amd.so (the C-ABI in include/hqc_amd.h).

Every wrapper takes torch tensors that already live on the GPU, checks shape,
dtype, device and contiguity on the host, passes raw device pointers and the
current HIP stream, and raises on a non-zero status.  There is no CPU
fallback: if the library is missing or no GPU is present the calls raise.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhqc_amd.so")
_ERRORS = {1: "bad argument", 2: "HIP error", 3: "size not supported"}

_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double

_SIGS = {
    "hqc_version": [],
    "hqc_swap_rows": [_P, _P, _I, _P, _P, _I, _P],
    "hqc_set_tuning": [_I, _I],
    "hqc_get_tuning": [_I],
    "hqc_gto_tables": [_P, _I, _I, _I, _D, _I, _P, _I, _I, _P, _P],
    "hqc_ao_values_f64": [_P, _I, _I, _I, _I, _I, _P, _P, _I, _I, _I, _P, _P],
    "hqc_ao_values_f32": [_P, _I, _I, _I, _I, _I, _P, _P, _I, _I, _I, _P, _P],
    "hqc_ao_values_f64_amax": [_P, _I, _I, _I, _I, _I, _P, _P, _I, _I, _I, _P, _P, _P],
    "hqc_ao_density_f64": [_P, _P, _P, _I, _I, _I, _P, _P, _P],
    "hqc_ao_density_f32": [_P, _P, _P, _I, _I, _I, _P, _P],
    "hqc_orb_density_f64": [_P, _P, _P, _I, _I, _I, _I, _P, _P],
    "hqc_orb_density_f32": [_P, _P, _P, _I, _I, _I, _I, _P, _P],
    "hqc_ao_vmat_workspace": [_I, _I, _I, _I, ctypes.POINTER(ctypes.c_size_t)],
    "hqc_ao_vmat_f64": [_P, _P, _P, _I, _I, _I, _D, _P, _P, ctypes.c_size_t, _P],
    "hqc_ao_vmat_f32": [_P, _P, _P, _I, _I, _I, _D, _P, _P, ctypes.c_size_t, _P],
    "hqc_ao_vmat_scratch_bytes": [],
    "hqc_ao_vmat_f64_noalloc": [_P, _P, _P, _I, _I, _I, _D, _P, _P],
    "hqc_phi_split_f16": [_P, ctypes.c_long, _D, _P, _P],
    "hqc_orb_density_h2": [_P, _P, _P, _I, _I, _I, _I, _D, _P, _P],
    "hqc_ao_vmat_h2_workspace": [_I, _I, _I, ctypes.POINTER(ctypes.c_size_t)],
    "hqc_ao_vmat_h2": [_P, _P, _P, _P, _I, _I, _I, _D, _D, _P, _P, ctypes.c_size_t, _P],
    "hqc_phi_slice_i8": [_P, _I, _I, _I, _I, _P, _P, _P, _P],
    "hqc_phi_slice_i8_h2": [_P, _I, _I, _I, _I, _D, _P, _P, _P, _P],
    "hqc_mixed_v_prep": [_P, _P, _P, _I, _I, _P, _P, _P, _P],
    "hqc_orb_density_q8": [_P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _P, _P],
    "hqc_fft3d_c2c_f64": [_P, _I, _I, _I, _P],
    "hqc_fft3d_c2c_f32": [_P, _I, _I, _I, _P],
    "hqc_pair_pack_c64": [_P, _P, _P, _I, _I, _I, _I, _I, _I, _P, _P],
    "hqc_pair_split_c64": [_P, _I, _I, _I, _I, _I, _P, _P],
    "hqc_pair_pack": [_P, _P, _P, _I, _I, _I, _I, _I, _I, _P, _P],
    "hqc_pair_pack_fft": [_P, _P, _P, _I, _I, _I, _I, _I, _I, _P, _I, _I, _I, _P],
    "hqc_pair_pack_fft_c64": [_P, _P, _P, _I, _I, _I, _I, _I, _I, _P, _I, _I, _I, _P],
    "hqc_pair_split_f64": [_P, _I, _I, _I, _I, _I, _P, _P],
    "hqc_pair_split_f32": [_P, _I, _I, _I, _I, _I, _P, _P],
    "hqc_vloc_G": [_P, _I, _I, _I, _D, _P, _P, _P],
    "hqc_real_to_complex": [_P, _I, _I, _I, _D, _P, _P],
    "hqc_complex_real": [_P, _I, _I, _I, _P, _P],
    "hqc_poisson_mul": [_P, _P, _I, _I, _I, _P, _P],
    "hqc_xc_blocks": [_I, _I],
    "hqc_xc_eval": [_P, _P, _I, _I, _I, _I, _I, _P, _P, _P],
    "hqc_real_pair_to_complex": [_P, _I, _I, _I, _D, _P, _P],
    "hqc_grad_G": [_P, _I, _I, _D, _P, _D, _P, _P],
    "hqc_div_G": [_P, _I, _I, _D, _P, _D, _P, _P],
    "hqc_gga_blocks": [_I],
    "hqc_xc_gga_eval": [_P, _P, _P, _I, _I, _I, _P, _P, _P, _P],
    "hqc_pbe_points": [_P, _P, _I, _P, _P],
    "hqc_occupations": [_P, _I, _I, _I, _D, _I, _I, _I, _D, _D, _P, _P, _P],
    "hqc_eigh_lds_bytes": [_I],
    "hqc_eigh_jacobi": [_P, _I, _I, _P, _I, _D, _P, _P, _P, _P],
    "hqc_overlap_kinetic": [_P, _I, _I, _P, _I, _P, _I, _P, _P, _I, _P, _P, _P, _P, _P],
    "hqc_overlap_kinetic_s": [_P, _I, _I, _P, _I, _P, _I, _P, _I, _P, _P, _P, _P, _P],
    "hqc_ewald_pairs": [_P, _I, _I, _P, _P, _I, _D, _D, _P, _P],
    "hqc_ewald_recip": [_P, _I, _I, _P, _P, _I, _P, _P, _P, _P],
    "hqc_force_blocks": [_I],
    "hqc_gto_tables_k": [_P, _P, _I, _I, _I, _D, _I, _P, _I, _I, _P, _P],
    "hqc_ao_values_k": [_P, _I, _I, _I, _I, _I, _P, _P, _I, _I, _I, _P, _P],
    "hqc_overlap_kinetic_k": [_P, _P, _I, _I, _P, _I, _P, _I, _P, _P, _I, _P, _P, _P, _P, _P],
    "hqc_grid_forces_k": [_P, _I, _I, _I, _I, _I, _P, _P, _P, _I, _I, _I, _P, _P, _P, _D, _P,
                          _P],
    "hqc_ao_points": [_P, _P, _P, _I, _I, _I, _D, _I, _P, _I, _I, _P, _P, _I, _P, _P],
    "hqc_grid_forces": [_P, _I, _I, _I, _I, _I, _P, _P, _P, _I, _I, _I, _P, _P, _P, _D, _P, _P],
    "hqc_vloc_force": [_P, _I, _I, _I, _D, _P, _P, _P, _P],
    "hqc_scf_grid": [_D, _D],
    "hqc_scf_create": [_P, _P],
    "hqc_scf_run": [_P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "hqc_scf_forces": [_P, _P, _I, _D, _I, _P, _P, _P, _P, _P],
    "hqc_scf_destroy": [_P],
}


_RESTYPES = {"hqc_eigh_lds_bytes": ctypes.c_size_t, "hqc_ao_vmat_scratch_bytes": ctypes.c_size_t,
             "hqc_scf_destroy": None}


def lib():
    """Load (once) and return the ctypes handle; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m hqc_amd.build`")
        h = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(h, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = h
    return _lib


def symbols():
    """Names of every entry point bound here (all declared in include/hqc_amd.h)."""
    return sorted(_SIGS)


def _call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise RuntimeError(f"{name} failed: {_ERRORS.get(rc, rc)}")


def _ptr(t, dtype=None, name="tensor"):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


F64 = torch.float64
F32 = torch.float32
I32 = torch.int32


TUNE_KEYS = {"vmat_blocks": 0, "vh_blocks": 1, "vmat_parts": 2, "vmat_tpw": 3, "q8_blocks": 4,
             "q8_min_tiles": 5, "fft_plane_threads": 6, "vbig_blocks": 7,
             "fft_two_buffer": 8, "fft_ip_items": 9, "eigh_recompute": 10, "xc_blocks": 11,
             "force_pointwise": 12, "ao_pointwise": 13}


def set_tuning(name, value):
    """Launch parameter `name` (TUNE_KEYS) of the calling host thread; 0 = default."""
    _call("hqc_set_tuning", TUNE_KEYS[name], int(value))


def get_tuning(name):
    return lib().hqc_get_tuning(TUNE_KEYS[name])


def gto_tables(xp, g, L, tmax, alpha, nl, out):
    B, n, _ = xp.shape
    na = alpha.numel()
    assert out.shape == (B, n, 3, g, na, nl)
    _call("hqc_gto_tables", _ptr(xp, F64, "xp"), B, n, g, float(L), int(tmax),
          _ptr(alpha, F64), na, nl, _ptr(out, F64, "P"), _stream())


def ao_values(P, coef, comp, nao_pad, ld, out, amax=None):
    """AO values into out (B, nao_pad, ld), FP64 or FP32; amax (B,) FP64 (FP64 out
    only), if given, is zero-filled here and receives max |phi| per geometry."""
    B, n, _, g, na, nl = P.shape
    nloc = coef.shape[1]
    assert out.shape == (B, nao_pad, ld) and coef.shape[0] == na
    if amax is not None:
        assert out.dtype == F64 and amax.dtype == F64 and amax.shape == (B,)
        amax.zero_()
        _call("hqc_ao_values_f64_amax", _ptr(P, F64, "P"), B, n, g, na, nl, _ptr(coef, F64),
              _ptr(comp, I32), nloc, nao_pad, ld, _ptr(out, F64, "phi"),
              _ptr(amax, F64, "amax"), _stream())
        return
    name = "hqc_ao_values_f64" if out.dtype == F64 else "hqc_ao_values_f32"
    _call(name, _ptr(P, F64, "P"), B, n, g, na, nl, _ptr(coef, F64), _ptr(comp, I32), nloc,
          nao_pad, ld, _ptr(out, None, "phi"), _stream())


def gto_tables_k(xp, kpt, g, L, tmax, alpha, nl, out):
    B, n, _ = xp.shape
    na = alpha.numel()
    assert kpt.shape == (B, 3) and out.dtype == torch.complex128
    assert out.shape == (B, n, 3, g, na, nl)
    _call("hqc_gto_tables_k", _ptr(xp, F64, "xp"), _ptr(kpt, F64, "kpt"), B, n, g, float(L),
          int(tmax), _ptr(alpha, F64), na, nl, _ptr(out, None, "P"), _stream())


def ao_values_k(P, coef, comp, nao_pad, ld, out):
    """Stacked complex AO values: out (B, 2 nao_pad, ld) FP64, Re rows then Im rows."""
    B, n, _, g, na, nl = P.shape
    nloc = coef.shape[1]
    assert P.dtype == torch.complex128 and out.shape == (B, 2 * nao_pad, ld)
    _call("hqc_ao_values_k", _ptr(P, None, "P"), B, n, g, na, nl, _ptr(coef, F64),
          _ptr(comp, I32), nloc, nao_pad, ld, _ptr(out, F64, "phi"), _stream())


def ao_points(xp, xe, kpt, L, tmax, alpha, nl, coef, comp, out):
    """AO values at points: xp (B,n,3), xe (B,ne,3) Bohr, kpt (B,3) or None;
    out (B, ne, n*nloc) FP64 (complex128 with kpt)."""
    B, n, _ = xp.shape
    ne = xe.shape[1]
    nloc = coef.shape[1]
    assert xe.shape == (B, ne, 3) and out.shape == (B, ne, n * nloc)
    assert out.dtype == (F64 if kpt is None else torch.complex128)
    assert kpt is None or kpt.shape == (B, 3)
    _call("hqc_ao_points", _ptr(xp, F64, "xp"), _ptr(xe, F64, "xe"), _ptr(kpt, F64, "kpt"), B,
          n, ne, float(L), int(tmax), _ptr(alpha, F64), alpha.numel(), int(nl), _ptr(coef, F64),
          _ptr(comp, I32), nloc, _ptr(out, None, "out"), _stream())


def _bidx(phi, bidx, B):
    """Geometry index list (int32 on the device) of a subset of phi's batch, or None."""
    if bidx is None:
        assert phi.shape[0] == B
        return None
    assert bidx.dtype == I32 and bidx.shape == (B,) and bidx.is_cuda
    return bidx


def ao_density(phi, D, rho, X=None, bidx=None):
    """rho (B, ld) [, X (B, nao_pad, ld)] = D-form density of phi[bidx] (bidx None: all)."""
    nao_pad, ld = phi.shape[1], phi.shape[2]
    B = D.shape[0]
    assert D.shape == (B, nao_pad, nao_pad) and rho.shape == (B, ld)
    bi = _bidx(phi, bidx, B)
    if phi.dtype == F64:
        if X is not None:
            assert X.shape == (B, nao_pad, ld)
        _call("hqc_ao_density_f64", _ptr(phi, F64, "phi"), _ptr(bi, I32, "bidx"), _ptr(D, F64, "D"),
              B, nao_pad, ld, _ptr(rho, F64, "rho"), _ptr(X, F64, "X"), _stream())
    else:
        assert X is None
        _call("hqc_ao_density_f32", _ptr(phi, F32, "phi"), _ptr(bi, I32, "bidx"), _ptr(D, F64, "D"),
              B, nao_pad, ld, _ptr(rho, F64, "rho"), _stream())


def orb_density(phi, At, rho, bidx=None, phi_scale=None):
    """phi FP64, FP32, or int32 FP16 pairs of phi_split_f16 (then phi_scale is its scale)."""
    nao_pad, ld = phi.shape[1], phi.shape[2]
    B, m_pad = At.shape[0], At.shape[2]
    assert At.shape == (B, nao_pad, m_pad) and rho.shape == (B, ld)
    bi = _bidx(phi, bidx, B)
    if phi.dtype == I32:
        _call("hqc_orb_density_h2", _ptr(phi, I32, "phi"), _ptr(bi, I32, "bidx"),
              _ptr(At, F64, "At"), B, nao_pad, m_pad, ld, float(phi_scale),
              _ptr(rho, F64, "rho"), _stream())
        return
    name = "hqc_orb_density_f64" if phi.dtype == F64 else "hqc_orb_density_f32"
    _call(name, _ptr(phi, None, "phi"), _ptr(bi, I32, "bidx"), _ptr(At, F64, "At"), B, nao_pad,
          m_pad, ld, _ptr(rho, F64, "rho"), _stream())


def vmat_workspace(B, nao_pad, ld, prec):
    n = ctypes.c_size_t(0)
    _call("hqc_ao_vmat_workspace", B, nao_pad, ld, prec, ctypes.byref(n))
    return n.value


def ao_vmat(phi, v, dV, V, work, bidx=None):
    nao_pad, ld = phi.shape[1], phi.shape[2]
    B = v.shape[0]
    assert v.shape == (B, ld) and V.shape == (B, nao_pad, nao_pad)
    bi = _bidx(phi, bidx, B)
    name = "hqc_ao_vmat_f64" if phi.dtype == F64 else "hqc_ao_vmat_f32"
    _call(name, _ptr(phi, None, "phi"), _ptr(bi, I32, "bidx"), _ptr(v, F64, "v"), B, nao_pad, ld,
          float(dV), _ptr(V, F64, "V"), _ptr(work, None, "work"),
          work.numel() * work.element_size(), _stream())


def ao_vmat_noalloc(phi, v, dV, V, bidx=None):
    """hqc_ao_vmat_f64 without a workspace (the library's own device array holds the
    partial tiles): for callers that cannot allocate.  Raises on sizes beyond it."""
    nao_pad, ld = phi.shape[1], phi.shape[2]
    B = v.shape[0]
    assert v.shape == (B, ld) and V.shape == (B, nao_pad, nao_pad)
    bi = _bidx(phi, bidx, B)
    rc = lib().hqc_ao_vmat_f64_noalloc(_ptr(phi, F64, "phi"), _ptr(bi, I32, "bidx"),
                                       _ptr(v, F64, "v"), B, nao_pad, ld, float(dV),
                                       _ptr(V, F64, "V"), _stream())
    if rc == 3:
        raise ValueError(f"hqc_ao_vmat_f64_noalloc: {B} geometries of {nao_pad} AOs need more "
                         f"than the {lib().hqc_ao_vmat_scratch_bytes() >> 20} MiB scratch array")
    if rc != 0:
        raise RuntimeError(f"hqc_ao_vmat_f64_noalloc failed: {_ERRORS.get(rc, rc)}")


def phi_split_f16(phi, scale, out):
    """out (int32 view of FP16 pairs) <- (hi, lo) with scale*phi = hi + lo."""
    assert phi.dtype == F64 and out.dtype == I32 and out.shape == phi.shape
    assert phi.is_contiguous() and out.is_contiguous()
    _call("hqc_phi_split_f16", _ptr(phi, F64, "phi"), phi.numel(),
          float(scale), _ptr(out, I32, "out"), _stream())


Q8_MAX_NAO_PAD = 128


def phi_slice_i8_h2(phi, nslice, scale, q, pe, h2):
    """phi_slice_i8's digits and exponents and, from the same read of phi, h2 (an
    int32 view of the FP16 pairs phi_split_f16(phi, scale) gives)."""
    G, nao_pad, ld = phi.shape
    assert q.dtype == torch.int8 and q.shape == (G, ld * nslice * nao_pad)
    assert pe.dtype == torch.int16 and pe.shape == (G, ld)
    assert h2.dtype == I32 and h2.shape == phi.shape
    _call("hqc_phi_slice_i8_h2", _ptr(phi, F64, "phi"), G, nao_pad, ld, nslice, float(scale),
          _ptr(q, torch.int8, "q"), _ptr(pe, torch.int16, "pe"), _ptr(h2, I32, "h2"),
          _stream())


def phi_slice_i8(phi, nslice, q, pe, amax=None):
    """q (G, ld*nslice*nao_pad) int8 digits and pe (G, ld) int16 exponents of phi
    (G, nao_pad, ld) FP64, in the tile layout the int8 density kernel reads;
    amax (G,) FP64, if given, is zero-filled here and receives max |phi| per
    geometry from the same pass."""
    G, nao_pad, ld = phi.shape
    assert q.dtype == torch.int8 and q.shape == (G, ld * nslice * nao_pad)
    assert pe.dtype == torch.int16 and pe.shape == (G, ld)
    if amax is not None:
        assert amax.dtype == F64 and amax.shape == (G,) and amax.is_contiguous()
        amax.zero_()
    _call("hqc_phi_slice_i8", _ptr(phi, F64, "phi"), G, nao_pad, ld, nslice,
          _ptr(q, torch.int8, "q"), _ptr(pe, torch.int16, "pe"), _ptr(amax, F64, "amax"),
          _stream())


def orb_density_q8(q, pe, At, rho, nslice, bidx=None, nuse=None, top=5):
    """rho (B, ld) = sum_i (At[:, :, i]^T phi)^2 from the int8 digits of phi_slice_i8
    (exact integer accumulation; bidx selects the geometries of q and pe; nuse of the
    nslice stored digits are used, all by default; digit products up to level top,
    4 or 5, or 3 with nuse 3 or 4)."""
    B, nao_pad, m_pad = At.shape
    ld = rho.shape[1]
    assert rho.shape == (B, ld) and q.shape[1] == ld * nslice * nao_pad and pe.shape[1] == ld
    assert q.shape[0] == pe.shape[0]
    bi = _bidx(q, bidx, B)
    _call("hqc_orb_density_q8", _ptr(q, torch.int8, "q"), _ptr(pe, torch.int16, "pe"),
          _ptr(bi, I32, "bidx"), _ptr(At, F64, "At"), B, nao_pad, m_pad, ld, nslice,
          nslice if nuse is None else int(nuse), int(top), _ptr(rho, F64, "rho"), _stream())


def vmat_h2_workspace(B, nao_pad, ld):
    n = ctypes.c_size_t(0)
    _call("hqc_ao_vmat_h2_workspace", B, nao_pad, ld, ctypes.byref(n))
    return n.value


def mixed_v_prep(v, va, direct, vv, vmax, dvmax):
    """vmax = max_r |v|, dvmax = max_r |v - va| per row, vv = v where direct else
    v - va: the mixed V step's operands in one pass over v and va (B, ld)."""
    B, ld = v.shape
    assert va.shape == (B, ld) and vv.shape == (B, ld)
    assert direct.dtype == torch.bool and direct.shape == (B,) and direct.is_contiguous()
    assert vmax.shape == (B,) and dvmax.shape == (B,)
    _call("hqc_mixed_v_prep", _ptr(v, F64, "v"), _ptr(va, F64, "va"),
          _ptr(direct.view(torch.uint8), torch.uint8, "direct"), B, ld, _ptr(vv, F64, "vv"),
          _ptr(vmax, F64, "vmax"), _ptr(dvmax, F64, "dvmax"), _stream())


def ao_vmat_h2(ph, v, scale, dV, V, work, bidx=None, vmax=None):
    """V = dV phi diag(v) phi^T from the FP16 pairs of phi_split_f16 (split products);
    vmax (B,) = max_r |v| per entry if the caller has it (else computed)."""
    nao_pad, ld = ph.shape[1], ph.shape[2]
    B = v.shape[0]
    assert ph.dtype == I32 and v.shape == (B, ld) and V.shape == (B, nao_pad, nao_pad)
    assert vmax is None or (vmax.shape == (B,) and vmax.is_contiguous())
    bi = _bidx(ph, bidx, B)
    _call("hqc_ao_vmat_h2", _ptr(ph, I32, "ph"), _ptr(bi, I32, "bidx"), _ptr(v, F64, "v"),
          _ptr(vmax, F64, "vmax"), B, nao_pad, ld, float(scale), float(dV), _ptr(V, F64, "V"),
          _ptr(work, None, "work"), work.numel() * work.element_size(), _stream())


SWAP_MAX_TENSORS = 64


def swap_rows(tensors, pairs):
    """Swap rows pairs[0, k] and pairs[1, k] of every tensor along its leading axis in
    one launch per SWAP_MAX_TENSORS tensors (hqc_swap_rows).  pairs: a (2, npairs)
    int64 GPU tensor, the pairs disjoint; tensors: contiguous GPU tensors whose
    leading axis holds every index named in pairs."""
    assert pairs.is_cuda and pairs.dtype == torch.int64 and pairs.dim() == 2 \
        and pairs.shape[0] == 2 and pairs.is_contiguous()
    npairs = pairs.shape[1]
    if npairs == 0 or not tensors:
        return
    pa = pairs.data_ptr()
    for i in range(0, len(tensors), SWAP_MAX_TENSORS):
        part = tensors[i:i + SWAP_MAX_TENSORS]
        bases = (ctypes.c_void_p * len(part))()
        rows = (ctypes.c_long * len(part))()
        for j, t in enumerate(part):
            bases[j] = _ptr(t, None, "swap_rows tensor").value
            assert t.dim() and t.numel(), "swap_rows tensor must have rows"
            rows[j] = t.numel() // t.shape[0] * t.element_size()
        _call("hqc_swap_rows", bases, rows, len(part), ctypes.c_void_p(pa),
              ctypes.c_void_p(pa + 8 * npairs), npairs, _stream())


def fft3d(c, sign):
    """In-place 3-D FFT of a complex128 or complex64 tensor (B, g, g, g)."""
    B, g = c.shape[0], c.shape[1]
    assert c.dtype in (torch.complex128, torch.complex64) and c.shape[1:] == (g, g, g)
    name = "hqc_fft3d_c2c_f64" if c.dtype == torch.complex128 else "hqc_fft3d_c2c_f32"
    _call(name, _ptr(c, None, "c"), B, g, int(sign), _stream())


# hqc_pair_pack_fft covers grids whose y-z plane fits the LDS with two buffers (g <= 69
# in FP64, whatever schedule is tuned) and at most this
# many (pair, AO) grids per call
PAIR_FFT_MAX_G = 69
PAIR_FFT_MAX_GRIDS = 65535


def pair_pack(phi, psi, table, nao, g, z, fft_sign=None):
    """z (mp, nao, g^3) complex128 <- phi[b,p] (psi[b,i] + i psi[b,j]) per row (b, i, j)
    of the host int32 table (validated here, then copied to the device).  fft_sign
    (+1 or -1): z is also 3-D transformed with that sign, the pair densities formed
    on the transform's first pass (hqc_pair_pack_fft); z may then be any
    (mp, rows >= nao, stride >= g^3) complex tensor, grid (m, p) in z[m, p, :g^3]."""
    B, nao_pad, ld = phi.shape
    nmo = psi.shape[1]
    mp = table.shape[0]
    assert not table.is_cuda and table.dtype == torch.int32 and table.shape == (mp, 3)
    assert psi.shape == (B, nmo, ld)
    assert z.dtype in (torch.complex128, torch.complex64) and z.is_contiguous()
    if fft_sign is None:
        assert z.shape == (mp, nao, g ** 3)
    else:
        assert z.dim() == 3 and z.shape[0] == mp and z.shape[1] >= nao and z.shape[2] >= g ** 3
    if mp == 0:
        return
    if not (0 <= int(table[:, 0].min()) and int(table[:, 0].max()) < B
            and 0 <= int(table[:, 1].min()) and int(table[:, 1].max()) < nmo
            and -1 <= int(table[:, 2].min()) and int(table[:, 2].max()) < nmo):
        raise ValueError("pair table entry out of range")
    td = table.to(phi.device)
    name = "hqc_pair_pack" if z.dtype == torch.complex128 else "hqc_pair_pack_c64"
    if fft_sign is None:
        _call(name, _ptr(phi, F64, "phi"), _ptr(psi, F64, "psi"),
              _ptr(td, torch.int32, "table"), mp, nao, nao_pad, nmo, g, ld, _ptr(z, None, "z"),
              _stream())
    else:
        assert fft_sign in (1, -1)
        _call(name.replace("_pack", "_pack_fft"), _ptr(phi, F64, "phi"), _ptr(psi, F64, "psi"),
              _ptr(td, torch.int32, "table"), mp, nao, nao_pad, nmo, g, ld, _ptr(z, None, "z"),
              z.shape[1], z.shape[2], int(fft_sign), _stream())


def pair_split(z, nao_pad, ld, M):
    """M (2 mp, nao_pad, ld) <- real half-spectrum blocks of the transformed pairs."""
    mp, nao, N = z.shape
    g = round(N ** (1 / 3))
    assert g ** 3 == N and M.shape == (2 * mp, nao_pad, ld)
    if z.dtype == torch.complex64:
        assert M.dtype == F32
        name = "hqc_pair_split_c64"
    else:
        name = "hqc_pair_split_f64" if M.dtype == F64 else "hqc_pair_split_f32"
    _call(name, _ptr(z, None, "z"), mp, nao, nao_pad, g, ld, _ptr(M, None, "M"), _stream())


def _lengths(L, B):
    """(scalar L, per-geometry tensor or None) from a float or a (B,) tensor."""
    if torch.is_tensor(L):
        assert L.shape == (B,)
        return 0.0, L
    return float(L), None


def grad_G(rhoG, g, L, scale, out):
    """L: float, or (B,) FP64 tensor of per-geometry cell lengths."""
    B = rhoG.shape[0]
    assert rhoG.dtype == torch.complex128 and rhoG.shape == (B, g, g, g)
    assert out.dtype == torch.complex128 and out.shape == (B, 3, g, g, g)
    L0, Lb = _lengths(L, B)
    _call("hqc_grad_G", _ptr(rhoG, None, "rhoG"), B, g, L0 if Lb is None else 1.0,
          _ptr(Lb, F64, "Lb"), float(scale), _ptr(out, None, "out"), _stream())


def div_G(wG, g, L, scale, out):
    B = wG.shape[0]
    assert wG.dtype == torch.complex128 and wG.shape == (B, 3, g, g, g)
    assert out.dtype == torch.complex128 and out.shape == (B, g, g, g)
    L0, Lb = _lengths(L, B)
    _call("hqc_div_G", _ptr(wG, None, "wG"), B, g, L0 if Lb is None else 1.0,
          _ptr(Lb, F64, "Lb"), float(scale), _ptr(out, None, "out"), _stream())


def gga_blocks(g):
    return lib().hqc_gga_blocks(int(g))


def xc_gga_eval(rho, grad, vh, g, vout, w, part):
    B, ld = rho.shape
    assert grad.shape == (B, 3, g, g, g) and w.shape == (B, 3, g, g, g)
    assert vh.shape == (B, g, g, g) and vout.shape == (B, ld)
    assert part.shape == (B, gga_blocks(g), 2)
    _call("hqc_xc_gga_eval", _ptr(rho, F64, "rho"), _ptr(grad, None, "grad"),
          _ptr(vh, None, "vh"), B, g, ld, _ptr(vout, F64, "vout"), _ptr(w, None, "w"),
          _ptr(part, F64, "part"), _stream())


def pbe_points(rho, sigma, out):
    n = rho.numel()
    assert sigma.numel() == n and out.shape == (n, 3)
    _call("hqc_pbe_points", _ptr(rho, F64, "rho"), _ptr(sigma, F64, "sigma"), n,
          _ptr(out, F64, "out"), _stream())


def vloc_G(xp, g, L, VG, out):
    B, n, _ = xp.shape
    assert out.dtype == torch.complex128 and out.shape == (B, g, g, g)
    _call("hqc_vloc_G", _ptr(xp, F64, "xp"), B, n, g, float(L), _ptr(VG, F64, "VG"),
          _ptr(out, None, "out"), _stream())


def real_to_complex(x, g, s, out):
    B, ld = x.shape
    _call("hqc_real_to_complex", _ptr(x, F64, "x"), B, g, ld, float(s), _ptr(out, None, "out"),
          _stream())


def real_pair_to_complex(x, g, s, out):
    """out ((B+1)/2, g, g, g) complex128 <- s (x[2p] + i x[2p+1])."""
    B, ld = x.shape
    assert out.dtype == torch.complex128 and out.shape == ((B + 1) // 2, g, g, g)
    _call("hqc_real_pair_to_complex", _ptr(x, F64, "x"), B, g, ld, float(s),
          _ptr(out, None, "out"), _stream())


def complex_real(c, ld, out):
    B, g = c.shape[0], c.shape[1]
    assert out.shape == (B, ld)
    _call("hqc_complex_real", _ptr(c, None, "c"), B, g, ld, _ptr(out, F64, "out"), _stream())


def poisson_mul(c, VG, saved=None):
    """c *= VG in G space; VG (g^3,) shared or (B, g^3) per geometry."""
    B, g = c.shape[0], c.shape[1]
    batched = VG.dim() == 2
    assert VG.shape[-1] == g ** 3 and (not batched or VG.shape[0] == B)
    _call("hqc_poisson_mul", _ptr(c, None, "c"), _ptr(VG, F64, "VG"), B, g, int(batched),
          _ptr(saved, None, "saved"), _stream())


def xc_blocks(B, g):
    """Per-geometry partial sums of hqc_xc_eval for a batch of B (the part array's
    middle axis), on this host thread."""
    return lib().hqc_xc_blocks(int(B), int(g))


def xc_eval(rho, vh, g, xc_id, vout, part, vh_paired=False):
    """vh: complex (B, g, g, g), or ((B+1)/2, g, g, g) geometry pairs (Re, Im) with
    vh_paired (the output of a transform of real_pair_to_complex)."""
    B, ld = rho.shape
    assert vh.dtype == torch.complex128 and vh.shape[0] == ((B + 1) // 2 if vh_paired else B)
    assert part.shape == (B, xc_blocks(B, g), 2) and vout.shape == (B, ld)
    _call("hqc_xc_eval", _ptr(rho, F64, "rho"), _ptr(vh, None, "vh"), int(bool(vh_paired)), B,
          g, ld, int(xc_id), _ptr(vout, F64, "vout"), _ptr(part, F64, "part"), _stream())


def occupations(w, nelec, sigma, method, search, max_cycle, tol, mu_shift, occ, mu):
    B, nmo = w.shape
    _call("hqc_occupations", _ptr(w, F64, "w"), B, nmo, int(nelec), float(sigma), int(method),
          int(search), int(max_cycle), float(tol), float(mu_shift), _ptr(occ, F64, "occ"),
          _ptr(mu, F64, "mu"), _stream())


EIGH_JACOBI_MAXN = 110
_PAIRS = {}


def jacobi_pairs(n, device):
    """Round-robin (circle method) pairing schedule for n (padded to even) indices."""
    key = (n, str(device))
    if key not in _PAIRS:
        npad = n + (n & 1)
        arr = list(range(npad))
        rows = []
        for _ in range(npad - 1):
            rows.append([x for i in range(npad // 2) for x in (arr[i], arr[npad - 1 - i])])
            arr = [arr[0]] + [arr[-1]] + arr[1:-1]
        _PAIRS[key] = torch.tensor(rows, dtype=torch.int32, device=device).contiguous()
    return _PAIRS[key]


def eigh_jacobi(A, max_sweeps=30, tol=1e-14, sweeps=None, pairs=None):
    """Ascending eigenvalues and column eigenvectors of symmetric A (B, n, n), n <= 110.
    pairs: a schedule table (jacobi_pairs) or None for the same schedule computed on
    the device."""
    B, n, _ = A.shape
    A = A.contiguous()
    w = torch.empty(B, n, dtype=F64, device=A.device)
    V = torch.empty(B, n, n, dtype=F64, device=A.device)
    _call("hqc_eigh_jacobi", _ptr(A, F64, "A"), B, n, _ptr(pairs, I32, "pairs"),
          int(max_sweeps), float(tol), _ptr(w, F64, "w"), _ptr(V, F64, "V"),
          _ptr(sweeps, I32, "sweeps"), _stream())
    return w, V


def overlap_kinetic(xp, lattice, alpha, coef, comp, S, T, dS=None, dT=None, s_only=False):
    """S, T (and centre derivatives) by lattice sums; s_only: the basis has s functions
    only (the caller's knowledge of comp), which takes the specialised kernel."""
    B, n, _ = xp.shape
    na, nloc = coef.shape
    if s_only and na == 4 and nloc <= 2:
        _call("hqc_overlap_kinetic_s", _ptr(xp, F64, "xp"), B, n, _ptr(lattice, F64, "lattice"),
              lattice.shape[0], _ptr(alpha, F64), na, _ptr(coef, F64), nloc, _ptr(S, F64, "S"),
              _ptr(T, F64, "T"), _ptr(dS, F64, "dS"), _ptr(dT, F64, "dT"), _stream())
        return
    _call("hqc_overlap_kinetic", _ptr(xp, F64, "xp"), B, n, _ptr(lattice, F64, "lattice"),
          lattice.shape[0], _ptr(alpha, F64), na, _ptr(coef, F64), _ptr(comp, I32), nloc,
          _ptr(S, F64, "S"), _ptr(T, F64, "T"), _ptr(dS, F64, "dS"), _ptr(dT, F64, "dT"),
          _stream())


def overlap_kinetic_k(xp, kpt, lattice, alpha, coef, comp, S, T, dS=None, dT=None):
    B, n, _ = xp.shape
    na, nloc = coef.shape
    for t in (S, T, dS, dT):
        assert t is None or t.dtype == torch.complex128
    assert kpt.shape == (B, 3)
    _call("hqc_overlap_kinetic_k", _ptr(xp, F64, "xp"), _ptr(kpt, F64, "kpt"), B, n,
          _ptr(lattice, F64, "lattice"), lattice.shape[0], _ptr(alpha, F64), na, _ptr(coef, F64),
          _ptr(comp, I32), nloc, _ptr(S, None, "S"), _ptr(T, None, "T"), _ptr(dS, None, "dS"),
          _ptr(dT, None, "dT"), _stream())


def ewald_pairs(rij, Gv, gk, kappa, g0, out):
    B, npair, _ = rij.shape
    _call("hqc_ewald_pairs", _ptr(rij, F64, "rij"), B, npair, _ptr(Gv, F64, "Gv"),
          _ptr(gk, F64, "gk"), Gv.shape[0], float(kappa), float(g0), _ptr(out, F64, "out"),
          _stream())


def ewald_recip(frac, Gv, gk):
    """(E (B), grad (B, n, 3)): the reciprocal pair sums of ewald_pairs summed over
    pairs, and their gradient in fractional coordinates, from structure factors."""
    B, n, _ = frac.shape
    nG = Gv.shape[0]
    E = torch.empty(B, dtype=F64, device=frac.device)
    grad = torch.empty(B, n, 3, dtype=F64, device=frac.device)
    work = torch.empty(B, nG, 2, dtype=F64, device=frac.device)
    _call("hqc_ewald_recip", _ptr(frac, F64, "frac"), B, n, _ptr(Gv, F64, "Gv"),
          _ptr(gk, F64, "gk"), nG, _ptr(E, F64, "E"), _ptr(grad, F64, "grad"),
          _ptr(work, F64, "work"), _stream())
    return E, grad


def force_blocks(g):
    return lib().hqc_force_blocks(g)


def grid_forces(P, alpha, coef, comp, nao_pad, ld, X, v, Zx, dV, part):
    B, n, _, g, na, nl = P.shape
    nloc = coef.shape[1]
    _call("hqc_grid_forces", _ptr(P, F64, "P"), B, n, g, na, nl, _ptr(alpha, F64),
          _ptr(coef, F64), _ptr(comp, I32), nloc, nao_pad, ld, _ptr(X, F64, "X"),
          _ptr(v, F64, "v"), _ptr(Zx, F64, "Zx"), float(dV), _ptr(part, F64, "part"), _stream())


def grid_forces_k(P, alpha, coef, comp, nao_pad, ld, X, v, Zx, dV, part):
    B, n, _, g, na, nl = P.shape
    nloc = coef.shape[1]
    assert P.dtype == torch.complex128 and X.shape == (B, 2 * nao_pad, ld)
    _call("hqc_grid_forces_k", _ptr(P, None, "P"), B, n, g, na, nl, _ptr(alpha, F64),
          _ptr(coef, F64), _ptr(comp, I32), nloc, nao_pad, ld, _ptr(X, F64, "X"),
          _ptr(v, F64, "v"), _ptr(Zx, F64, "Zx"), float(dV), _ptr(part, F64, "part"), _stream())


def vloc_force(xp, g, L, VG, rhoG, dE):
    B, n, _ = xp.shape
    _call("hqc_vloc_force", _ptr(xp, F64, "xp"), B, n, g, float(L), _ptr(VG, F64, "VG"),
          _ptr(rhoG, None, "rhoG"), _ptr(dE, F64, "dE"), _stream())


# --------------------------------------------------------------------------- SCF driver
class ScfCell(ctypes.Structure):
    """hqc_scf_cell of include/hqc_amd.h."""
    _fields_ = [("n", _I), ("L", _D), ("g", _I), ("na", _I), ("nloc", _I), ("nl", _I),
                ("alpha", ctypes.POINTER(_D)), ("coef", ctypes.POINTER(_D)),
                ("comp", ctypes.POINTER(_I)), ("rcut", _D), ("xc_id", _I), ("smearing", _I),
                ("smearing_method", _I), ("sigma", _D), ("search", _I), ("search_cycle", _I),
                ("search_tol", _D), ("tol", _D), ("max_cycle", _I), ("diis_space", _I),
                ("diis_start_cycle", _I), ("diis_damp", _D)]


def scf_grid(L, grid_length):
    """The reference's odd grid for a cell of L Bohr (hqc_scf_grid, host only)."""
    return lib().hqc_scf_grid(float(L), float(grid_length))


class ScfDriver:
    """The C SCF driver (hqc_scf_create / run / destroy) through ctypes: the binding
    a host without Python makes.  Basis arrays alpha (na), coef (na, nloc), comp
    (nloc) as hqc_amd.basis gives them; L in Bohr."""

    def __init__(self, n, L, g, alpha, coef, comp, nl, rcut=24.0, xc_id=1, smearing=False,
                 smearing_method="fermi", sigma=0.0, search="bisect", search_cycle=100,
                 search_tol=1e-7, tol=1e-7, max_cycle=100, diis_space=8, diis_start_cycle=1,
                 diis_damp=0.0):
        import numpy as np
        a = np.ascontiguousarray(alpha, dtype=np.float64)
        c = np.ascontiguousarray(coef, dtype=np.float64)
        p = np.ascontiguousarray(comp, dtype=np.int32)
        self.n, self.nao = int(n), int(n) * c.shape[1]
        cell = ScfCell(int(n), float(L), int(g), a.shape[0], c.shape[1], int(nl),
                       a.ctypes.data_as(ctypes.POINTER(_D)), c.ctypes.data_as(ctypes.POINTER(_D)),
                       p.ctypes.data_as(ctypes.POINTER(_I)), float(rcut), int(xc_id),
                       int(bool(smearing)), 0 if smearing_method == "fermi" else 1, float(sigma),
                       0 if search == "bisect" else 1, int(search_cycle), float(search_tol),
                       float(tol), int(max_cycle), int(diis_space), int(diis_start_cycle),
                       float(diis_damp))
        self._ctx = ctypes.c_void_p()
        self.max_cycle = int(max_cycle)
        _call("hqc_scf_create", ctypes.byref(cell), ctypes.byref(self._ctx))

    def run(self, xp):
        """xp (B, n, 3) Bohr on the GPU -> dict of make_solver's outputs (mo, dm,
        bands [Ry], E, Ki, Vep, Vee [Ry], Se, converged) plus n_iter."""
        B = xp.shape[0]
        dev = xp.device
        f64 = torch.float64
        out = {k: torch.empty(B, self.nao, self.nao, dtype=f64, device=dev) for k in ("mo", "dm")}
        out["bands"] = torch.empty(B, self.nao, dtype=f64, device=dev)
        for k in ("E", "Ki", "Vep", "Vee", "Se"):
            out[k] = torch.empty(B, dtype=f64, device=dev)
        for k in ("converged", "n_iter"):
            out[k] = torch.empty(B, dtype=torch.int32, device=dev)
        _call("hqc_scf_run", self._ctx, _ptr(xp.contiguous(), F64, "xp"), B,
              *[_ptr(out[k]) for k in ("mo", "dm", "bands", "E", "Ki", "Vep", "Vee", "Se",
                                       "converged", "n_iter")], _stream())
        out["converged"] = out["converged"].bool()
        return out

    def forces(self, xp, kappa=10.0, Gmax=15):
        """xp (B, n, 3) Bohr on the GPU -> (E_tot (B) Ry, F (B, n, 3) Ha/Bohr, converged,
        n_iter): make_pes(...).energy_and_forces from the C driver."""
        B, n = xp.shape[0], xp.shape[1]
        dev = xp.device
        E = torch.empty(B, dtype=F64, device=dev)
        F = torch.empty(B, n, 3, dtype=F64, device=dev)
        conv = torch.empty(B, dtype=torch.int32, device=dev)
        nit = torch.empty(B, dtype=torch.int32, device=dev)
        _call("hqc_scf_forces", self._ctx, _ptr(xp.contiguous(), F64, "xp"), B, float(kappa),
              int(Gmax), _ptr(E), _ptr(F), _ptr(conv), _ptr(nit), _stream())
        return E, F, conv.bool(), nit

    def close(self):
        if self._ctx:
            lib().hqc_scf_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
