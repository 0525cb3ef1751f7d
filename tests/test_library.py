"""CPU checks of the C-ABI boundary: the library loads, exports every symbol
that include/hqc_amd.h declares, and the ctypes binding covers the same set.
No compute call is made (there is no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "hqc_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|size_t|void)\s+(hqc_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def libpath():
    from hqc_amd import _lib, build
    if not os.path.isfile(_lib.LIB_PATH):
        build.build()
    return _lib.LIB_PATH


def test_library_exports_every_declared_symbol(libpath):
    h = ctypes.CDLL(libpath)
    missing = [s for s in _declared() if not hasattr(h, s)]
    assert not missing, missing


def test_binding_matches_header():
    from hqc_amd import _lib
    assert _lib.symbols() == _declared()


def test_version_and_size_queries(libpath):
    from hqc_amd import _lib
    assert _lib.lib().hqc_version() >= 100
    # pure host-side queries (no device work)
    assert _lib.xc_blocks(1, 49) == 256 and _lib.xc_blocks(512, 49) == 16 and _lib.force_blocks(49) > 0
    n = _lib.vmat_workspace(2, 112, 117696, 64)
    assert n > 0 and n % (256 * 8) == 0
    with pytest.raises(RuntimeError):
        _lib.vmat_workspace(2, 100, 117696, 64)     # nao_pad must be a multiple of 16


def test_no_cpu_fallback_without_gpu():
    """Calling a compute entry point with CPU tensors raises; nothing silently
    falls back to a host implementation."""
    import torch
    from hqc_amd import _lib
    x = torch.zeros(1, 5, 5, 5, dtype=torch.complex128)
    with pytest.raises(ValueError):
        _lib.fft3d(x, -1)


def test_batch_buckets_by_atom_count_and_grid():
    """Systems are grouped by (n, odd grid); lengths on the same grid share a batch."""
    from hqc_amd.pbc.batch import BatchPES, uniform_gas_length
    pes = BatchPES("gth-dzv", grid_length=0.3)
    ns = [3, 3, 2, 3, 3]
    rs = [1.30, 1.35, 1.30, 1.50, 1.32]
    Lb = [uniform_gas_length(n) * r for n, r in zip(ns, rs)]
    b = pes.buckets(ns, Lb)
    assert b == {(3, 11): [0, 1, 4], (2, 9): [2], (3, 13): [3]}


def test_tuning_table_is_per_thread_and_checked(libpath):
    """The launch-parameter table (host only): values are kept per host thread,
    bad keys and out-of-range values are refused, 0 restores the default."""
    import threading
    from hqc_amd import _lib
    _lib.set_tuning("vmat_tpw", 6)
    seen = []
    t = threading.Thread(target=lambda: seen.append(_lib.get_tuning("vmat_tpw")))
    t.start()
    t.join()
    assert _lib.get_tuning("vmat_tpw") == 6 and seen == [0]
    with pytest.raises(RuntimeError):
        _lib.set_tuning("vmat_tpw", 9)
    with pytest.raises(RuntimeError):
        _lib.set_tuning("fft_plane_threads", 384)
    with pytest.raises(RuntimeError):
        _lib.set_tuning("fft_two_buffer", 2)
    assert _lib.lib().hqc_set_tuning(99, 1) == 1 and _lib.lib().hqc_get_tuning(99) == -1
    _lib.set_tuning("vmat_tpw", 0)
    assert _lib.get_tuning("vmat_tpw") == 0


def test_scf_grid_rounds_like_the_reference(libpath):
    """hqc_scf_grid is solver.py:159's odd grid, Python's round (ties to even)."""
    from hqc_amd import _lib
    for L, gl in [(7.7, 0.16), (4.03 * 1.25, 0.16), (10.0, 0.3), (0.32 * 5, 0.32),
                  (0.32 * 7, 0.32), (9.6, 0.12)]:
        assert _lib.scf_grid(L, gl) == round(L / gl / 2) * 2 + 1, (L, gl)


def test_scf_driver_rejects_what_it_does_not_cover(libpath):
    """Host-side argument checks of hqc_scf_create (no device call is reached):
    Hartree-Fock (xc_id 0) and n_ao > 110 are refused with the library's codes."""
    from hqc_amd import _lib
    from hqc_amd.basis import HBasis
    b = HBasis("gth-dzv")
    with pytest.raises(RuntimeError, match="bad argument"):
        _lib.ScfDriver(8, 7.7, 25, b.alpha, b.coef, b.comp, 1, xc_id=0)
    with pytest.raises(RuntimeError, match="size not supported"):
        _lib.ScfDriver(64, 12.0, 25, b.alpha, b.coef, b.comp, 1, xc_id=1)
