"""Autotuner for the launch parameters of the dominant kernels.

For an engine's workload shape (geometries per launch, AOs, grid) it times each
candidate setting of the library's tunable launch parameters (hqc_set_tuning,
include/hqc_amd.h) on synthetic operands of that shape with HIP events on the
stream the kernels run on, and keeps the fastest:

* potential matrix (split-FP16 products for n_ao <= 128, else FP32/FP64 MFMA):
  workgroups aimed at per launch, grid chunks per geometry at most, upper-triangle
  tiles per wave;
* int8-digit orbital density: workgroups aimed at, point tiles per wave at least;
* 3-D FFT: threads per y-z plane, in-place or two-buffer LDS passes (the same
  outputs bit for bit either way).

The choice changes how work is split across workgroups, not what is computed
(the potential matrix's partial sums are then added in another grouping).  It is
recorded as JSON ({"shape", "device", "choice", "timings"}) and applied with
apply(choice) on the calling host thread, or passed to an engine as
Engine(..., tuning=choice).

    python -m hqc_amd.tune --n 54 --rs 1.25 --basis gth-dzv --grid-length 0.16 --batch 512
"""
import argparse
import itertools
import json
import math
import sys

import torch

from . import _lib as K

DEFAULTS = {k: 0 for k in K.TUNE_KEYS}


def apply(choice):
    """Set every tunable parameter on this host thread (missing keys: default)."""
    for name in K.TUNE_KEYS:
        K.set_tuning(name, int((choice or {}).get(name, 0)))


def _time(fn, reps):
    s = torch.cuda.current_stream()
    fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _sweep(grid, fn, reps):
    """{setting: ms} over the cartesian product of grid {name: [values]}."""
    names = list(grid)
    out = {}
    for vals in itertools.product(*(grid[n] for n in names)):
        setting = dict(zip(names, vals))
        apply(setting)
        out[json.dumps(setting, sort_keys=True)] = _time(fn, reps)
    apply({})
    return out


def autotune(engine, B, reps=5, seed=0):
    """Tune the kernels an engine's SCF runs, at batch B; returns the record."""
    c = engine.c
    dev = c.device
    gen = torch.Generator(device="cpu").manual_seed(seed)
    rec = {"shape": {"B": B, "nao": c.nao, "nao_pad": c.nao_pad, "g": c.g, "ld": c.ld,
                     "precision": engine.precision},
           "device": torch.cuda.get_device_name(dev), "timings": {}, "choice": {}}
    phi = torch.zeros(B, c.nao_pad, c.ld, dtype=torch.float64)
    phi[:, :c.nao, :c.N] = torch.rand(B, c.nao, c.N, generator=gen, dtype=torch.float64)
    phi = phi.to(dev)
    v = torch.zeros(B, c.ld, dtype=torch.float64, device=dev)
    v[:, :c.N] = torch.rand(B, c.N, generator=gen, dtype=torch.float64).to(dev) - 0.5
    V = torch.empty(B, c.nao_pad, c.nao_pad, dtype=torch.float64, device=dev)
    small = c.nao_pad <= 128
    mixed = engine.precision == "mixed"
    # potential matrix in the precision the path's repeated V runs in
    if mixed and small:
        ph = torch.empty(B, c.nao_pad, c.ld, dtype=torch.int32, device=dev)
        scale = 2.0 ** math.floor(math.log2(2.0 ** 14))
        K.phi_split_f16(phi, scale, ph)
        vmax = v.abs().amax(1).contiguous()

        def run_v():
            work = engine._workspace(B, 16)
            K.ao_vmat_h2(ph, v, scale, 1.0, V, work, None, vmax)
        grid = {"vh_blocks": [1024, 2048, 4096], "vmat_parts": [16, 32, 64],
                "vmat_tpw": [0, 4, 5, 6, 8]}
        vname = "vmat_h2"
    else:
        pv = phi if not mixed else phi.float()
        prec = 64 if pv.dtype == torch.float64 else 32

        def run_v():
            K.ao_vmat(pv, v, 1.0, V, engine._workspace(B, prec))
        grid = ({"vmat_blocks": [1024, 2048, 4096], "vmat_parts": [16, 32, 64],
                 "vmat_tpw": [0, 4, 5, 6, 8]} if small else
                {"vbig_blocks": [512, 1024, 2048], "vmat_parts": [16, 32, 64]})
        vname = "vmat_f64" if prec == 64 else "vmat_f32"
    t = _sweep(grid, run_v, reps)
    rec["timings"][vname] = t
    best = json.loads(min(t, key=t.get))
    rec["choice"].update(best)
    # int8-digit density (the mixed path's accurate density, n_ao <= 128)
    if mixed and small and engine.MIXED_Q8:
        ns = engine.Q8_SLICES
        q = torch.empty(B, c.ld * ns * c.nao_pad, dtype=torch.int8, device=dev)
        pe = torch.empty(B, c.ld, dtype=torch.int16, device=dev)
        K.phi_slice_i8(phi, ns, q, pe)
        # orbitals above the occupation cut: about n/2 + 8 at 10^4 K (62 of 108 at 54 H)
        m_pad = (c.n // 2 + 8 + 15) // 16 * 16
        At = torch.zeros(B, c.nao_pad, m_pad, dtype=torch.float64, device=dev)
        At[:, :c.nao] = torch.rand(B, c.nao, m_pad, generator=gen,
                                   dtype=torch.float64).to(dev) - 0.5
        rho = torch.empty(B, c.ld, dtype=torch.float64, device=dev)
        t = _sweep({"q8_blocks": [1024, 2048, 4096], "q8_min_tiles": [4, 8, 16]},
                   lambda: K.orb_density_q8(q, pe, At, rho, ns, top=4), reps)
        rec["timings"]["orb_density_q8"] = t
        rec["choice"].update(json.loads(min(t, key=t.get)))
        del q, pe
    # FFTs of the Poisson solve (two geometries per complex grid)
    cb = torch.randn((B + 1) // 2, c.g, c.g, c.g, dtype=torch.complex128, generator=None,
                     device=dev)
    t = _sweep({"fft_plane_threads": [256, 512, 1024], "fft_two_buffer": [0, 1]},
               lambda: K.fft3d(cb, -1), reps)
    rec["timings"]["fft3d"] = t
    rec["choice"].update(json.loads(min(t, key=t.get)))
    return rec


def main(argv=None):
    from .pbc.engine import Cell, Engine
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=54)
    ap.add_argument("--rs", type=float, default=1.25)
    ap.add_argument("--basis", default="gth-dzv")
    ap.add_argument("--grid-length", type=float, default=0.16)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=None, help="write the record here (JSON)")
    a = ap.parse_args(argv)
    L = (4 / 3 * math.pi * a.n) ** (1 / 3)
    cell = Cell(a.n, L, a.rs, a.basis, grid_length=a.grid_length)
    rec = autotune(Engine(cell, dft=True, precision=a.precision), a.batch, a.reps)
    text = json.dumps(rec, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text + "\n")
    print(json.dumps({"shape": rec["shape"], "choice": rec["choice"]}))
    return rec


if __name__ == "__main__":
    main(sys.argv[1:])
