"""Throughput benchmark: energy+force geometries/sec for 54 H / gth-dzv.

Workload (one "step"): a batch of B independent 54-proton geometries
(rs = 1.25 Bohr, L = (4 pi 54/3)^(1/3) rs = 7.6162 Bohr, protons uniform in
the cell from numpy default_rng(seed), seeds 0..B-1), gth-dzv basis
(108 AOs), LDA (Slater + VWN), Fermi-Dirac smearing at T = 10^4 K
(sigma = 1/(2 beta), beta = 157888.088922572/T), the reference's odd grid for
grid_length 0.16 (49^3), SCF tol 1e-8 Ha, bisection mu search (make_pes
defaults otherwise); energy (E + V_pp) plus analytic forces for every
geometry.  Inputs are already on the GPU when the timed region starts.

    python bench.py [--steps K] [--warmup W] [--batch B] [--path fp64|mixed]

Prints ONE JSON line.  Only one GPU is supported (--gpus N>1 exits).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("energy+force geometries/sec, 54 H gth-dzv: mixed vs FP64 at "
          "ΔE≤1µHa/atom")
N_ATOMS, RS, BASIS = 54, 1.25, "gth-dzv"
T_K = 1.0e4
SIGMA = 1.0 / (157888.088922572 / T_K) / 2.0
GRID_LENGTH = 0.16
TOL = 1e-8
MI355X_PEAK = {"f64": 78.6, "f32": 157.3}   # dense MFMA TFLOP/s


def workload_xp(B, seed0=0):
    L = (4 / 3 * math.pi * N_ATOMS) ** (1 / 3)
    return np.stack([np.random.default_rng(seed0 + b).uniform(0, L, (N_ATOMS, 3))
                     for b in range(B)]), L


def make(precision):
    from hqc_amd.pbc import make_pes
    L = (4 / 3 * math.pi * N_ATOMS) ** (1 / 3)
    return make_pes(N_ATOMS, L, RS, BASIS, grid_length=GRID_LENGTH, dft=True, xc="lda,vwn",
                    smearing=True, smearing_method="fermi", smearing_sigma=SIGMA,
                    search_method="bisect", tol=TOL, precision=precision)


def time_vmat(pes, B, reps=20):
    """Average duration of the AO x grid potential-matrix kernel at the workload
    shape, from HIP events on the stream it is launched on (torch's current)."""
    import torch
    eng = pes.solver.engine
    c = eng.c
    dt = torch.float64 if eng.precision == "fp64" else torch.float32
    phi = torch.rand(B, c.nao_pad, c.ld, dtype=torch.float64, device="cuda").to(dt)
    phi[:, c.nao:] = 0
    phi[:, :, c.N:] = 0
    v = torch.zeros(B, c.ld, dtype=torch.float64, device="cuda")
    v[:, :c.N] = torch.rand(B, c.N, dtype=torch.float64, device="cuda")
    for _ in range(3):
        eng.vmat(phi, v)
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        eng.vmat(phi, v)
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    flops = 2.0 * c.nao ** 2 * c.N * B             # algorithmic, no symmetry credit
    return ms, flops


def cpu_baseline(sample_geoms=1):
    """Oracle (NumPy FP64, FFT-Poisson restatement) on the host: energy + analytic
    forces of `sample_geoms` config geometries."""
    from oracle import cpu_ref as R
    xps, L = workload_xp(sample_geoms)
    t0 = time.perf_counter()
    for b in range(sample_geoms):
        s = R.Solver(N_ATOMS, L, RS, BASIS, grid_length=GRID_LENGTH, dft=True, smearing=True,
                     smearing_sigma=SIGMA, tol=TOL)
        s.solve(xps[b])
        R.forces(s, xps[b])
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": sample_geoms / dt, "unit": "geometries/s", "cores": cores, "kind": "port",
            "sample": f"{sample_geoms} geometry (seed 0), full SCF + forces, oracle FFT route, "
                      f"{dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--path", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--compare", action="store_true",
                    help="also run the other precision path once and report its error")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (--gpus 1); "
              f"--gpus {a.gpus} is not supported.", file=sys.stderr)
        sys.exit(2)
    import torch
    from hqc_amd.pbc.engine import RY

    B = a.batch
    xps, _ = workload_xp(B)
    x = torch.tensor(xps, dtype=torch.float64, device="cuda")
    pes = make(a.path)
    for _ in range(a.warmup):
        pes.energy_and_forces(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        E, F = pes.energy_and_forces(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    st = pes.solver.last
    nit = st["nit"].double().mean().item()
    conv = bool(st["conv"].all().item())
    ms, flops = time_vmat(pes, B)
    prec = "f64" if a.path == "fp64" else "f32"
    achieved = flops / (ms * 1e-3) / 1e12
    peak = MI355X_PEAK[prec]
    out = {
        "metric": METRIC, "value": B / dt, "unit": "geometries/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.path == "fp64" else "f64+f32(increments)",
        "data": "synthetic: uniform-random protons, numpy default_rng(seed), seeds 0..B-1",
        "config": {"workload": f"54 H rs=1.25 gth-dzv LDA(vwn) Fermi T=1e4K, grid 49^3 "
                               f"(grid_length 0.16), tol 1e-8, energy+forces, batch {B}",
                   "path": a.path, "batch": B, "n_iter_mean": nit, "all_converged": conv},
        "roofline": {"bound": "mfma", "kernel": "k_vmat (dV phi diag(v) phi^T)",
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "ms_per_launch": ms, "flops_per_launch": flops},
    }
    eng = pes.solver.engine
    out["config"]["contractions"] = dict(eng.stats)
    if eng.eigh_calls:
        out["config"]["jacobi_sweeps_per_eigh"] = eng.sweeps.item() / eng.eigh_calls
    if a.compare:
        other = make("mixed" if a.path == "fp64" else "fp64")
        E2, F2 = other.energy_and_forces(x)
        w_a, w_b = st["w1"], other.solver.last["w1"]
        out["compare"] = {
            "against": other.solver.engine.precision,
            "max_dE_per_atom_Ha": (E - E2).abs().max().item() / RY / N_ATOMS,
            "max_dF_Ha_per_Bohr": (F - F2).abs().max().item(),
            "max_dbands_Ha": (w_a - w_b).abs().max().item(),
            "n_iter_mean_other": other.solver.last["nit"].double().mean().item()}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
