"""Throughput benchmark: energy+force geometries/sec for 54 H / gth-dzv.

Workload (one "step"): a batch of B independent 54-proton geometries
(rs = 1.25 Bohr, L = (4 pi 54/3)^(1/3) rs = 7.6162 Bohr, protons uniform in
the cell from numpy default_rng(seed), seeds 0..B-1), gth-dzv basis
(108 AOs), LDA (Slater + VWN), Fermi-Dirac smearing at T = 10^4 K
(sigma = 1/(2 beta), beta = 157888.088922572/T), the reference's odd grid for
grid_length 0.16 (49^3), SCF tol 1e-8 Ha, bisection mu search (make_pes
defaults otherwise); energy (E + V_pp) plus analytic forces for every
geometry.  Inputs are already on the GPU when the timed region starts.

`value` is the throughput of --path (default: the mixed-precision batched
path, batch 512: the batch's state, ~110 GB of AO values and digits plus the
force pass's 54 GB, fits the 288 GB of HBM, and the SCF's last iterations, where a
few slow geometries run alone, cost about the same for any batch size, so they
weigh half as much per geometry as at 256).  Beside it the run measures,
untimed for `value`:
  * the FP64 path on the same batch, whose results also give the accuracy of
    the mixed path (max |dE| per atom, |dF|, |d bands| against FP64);
  * the plain FP64 path, one geometry per call (seeds 0..3);
  * the roofline of the dominant kernel, from HIP events on its stream;
  * the oracle on the host for one geometry (cpu_baseline).

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
# dense MFMA TFLOP/s; MI355X's FP64 vector FMA peak is the same 78.6 TFLOP/s
MI355X_PEAK = {"f64": 78.6, "f32": 157.3, "i8": 5000.0}   # TFLOP/s (int8: TOP/s)
MI355X_HBM_GBS = 8000.0


def workload_xp(B, seed0=0):
    L = (4 / 3 * math.pi * N_ATOMS) ** (1 / 3)
    return np.stack([np.random.default_rng(seed0 + b).uniform(0, L, (N_ATOMS, 3))
                     for b in range(B)]), L


TUNING = {}       # path -> autotuned launch parameters (--autotune), else library defaults


def make(precision):
    from hqc_amd.pbc import make_pes
    L = (4 / 3 * math.pi * N_ATOMS) ** (1 / 3)
    return make_pes(N_ATOMS, L, RS, BASIS, grid_length=GRID_LENGTH, dft=True, xc="lda,vwn",
                    smearing=True, smearing_method="fermi", smearing_sigma=SIGMA,
                    search_method="bisect", tol=TOL, precision=precision,
                    tuning=TUNING.get(precision))


def vmat_kind(pes):
    """Which potential-matrix kernel the path's low-precision (or only) V runs on."""
    eng = pes.solver.engine
    if eng.precision == "fp64":
        return "f64"
    return "h2" if eng.MIXED_SPLIT16 and eng.c.nao_pad <= 128 else "f32"


def time_vmat(pes, B, reps=20):
    """Average duration of the AO x grid potential-matrix kernel at the workload
    shape, from HIP events on the stream it is launched on (torch's current):
    FP64 for the FP64 path, the FP16 split-product kernel (or FP32) for the mixed
    path's increments."""
    import torch
    from hqc_amd import _lib as K
    eng = pes.solver.engine
    c = eng.c
    kind = vmat_kind(pes)
    phi = torch.rand(B, c.nao_pad, c.ld, dtype=torch.float64, device="cuda")
    phi[:, c.nao:] = 0
    phi[:, :, c.N:] = 0
    if kind == "h2":
        ph = torch.empty(B, c.nao_pad, c.ld, dtype=torch.int32, device="cuda")
        K.phi_split_f16(phi, eng._phi_scale, ph)
        phi = ph
    elif kind == "f32":
        phi = phi.float()
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
    # algorithmic count of a symmetric rank-N update (V is symmetric): n(n+1) N
    # multiply-adds... as flops: 2 * n(n+1)/2 * N per geometry
    flops = 1.0 * c.nao * (c.nao + 1) * c.N * B
    return ms, flops


def time_orb_density(pes, B, q8=False, reps=20, top=5):
    """Average duration of the accurate orbital-form density kernel (rho = sum_i f_i
    psi_i^2) at the workload shape with the occupations of the last run: FP64 vector
    FMAs, or (q8) int8 digits of phi with exact integer products.  Returns (ms,
    flops, orbitals, phi bytes per launch, int8 MFMA ops per launch)."""
    import torch
    from hqc_amd import _lib as K
    eng = pes.solver.engine
    c = eng.c
    last = pes.solver.last
    mo, occ = last["mo"][:B], last["occ"][:B]
    keep = max(int((occ > eng.OCC_CUT).sum(-1).max().item()), 1)
    m_pad = (keep + 15) // 16 * 16
    At = torch.zeros(B, c.nao_pad, m_pad, dtype=torch.float64, device="cuda")
    At[:, :c.nao, :keep] = mo[:, :, :keep] * occ[:, None, :keep].clamp(min=0).sqrt()
    phi = torch.rand(B, c.nao_pad, c.ld, dtype=torch.float64, device="cuda")
    rho = torch.zeros(B, c.ld, dtype=torch.float64, device="cuda")
    ns = eng.Q8_SLICES
    if q8:
        q = torch.empty(B, c.ld * ns * c.nao_pad, dtype=torch.int8, device="cuda")
        pe = torch.empty(B, c.ld, dtype=torch.int16, device="cuda")
        K.phi_slice_i8(phi, ns, q, pe)
        del phi
        run = lambda: K.orb_density_q8(q, pe, At, rho, ns, top=top)
    else:
        run = lambda: K.orb_density(phi, At, rho)
    for _ in range(3):
        run()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        run()
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    flops = 2.0 * (c.nao + 1) * keep * c.N * B       # psi = C^T phi, then sum f psi^2
    # int8 digits: ns bytes per phi value + a 2-byte exponent per point in, rho out
    nbytes = B * c.ld * (ns * c.nao_pad + 2.0 + 8.0)
    # digit pairs (k, l) with l < ns, k <= top - l, each a 16x16x64 product over the
    # AO depth padded to 64 and the orbitals padded to 16
    pairs = sum(top + 1 - l for l in range(ns))
    kdepth = (c.nao_pad + 63) // 64 * 64
    ops = 2.0 * m_pad * kdepth * c.ld * B * pairs
    return ms, flops, keep, nbytes, ops


TRAFFIC_BATCH = 512
TRAFFIC_FILE = os.path.join(ROOT, "profiles", f"pmc_traffic_54H_dzv_b{TRAFFIC_BATCH}.txt")


def measured_traffic(kernel_prefix, B):
    """HBM bytes per launch of a kernel at this workload with batch TRAFFIC_BATCH, from
    the calibrated rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE table kept with the profiles
    (None for other batch sizes or when the table is absent)."""
    if B != TRAFFIC_BATCH or not os.path.isfile(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        parts = line.split()
        if (len(parts) >= 5 and not line.startswith("#")
                and line.split(None, 1)[1].startswith(kernel_prefix)):
            return float(parts[-1])
    return None


MFMA_FILE = os.path.join(ROOT, "profiles", f"pmc_mfma_54H_dzv_b{TRAFFIC_BATCH}.txt")


def measured_mfma_util(kernel_prefix, B):
    """rocprofv3 MfmaUtil (%) of a kernel at this workload with batch TRAFFIC_BATCH,
    from the table kept with the profiles (None for other batch sizes or when absent)."""
    if B != TRAFFIC_BATCH or not os.path.isfile(MFMA_FILE):
        return None
    for line in open(MFMA_FILE):
        parts = line.split()
        if len(parts) >= 4 and not line.startswith("#") and \
                line.split(None, 1)[1].startswith(kernel_prefix):
            for i, tok in enumerate(parts):      # the first number after the kernel name
                if tok.endswith(">"):
                    return float(parts[i + 1])
    return None


def _converged_diffs(E, F, w, E2, F2, w2, both):
    """Largest differences between two runs over the geometries both converged."""
    from hqc_amd.pbc.engine import RY
    if not bool(both.any()):
        return {}
    return {"max_dE_per_atom_Ha_converged": (E - E2)[both].abs().max().item() / RY / N_ATOMS,
            "max_dF_Ha_per_Bohr_converged": (F - F2)[both].abs().max().item(),
            "max_dbands_Ha_converged": (w - w2)[both].abs().max().item()}


def run_path(path, x, steps, warmup):
    """(pes, seconds per call of energy_and_forces(x), E, F)."""
    import torch
    pes = make(path)
    for _ in range(warmup):
        pes.energy_and_forces(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        E, F = pes.energy_and_forces(x)
    torch.cuda.synchronize()
    return pes, (time.perf_counter() - t0) / steps, E, F


def plain_fp64(ngeom=4):
    """The plain FP64 path: one geometry per call, seeds 0..ngeom-1 (after one
    untimed call)."""
    import torch
    xps, _ = workload_xp(ngeom)
    pes = make("fp64")
    pes.energy_and_forces(torch.tensor(xps[:1], dtype=torch.float64, device="cuda"))
    xs = [torch.tensor(xps[b:b + 1], dtype=torch.float64, device="cuda") for b in range(ngeom)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for x in xs:
        pes.energy_and_forces(x)
    torch.cuda.synchronize()
    return ngeom / (time.perf_counter() - t0)


def cpu_model():
    """Host CPU model name (from /proc/cpuinfo) and the thread count the oracle uses."""
    name = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                name = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return name, threads


def cpu_baseline(sample_geoms=1):
    """Oracle (NumPy FP64, FFT-Poisson restatement) on the host: energy + analytic
    forces of `sample_geoms` config geometries.  Returns (baseline record, oracle
    results of seed 0: E + V_pp [Ry], bands [Ha], F [Ha/Bohr])."""
    from oracle import cpu_ref as R
    xps, L = workload_xp(sample_geoms)
    t0 = time.perf_counter()
    res = None
    for b in range(sample_geoms):
        s = R.Solver(N_ATOMS, L, RS, BASIS, grid_length=GRID_LENGTH, dft=True, smearing=True,
                     smearing_sigma=SIGMA, tol=TOL)
        o = s.solve(xps[b])
        F = R.forces(s, xps[b])
        if b == 0:
            res = dict(E=o[3] + R.ewald_pp(xps[b], L, RS), w=o[2] / R.RY, F=F,
                       n_iter=s.last["n_iter"], converged=bool(o[8]))
    dt = time.perf_counter() - t0
    model, threads = cpu_model()
    return ({"value": sample_geoms / dt, "unit": "geometries/s", "cores": threads,
             "cpu_model": model, "kind": "port",
             "sample": f"{sample_geoms} geometry (seed 0) of the workload, full SCF + forces, "
                       f"oracle FFT-Poisson route, {dt:.1f} s"}, res)


LITERAL_N = 16


def cpu_baseline_literal():
    """The reference's own algorithm on the host: the oracle's literal route builds the
    pair-density tensor rhoG (g^3 n_ao^2 doubles) and takes J from it every iteration
    (hqc/pbc/solver.py:391-479, 616-629), energy only (the reference has no forces).
    At 54 H (11 GB rhoG, ~40 iterations over it) one geometry takes ~10 minutes on a
    few cores, so the timed sample is the same settings at 16 H (33^3); the 54 H time
    is projected from the two scalings (rhoG build ~ n^2 g^3, each J ~ n_ao^2 g^3)."""
    from oracle import cpu_ref as R
    n = LITERAL_N
    L = (4 / 3 * math.pi * n) ** (1 / 3)
    xp = np.random.default_rng(0).uniform(0, L, (n, 3))
    t0 = time.perf_counter()
    s = R.Solver(n, L, RS, BASIS, grid_length=GRID_LENGTH, dft=True, smearing=True,
                 smearing_sigma=SIGMA, tol=TOL, literal=True)
    s.solve(xp)
    dt = time.perf_counter() - t0
    g16, g54 = s.sys.g, 49
    scale = (N_ATOMS / n) ** 2 * (g54 / g16) ** 3
    model, threads = cpu_model()
    return {"value": 1.0 / dt, "unit": "geometries/s (energy only)", "cores": threads,
            "cpu_model": model, "kind": "port",
            "sample": f"literal rhoG route, {n} H gth-dzv rs 1.25 LDA Fermi 1e4 K, {g16}^3, "
                      f"seed 0, {s.last['n_iter']} SCF iterations, {dt:.1f} s",
            "projected_54H_s_per_geometry": dt * scale,
            "projection": f"x (54/{n})^2 (49/{g16})^3 = {scale:.1f}"}


def kernel_probe(name, B, reps=5):
    """Launch one kernel `reps` times at the workload shape (B geometries) and exit:
    the program that rocprofv3 --pmc counts.  'copy' is the calibration kernel: a
    device-to-device copy of exactly 1 GiB (read 1 GiB, write 1 GiB)."""
    import torch
    from hqc_amd import _lib as K
    if name == "copy":
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(reps):
            dst.copy_(src)
        torch.cuda.synchronize()
        return
    pes = make("mixed" if name in ("vmat_f32", "vmat_h2") else "fp64")
    c = pes.solver.engine.c
    if name == "vmat_h2":
        phi = torch.rand(B, c.nao_pad, c.ld, dtype=torch.float64, device="cuda")
        ph = torch.empty(B, c.nao_pad, c.ld, dtype=torch.int32, device="cuda")
        pes.solver.engine._phi_scale = 2.0 ** 13         # |phi| < 1
        K.phi_split_f16(phi, pes.solver.engine._phi_scale, ph)
        del phi
        v = torch.rand(B, c.ld, dtype=torch.float64, device="cuda")
        for _ in range(reps):
            pes.solver.engine.vmat(ph, v)
    elif name in ("vmat_f64", "vmat_f32"):
        dt = torch.float64 if name == "vmat_f64" else torch.float32
        phi = torch.rand(B, c.nao_pad, c.ld, dtype=torch.float64, device="cuda").to(dt)
        v = torch.rand(B, c.ld, dtype=torch.float64, device="cuda")
        for _ in range(reps):
            pes.solver.engine.vmat(phi, v)
    elif name == "orb_density":
        keep = 64
        At = torch.rand(B, c.nao_pad, keep, dtype=torch.float64, device="cuda")
        phi = torch.rand(B, c.nao_pad, c.ld, dtype=torch.float64, device="cuda")
        rho = torch.zeros(B, c.ld, dtype=torch.float64, device="cuda")
        for _ in range(reps):
            K.orb_density(phi, At, rho)
    elif name in ("orb_density_q8", "orb_density_q8_top4"):
        keep, ns = 64, pes.solver.engine.Q8_SLICES
        top = 4 if name.endswith("top4") else 5
        At = torch.rand(B, c.nao_pad, keep, dtype=torch.float64, device="cuda")
        phi = torch.rand(B, c.nao_pad, c.ld, dtype=torch.float64, device="cuda")
        q = torch.empty(B, c.ld * ns * c.nao_pad, dtype=torch.int8, device="cuda")
        pe = torch.empty(B, c.ld, dtype=torch.int16, device="cuda")
        K.phi_slice_i8(phi, ns, q, pe)
        del phi
        rho = torch.zeros(B, c.ld, dtype=torch.float64, device="cuda")
        for _ in range(reps):
            K.orb_density_q8(q, pe, At, rho, ns, top=top)
    elif name in ("exchange", "exchange_c64", "exchange_fused", "exchange_fused_c64"):
        # the Hartree-Fock exchange pass of one geometry: 16 orbital pairs packed
        # against the AO values, 3-D transformed (complex128, or complex64 for the
        # mixed path's early iterations) and split into the real contraction rows;
        # _fused: packed on the transform's first pass, as the solver does
        mp = 16
        c64 = name.endswith("_c64")
        fused = "fused" in name
        phi = torch.rand(1, c.nao_pad, c.ld, dtype=torch.float64, device="cuda")
        psi = torch.rand(1, 2 * mp, c.ld, dtype=torch.float64, device="cuda")
        table = torch.tensor([[0, 2 * k, 2 * k + 1] for k in range(mp)], dtype=torch.int32)
        z = torch.empty(mp, c.nao, c.N, device="cuda",
                        dtype=torch.complex64 if c64 else torch.complex128)
        M = torch.empty(2 * mp, c.nao_pad, c.ld, device="cuda",
                        dtype=torch.float32 if c64 else torch.float64)
        for _ in range(reps):
            if fused:
                K.pair_pack(phi, psi, table, c.nao, c.g, z, fft_sign=-1)
            else:
                K.pair_pack(phi, psi, table, c.nao, c.g, z)
                K.fft3d(z.view(mp * c.nao, c.g, c.g, c.g), -1)
            K.pair_split(z, c.nao_pad, c.ld, M)
    else:
        raise SystemExit(f"unknown kernel {name}")
    torch.cuda.synchronize()


def summarize_trace(d, top=25):
    """Text table (kernel, calls, avg us, total ms, %) from a rocprofv3 results db
    (its top_kernels view holds durations in microseconds)."""
    import glob
    import re
    import sqlite3
    db = glob.glob(os.path.join(d, "**", "*.db"), recursive=True)[0]
    cur = sqlite3.connect(db).cursor()
    rows = cur.execute("select name,total_calls,total_duration,average,percentage "
                       "from top_kernels").fetchall()

    def short(n):
        n = re.sub(r"\(.*", "", n)
        return ("rocBLAS " + n[:40] + "...") if n.startswith("Cijk_") else n[:70]
    tot = sum(r[2] for r in rows)
    lines = [f"{'kernel':72s} {'calls':>6s} {'avg_us':>9s} {'total_ms':>9s} {'%':>6s}"]
    for r in rows[:top]:
        lines.append(f"{short(r[0]):72s} {r[1]:6d} {r[3]:9.1f} {r[2] / 1e3:9.2f} "
                     f"{r[4]:6.2f}")
    lines.append(f"total GPU kernel time: {tot / 1e3:.1f} ms over "
                 f"{sum(r[1] for r in rows)} dispatches")
    return "\n".join(lines)


def summarize_pmc(d, n_xcd=8, n_simd=1024):
    """Mean value per dispatch of each counter, per kernel, from rocprofv3 --pmc
    csv output; then per kernel the mean dispatch duration (the csv's timestamps)
    and, where the counters are there, the engine clock (GRBM_GUI_ACTIVE, which sums
    the XCDs, per XCD over the duration) and the MFMA pipe's busy share
    (SQ_VALU_MFMA_BUSY_CYCLES over GRBM_GUI_ACTIVE per XCD x the SIMDs)."""
    import csv
    import glob
    acc, dur = {}, {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for row in csv.DictReader(open(f)):
            k = row["Kernel_Name"][:60]
            acc.setdefault((k, row["Counter_Name"]), []).append(float(row["Counter_Value"]))
            dur.setdefault(k, {})[row["Dispatch_Id"]] = (
                int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) * 1e-9
    lines = [f"{'kernel':62s} {'counter':14s} {'dispatches':>10s} {'mean':>14s}"]
    for (k, cn), vals in sorted(acc.items()):
        lines.append(f"{k:62s} {cn:14s} {len(vals):10d} {sum(vals) / len(vals):14.1f}")
    lines.append(f"{'kernel':62s} {'duration_ms':>11s} {'clock_GHz':>9s} {'mfma_busy_%':>11s}")
    for k in sorted(dur):
        t = sum(dur[k].values()) / len(dur[k])
        mean = lambda cn: (sum(acc[(k, cn)]) / len(acc[(k, cn)])) if (k, cn) in acc else None
        gui, busy = mean("GRBM_GUI_ACTIVE"), mean("SQ_VALU_MFMA_BUSY_CYCLES")
        clk = f"{gui / n_xcd / t / 1e9:9.2f}" if gui and t > 0 else f"{'-':>9s}"
        util = f"{100 * busy / (gui / n_xcd * n_simd):11.1f}" if gui and busy is not None \
            else f"{'-':>11s}"
        lines.append(f"{k:62s} {t * 1e3:11.3f} {clk} {util}")
    return "\n".join(lines)


# ---------------------------------------------------------------------------
# Other workloads (non-default; --workload): the uneven Hartree-Fock PES batch, one
# 128 H gth-dzvp Hartree-Fock cell and the 54 H PBE 2x2x2 k-mesh.  Each prints one
# JSON line of the same shape with its own metric; `value` is again whole-job
# throughput with the positions already on the GPU.
# ---------------------------------------------------------------------------
def hf_batch_workload(B, seed=0):
    """B geometries, atom counts uniform from {16, 32, 54, 64}, rs uniform in [1, 1.5]
    Bohr, protons uniform in the uniform-gas cell (rs units), from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    ns = rng.choice([16, 32, 54, 64], B)
    rss = rng.uniform(1.0, 1.5, B)
    xps = [rng.uniform(0, (4 / 3 * math.pi * n) ** (1 / 3), (n, 3)) for n in ns]
    return ns, rss, xps


def run_workload(a):
    import torch
    from hqc_amd.pbc import make_pes, make_pes_batch
    out = {"n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "data": "synthetic: uniform-random protons"}
    if a.workload == "hf-batch":
        ns, rss, xps = hf_batch_workload(a.batch)
        t_start = time.perf_counter()

        def progress(n, g, k):
            print(f"bucket n={n} g={g}: {k} geometries done "
                  f"({time.perf_counter() - t_start:.1f} s)", file=sys.stderr, flush=True)
        pes = make_pes_batch("gth-dzv", grid_length=GRID_LENGTH, dft=False, tol=TOL,
                             precision=a.path, workers=a.workers, progress=progress)
        run = lambda: pes.energy_and_forces(xps, rss)
        units = len(xps)
        out.update(metric="energy+force geometries/sec, uneven HF PES batch (gth-dzv)",
                   unit="geometries/s",
                   config={"workload": f"{units} geometries, n in {{16,32,54,64}}, rs in [1,1.5], "
                                       "restricted HF at Gamma, gth-dzv, grid spacing 0.16 Bohr, "
                                       f"tol {TOL}, energy+forces", "path": a.path,
                           "workers": a.workers})
    elif a.workload == "hf-128":
        n, rs = 128, 1.25
        L = (4 / 3 * math.pi * n) ** (1 / 3)
        xp = torch.tensor(np.random.default_rng(1).uniform(0, L, (1, n, 3)), device="cuda")
        pes = make_pes(n, L, rs, "gth-dzvp", grid_length=GRID_LENGTH, dft=False, tol=TOL,
                       ovlp_cut=1e-4, max_cycle=200, precision=a.path)
        run = lambda: pes.energy_and_forces(xp)
        units = 1
        out.update(metric="energy+force geometries/sec, 128 H gth-dzvp HF (640 AOs, 65^3)",
                   unit="geometries/s",
                   config={"workload": "128 H rs 1.25 gth-dzvp restricted HF, 65^3, full "
                                       "exchange, ovlp_cut 1e-4, seed 1, energy+forces",
                           "path": a.path})
    elif a.workload == "kmesh":
        n, rs = 54, 1.4
        L = (4 / 3 * math.pi * n) ** (1 / 3)
        xp = np.random.default_rng(0).uniform(0, L, (n, 3))
        mp = np.array([[i, j, k] for i in (-0.25, 0.25) for j in (-0.25, 0.25)
                       for k in (-0.25, 0.25)])
        kpts = torch.tensor(mp * 2 * np.pi / (L * rs), device="cuda")
        x8 = torch.tensor(np.repeat(xp[None], 8, 0), device="cuda")
        pes = make_pes(n, L, rs, "gth-dzv", grid_length=0.178, dft=True, xc="pbe,pbe",
                       smearing=True, smearing_sigma=1.0 / (157888.088922572 / 5000.0) / 2.0,
                       tol=TOL, search_tol=1e-12, search_cycle=200, max_cycle=300,
                       gamma=False, precision=a.path)
        run = lambda: pes.energy_and_forces(x8, kpts)
        units = 8
        out.update(metric="energy+force k-points/sec, 54 H PBE 2x2x2 Monkhorst-Pack mesh",
                   unit="k-points/s",
                   config={"workload": "54 H rs 1.4 gth-dzv PBE Fermi 5000 K, 2x2x2 MP mesh "
                                       "of one geometry, 49^3, energy+forces", "path": a.path})
    else:
        raise SystemExit(f"unknown workload {a.workload}")
    for _ in range(a.warmup):
        run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res = run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(a.steps, 1)
    if a.workload == "hf-batch":
        info = res[2]
        out["config"]["converged"] = int(sum(info["converged"]))
        out["config"]["n_iter_mean"] = float(np.mean(info["n_iter"]))
        out["config"]["buckets"] = len(info["buckets"])
    else:
        st = pes.solver.last
        out["config"]["n_iter"] = st["nit"].tolist()
        out["config"]["converged"] = bool(st["conv"].all())
    out.update(value=units / dt, ms_per_step=dt * 1e3,
               dtype="f64" if a.path == "fp64" else "f64+f32")
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--path", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-literal", action="store_true",
                    help="skip the CPU row of the reference's literal rhoG route")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the FP64 batched and plain runs (accuracy and speedups)")
    ap.add_argument("--kernel-probe", default=None,
                    help="vmat_f64|vmat_f32|vmat_h2|orb_density|orb_density_q8[_top4]|exchange|"
                         "exchange_c64|exchange_fused|exchange_fused_c64|copy: launch that "
                         "kernel (or the exchange pass) only")
    ap.add_argument("--workload", default="metric",
                    choices=["metric", "hf-batch", "hf-128", "kmesh"],
                    help="metric (default): the 54 H LDA batch; the others are the uneven "
                         "HF PES batch (--batch geometries), 128 H dzvp HF and the PBE k-mesh")
    ap.add_argument("--autotune", action="store_true",
                    help="sweep the kernels' launch parameters at the workload shape first "
                         "(untimed) and run both paths with the choice")
    ap.add_argument("--workers", type=int, default=1,
                    help="hf-batch: (n, grid) buckets solved concurrently on their own streams")
    ap.add_argument("--summarize-trace", default=None, metavar="DIR")
    ap.add_argument("--summarize-pmc", default=None, metavar="DIR")
    ap.add_argument("--top", type=int, default=25, help="rows of --summarize-trace")
    a = ap.parse_args()
    if a.summarize_trace:
        print(summarize_trace(a.summarize_trace, a.top))
        return
    if a.summarize_pmc:
        print(summarize_pmc(a.summarize_pmc))
        return
    if a.gpus != 1:
        print("bench.py measures one GPU only (--gpus 1); "
              f"--gpus {a.gpus} is not supported.", file=sys.stderr)
        sys.exit(2)
    import torch
    from hqc_amd.pbc.engine import RY

    if a.workload != "metric":
        run_workload(a)
        return
    B = a.batch
    if a.kernel_probe:
        kernel_probe(a.kernel_probe, B)
        return
    xps, _ = workload_xp(B)
    x = torch.tensor(xps, dtype=torch.float64, device="cuda")
    tuned = None
    if a.autotune:
        # untimed: sweep the launch parameters at the workload shape for both paths
        from hqc_amd import tune
        tuned = {}
        for path in ("mixed", "fp64"):
            rec = tune.autotune(make(path).solver.engine, B)
            TUNING[path] = rec["choice"]
            tuned[path] = rec["choice"]
        torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    pes, dt, E, F = run_path(a.path, x, a.steps, a.warmup)
    hbm_peak = torch.cuda.max_memory_allocated() / 1e9
    st = pes.solver.last
    eng = pes.solver.engine
    nit = st["nit"].double().mean().item()
    conv = bool(st["conv"].all().item())
    stats = dict(eng.stats)
    sweeps = eng.sweeps.item() / eng.eigh_calls if eng.eigh_calls else None
    # DIIS iterations of the last call and how many geometries each one carried
    act = list(eng.active)
    tail = {"iterations": len(act), "geometry_iterations": int(sum(act)),
            "iterations_below_16_geometries": sum(1 for a in act if a < 16)}
    # dominant kernels: potential matrix (FP64 or FP32) and FP64 orbital density
    kind = vmat_kind(pes)
    ms_v, fl_v = time_vmat(pes, B)
    q8 = a.path == "mixed" and sum(stats.get(k, 0) for k in ("rhoq8", "rhoq8_top4",
                                                             "rhoq8_top3")) > 0
    n_vmat = stats["v64"] if a.path == "fp64" else stats["v32"]
    nsteps = max(a.steps + a.warmup, 1)
    c = eng.c
    if kind == "h2":
        # FP16 pairs (4 bytes per phi value) + v (8 bytes per point) per launch: HBM roof
        vname = "k_vmat_h2"
        vk = {"kernel": "k_vmat_h2 (dV phi diag(v) phi^T from FP16 hi/lo split products, "
                        "FP32 accumulation)", "bound": "hbm", "unit": "GB/s",
              "bytes_per_launch": B * c.ld * (4.0 * c.nao_pad + 8.0),
              "peak": MI355X_HBM_GBS}
    else:
        prec_v = "f64" if kind == "f64" else "f32"
        vname = "k_vmat_pipe<" + ("double" if prec_v == "f64" else "float")
        vk = {"kernel": vname + "> (dV phi diag(v) phi^T, symmetric: n(n+1)N flops)",
              "bound": "mfma", "unit": "TFLOP/s", "peak": MI355X_PEAK[prec_v]}
    vk.update({"ms_per_launch": ms_v, "flops_per_launch": fl_v,
               "calls_per_step": n_vmat / nsteps, "traffic": measured_traffic(vname, B),
               "mfma_util_pct": measured_mfma_util(vname, B)})
    cand = [vk]
    if q8:
        # bound by the int8 MFMA pipe (16 busy cycles per 16x16x64 instruction, busy
        # ~65% of the time at the clock the load allows), not by its stream of int8
        # digits (4 bytes per phi value), which is reported beside it; units of the
        # MFMA roof: tera int8 operations per second.  One entry per digit-product
        # level the SCF ran (level 3 while the DIIS residual is >= 1e-6, level 4 after,
        # by default).
        for top, key in ((3, "rhoq8_top3"), (4, "rhoq8_top4"), (5, "rhoq8")):
            if not stats.get(key, 0):
                continue
            ms_d, fl_d, keep, by_d, ops_d = time_orb_density(pes, B, True, top=top)
            kname = f"k_orb_density_q8<4, 2, {top}>"
            cand.append({
                "kernel": "k_orb_density_q8 (sum_i f_i (C^T phi)_i^2 from int8 digits, exact "
                          f"int32 products to level {top}; int8 MFMA ops)",
                "bound": "mfma", "unit": "TFLOP/s", "ops_per_launch": ops_d,
                "peak": MI355X_PEAK["i8"], "bytes_per_launch": by_d,
                "hbm_gbs": by_d / (ms_d * 1e-3) / 1e9,
                "hbm_frac": by_d / (ms_d * 1e-3) / 1e9 / MI355X_HBM_GBS,
                "traffic": measured_traffic(kname, B),
                "mfma_util_pct": measured_mfma_util(kname, B),
                "ms_per_launch": ms_d, "flops_per_launch": fl_d, "orbitals": keep,
                "calls_per_step": stats[key] / nsteps})
    else:
        ms_d, fl_d, keep, by_d, ops_d = time_orb_density(pes, B, False)
        cand.append({"kernel": "k_orb_density_valu (sum_i f_i (C^T phi)_i^2, FP64 vector FMA)",
                     "bound": "mfma", "unit": "TFLOP/s", "peak": MI355X_PEAK["f64"],
                     "traffic": measured_traffic("k_orb_density_valu", B),
                     "ms_per_launch": ms_d, "flops_per_launch": fl_d, "orbitals": keep,
                     "calls_per_step": stats["rho64"] / nsteps})
    for k in cand:
        if k["unit"] == "GB/s":
            k["achieved"] = k["bytes_per_launch"] / (k["ms_per_launch"] * 1e-3) / 1e9
        else:
            work = k.get("ops_per_launch", k["flops_per_launch"])
            k["achieved"] = work / (k["ms_per_launch"] * 1e-3) / 1e12
        k["frac"] = k["achieved"] / k["peak"]
    dom = max(cand, key=lambda k: k["ms_per_launch"] * k["calls_per_step"])
    w_main = st["w1"].clone()
    conv_main = st["conv"].bool().clone()
    del pes, st, eng
    torch.cuda.empty_cache()
    out = {
        "metric": METRIC, "value": B / dt, "unit": "geometries/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.path == "fp64" else
                 ("f64+f16x2(split increments)" if kind == "h2" else "f64+f32(increments)")
                 + ("+i8(digit density)" if q8 else ""),
        "data": "synthetic: uniform-random protons, numpy default_rng(seed), seeds 0..B-1",
        "config": {"workload": f"54 H rs=1.25 gth-dzv LDA(vwn) Fermi T=1e4K, grid 49^3 "
                               f"(grid_length 0.16), tol 1e-8, energy+forces, batch {B}",
                   "path": a.path, "batch": B, "n_iter_mean": nit, "all_converged": conv,
                   "contractions": stats, "jacobi_sweeps_per_eigh": sweeps,
                   "diis_loop": tail, "hbm_peak_gb": round(hbm_peak, 1)},
        # the dominant kernel's own roof: HBM for the split-FP16 V, the int8 MFMA pipe
        # for the digit density, MFMA / FP64 vector FMAs (same 78.6 TFLOP/s peak) for
        # the FP64 forms
        "roofline": {"bound": dom["bound"], "kernel": dom["kernel"], "achieved": dom["achieved"],
                     "peak": dom["peak"], "unit": dom["unit"], "frac": dom["frac"],
                     "traffic": dom["traffic"],
                     "traffic_source": None if dom["traffic"] is None else
                     "profiles/" + os.path.basename(TRAFFIC_FILE) + " (rocprofv3 --pmc "
                     "FETCH_SIZE x2.0, WRITE_SIZE x1.0 from a 1 GiB copy; not measured in "
                     "this run)",
                     "mfma_util_pct": dom.get("mfma_util_pct"),
                     "mfma_util_source": None if dom.get("mfma_util_pct") is None else
                     "profiles/" + os.path.basename(MFMA_FILE) + " (rocprofv3 --pmc MfmaUtil "
                     "of the kernel probe; not measured in this run)",
                     "ms_per_launch": dom["ms_per_launch"],
                     "flops_per_launch": dom["flops_per_launch"],
                     **{k: dom[k] for k in ("ops_per_launch", "bytes_per_launch", "hbm_gbs",
                                            "hbm_frac") if k in dom}},
        "kernels": cand,
        "tuning": tuned if tuned is not None else "library defaults (no --autotune)",
    }
    seed0 = {a.path: (E[0].item(), F[0].cpu().numpy(), w_main[0].cpu().numpy())}
    if not a.no_compare:
        other = "fp64" if a.path == "mixed" else "mixed"
        pes2, dt2, E2, F2 = run_path(other, x, 1, 1)
        w_a, w_b = w_main, pes2.solver.last["w1"]
        seed0[other] = (E2[0].item(), F2[0].cpu().numpy(), w_b[0].cpu().numpy())
        out["compare"] = {
            "against": other, "batched_value": B / dt2,
            "max_dE_per_atom_Ha": (E - E2).abs().max().item() / RY / N_ATOMS,
            "max_dF_Ha_per_Bohr": (F - F2).abs().max().item(),
            "max_dbands_Ha": (w_a - w_b).abs().max().item(),
            # the same over the geometries both runs converged (a geometry left at
            # max_cycle stops wherever its last iteration was, in either run)
            "converged_both": int((conv_main & pes2.solver.last["conv"].bool()).sum().item()),
            **_converged_diffs(E, F, w_a, E2, F2, w_b,
                               conv_main & pes2.solver.last["conv"].bool()),
            "n_iter_mean_other": pes2.solver.last["nit"].double().mean().item(),
            "all_converged_other": bool(pes2.solver.last["conv"].all().item())}
        del pes2
        torch.cuda.empty_cache()
        plain = plain_fp64()
        out["compare"]["fp64_plain_value"] = plain
        out["speedup_vs_fp64_plain"] = out["value"] / plain
        fp64_batched = out["value"] if a.path == "fp64" else B / dt2
        out["speedup_vs_fp64_batched"] = out["value"] / fp64_batched
    if not a.no_cpu:
        base, ref = cpu_baseline(1)
        out["cpu_baseline"] = base
        # both GPU paths' seed-0 results against the oracle (the accuracy bar of the
        # metric is 1 uHa/atom, 1e-5 Ha/Bohr, 1e-6 Ha bands)
        chk = {"seed": 0, "oracle_n_iter": ref["n_iter"], "oracle_converged": ref["converged"]}
        for path, (Eg, Fg, wg) in seed0.items():
            chk[path] = {"dE_per_atom_Ha": abs(Eg - ref["E"]) / RY / N_ATOMS,
                         "dF_Ha_per_Bohr": float(np.abs(Fg - ref["F"]).max()),
                         "dbands_Ha": float(np.abs(wg - ref["w"]).max())}
        out["oracle_check"] = chk
        if not a.no_literal:
            out["cpu_baseline_literal"] = cpu_baseline_literal()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
