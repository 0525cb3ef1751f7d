"""Batched PES over geometries of different atom counts and cell sizes.

The reference evaluates one geometry per call (hqc/pbc/pes.py:78-86) and a
sweep over systems of different sizes needs one factory, and one compile, per
(n, L).  Here a list of geometries is split into buckets of equal atom count n
and equal odd grid g (hqc/pbc/solver.py:159); each bucket runs as one batched
SCF in which every geometry keeps its own cell length L (so its own dV, VG,
lattice sums and image count), then the batched force adjoint.

    pes = make_pes_batch("gth-dzv", grid_length=0.16, dft=False)
    E, F, info = pes.energy_and_forces(xps, rs)   # lists, input order

xps[i] (n_i, 3) in rs units, rs[i] Bohr, L[i] in rs units (default the
uniform-gas cell (4 pi n_i / 3)^(1/3)); E[i] = E_elec + V_pp in Ry,
F[i] (n_i, 3) in Ha/Bohr (forces of the free energy, as PES.energy_and_forces).
"""
import concurrent.futures as concurrent_futures
import contextlib
import math
import threading

import numpy as np
import torch

from .engine import RY, Cell, Engine
from .forces import check_supported, compute_forces


def uniform_gas_length(n):
    """Cell length in rs units holding n electrons at density 3/(4 pi rs^3)."""
    return (4.0 / 3.0 * math.pi * n) ** (1.0 / 3.0)


class BatchPES:
    """Buckets of equal (n, odd grid) solved as batched SCFs.

    workers > 1: the buckets run concurrently, each in a worker thread on its own
    HIP stream (default priority), largest estimated cost first.  A bucket's SCF
    stops the host at every convergence check; with one stream those stops leave
    the GPU idle between small buckets, with several the other buckets' queued work
    fills them.  Results are the same numbers as with workers = 1 (the same kernels
    on the same data; only the interleaving differs)."""

    def __init__(self, basis, rcut=24, grid_length=0.12, max_batch=64, device="cuda",
                 kappa=10, Gmax=15, workers=1, progress=None, **engine_kw):
        self.basis, self.rcut, self.grid_length = basis, rcut, grid_length
        self.max_batch, self.device = int(max_batch), device
        self.kappa, self.Gmax = kappa, Gmax
        self.workers = max(1, int(workers))
        self.progress = progress      # called with (n, g, geometries) after each bucket
        self.engine_kw = engine_kw
        self._engines = {}
        self.last_states = {}
        self._local = threading.local()

    def _engine(self, n, L, rs, g, cache=True):
        key = (n, g)
        if not cache or key not in self._engines:
            cell = Cell(n, L, rs, self.basis, rcut=self.rcut, grid_length=self.grid_length,
                        device=self.device)
            assert cell.g == g
            eng = Engine(cell, **self.engine_kw)
            if not cache:
                return eng
            self._engines[key] = eng
        return self._engines[key]

    def buckets(self, ns, Lb_bohr):
        """{(n, g): [indices]} in input order."""
        out = {}
        for i, (n, Lb) in enumerate(zip(ns, Lb_bohr)):
            g = round(Lb / self.grid_length / 2) * 2 + 1
            out.setdefault((n, g), []).append(i)
        return out

    def _stream(self):
        s = getattr(self._local, "stream", None)
        if s is None:
            s = torch.cuda.Stream(device=self.device)        # default priority
            self._local.stream = s
        return s

    def _solve_bucket(self, n, g, idx, xps, rs, Lrs, Lb, concurrent):
        """Solve the geometries idx of one bucket in chunks of max_batch; returns
        [(chunk indices, E, F, converged, n_iter, light state)]."""
        out = []
        stream = self._stream() if concurrent else None
        ctx = torch.cuda.stream(stream) if concurrent else contextlib.nullcontext()
        with ctx:
            for s0 in range(0, len(idx), self.max_batch):
                part = idx[s0:s0 + self.max_batch]
                # a worker gets engines of its own: an engine's buffers belong to one stream
                eng = self._engine(n, Lrs[part[0]], rs[part[0]], g, cache=not concurrent)
                check_supported(eng)
                xpb = torch.tensor(np.stack([np.asarray(xps[i], dtype=float) * rs[i]
                                             for i in part]), dtype=torch.float64,
                                   device=eng.c.device)
                Lt = torch.tensor(Lb[part], dtype=torch.float64, device=eng.c.device)
                st = eng.solve(xpb, Lb=Lt, forces=True)
                Fb = compute_forces(eng, st, self.kappa, self.Gmax)
                Eb = st["E"] * RY + st["Epp"]
                # the n_ao x n_ao results of the chunk (S and the slot order as the SCF
                # left them; D, F, bands in the chunk's input order), not the grid arrays
                light = {k: st[k] for k in ("S", "D", "F", "w1", "order", "conv", "nit")
                         if k in st}
                light["index"] = list(part)
                out.append((s0, part, Eb, Fb, st["conv"].tolist(), st["nit"].tolist(), light))
            if concurrent:
                stream.synchronize()
        if self.progress is not None:
            self.progress(n, g, len(idx))
        return out

    def energy_and_forces(self, xps, rs, L=None):
        ng = len(xps)
        rs = np.broadcast_to(np.asarray(rs, dtype=float), (ng,))
        ns = [int(np.asarray(x).shape[0]) for x in xps]
        Lrs = np.array([uniform_gas_length(n) for n in ns]) if L is None else \
            np.broadcast_to(np.asarray(L, dtype=float), (ng,))
        Lb = Lrs * rs
        E, F = [None] * ng, [None] * ng
        conv, nit = [None] * ng, [None] * ng
        self.last_states = {}
        bk = self.buckets(ns, Lb)
        # largest estimated cost first (n_ao^2 grid points per geometry x count)
        order = sorted(bk.items(), key=lambda kv: -(kv[0][0] ** 2 * kv[0][1] ** 3 * len(kv[1])))
        concurrent = self.workers > 1 and len(order) > 1
        if concurrent:
            main = torch.cuda.current_stream()
            streams = []

            def task(item):
                (n, g), idx = item
                res = self._solve_bucket(n, g, idx, xps, rs, Lrs, Lb, True)
                streams.append(self._stream())
                return (n, g), res
            with concurrent_futures.ThreadPoolExecutor(self.workers) as ex:
                results = list(ex.map(task, order))
            for s in set(streams):
                main.wait_stream(s)
        else:
            results = [((n, g), self._solve_bucket(n, g, idx, xps, rs, Lrs, Lb, False))
                       for (n, g), idx in order]
        for (n, g), res in results:
            for s0, part, Eb, Fb, cb, nb, light in res:
                if concurrent:                       # read later on the caller's stream
                    for t in [Eb, Fb] + [v for v in light.values() if torch.is_tensor(v)]:
                        t.record_stream(torch.cuda.current_stream())
                self.last_states[(n, g, s0)] = light
                for j, i in enumerate(part):
                    E[i], F[i] = Eb[j], Fb[j]
                    conv[i], nit[i] = bool(cb[j]), int(nb[j])
        return E, F, {"converged": conv, "n_iter": nit,
                      "buckets": {k: len(v) for k, v in bk.items()}}


def make_pes_batch(basis, rcut=24, tol=1e-7, max_cycle=100, grid_length=0.12, diis=True,
                   diis_space=8, diis_start_cycle=1, diis_damp=0.0, dft=False, xc="lda,vwn",
                   smearing=False, smearing_method="fermi", smearing_sigma=0.0,
                   search_method="bisect", search_cycle=100, search_tol=1e-7, Gmax=15,
                   kappa=10, precision="fp64", max_batch=64, ovlp_cut=0.0, workers=1,
                   tuning=None, progress=None):
    """make_pes's options (hqc/pbc/pes.py:10-31) for a list of systems of any size;
    workers > 1 solves that many (n, grid) buckets concurrently on their own streams."""
    return BatchPES(basis, rcut=rcut, grid_length=grid_length, max_batch=max_batch,
                    kappa=kappa, Gmax=Gmax, workers=workers, progress=progress, dft=dft, xc=xc,
                    smearing=smearing,
                    smearing_method=smearing_method, smearing_sigma=smearing_sigma,
                    search_method=search_method, search_cycle=search_cycle,
                    search_tol=search_tol, tol=tol, max_cycle=max_cycle, diis=diis,
                    diis_space=diis_space, diis_start_cycle=diis_start_cycle,
                    diis_damp=diis_damp, precision=precision, ovlp_cut=ovlp_cut, tuning=tuning)
