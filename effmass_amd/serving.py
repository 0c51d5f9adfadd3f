"""Serving loop for effective-mass queries (the config-4 workload).

A query is a cell of 2-64 atoms (diamond-derived, 135 bohr^3 per atom), a
cutoff, a band count and 1-4 k-points; its answer is the k.p effective-mass
tensor of a band at every k-point.  The server keeps each admitted query's
data resident on the device (V_loc, projector tables and matrices, psi) for
as long as the query runs, and runs queries concurrently:

* admission by HBM footprint: a query starts only when a worker slot is free
  and its estimated device bytes fit under the server's budget next to the
  queries already running (FIFO order, so no query starves);
* continuous batching: queries that arrive while others run are admitted
  as soon as slots and memory free up.  Each running query issues its
  kernels on its own HIP stream (default priority), so their work
  overlaps on the device;
* every query gets its own engine context with an HBM budget equal to its
  estimate, so a wrong estimate fails that query loudly instead of
  starving the others.

Results do not depend on what else runs: all kernels are deterministic
(no atomics; split-K partial sums are reduced in a fixed order), so a query
returns the same numbers alone and in the mix.
"""
from __future__ import annotations

import math
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import effmass, eigensolver, synth
from .hamiltonian import Hamiltonian

VOLUME_PER_ATOM = 135.0
SIZES = (2, 8, 16, 32, 64)
ECUTS = (30.0, 40.0, 50.0)


def smooth_size(n):
    """Smallest m >= n whose prime factors are 2, 3, 5 (FFT-friendly)."""
    m = max(1, n)
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def query_cell(natom: int):
    """Lattice and fractional positions for natom in {2, 8, 16, 32, 64} with
    135 bohr^3 per atom: fcc primitive for 2, diamond cubic supercells
    (1x1x1, 2x1x1, 2x2x1, 2x2x2) otherwise."""
    if natom == 2:
        a = (4 * 2 * VOLUME_PER_ATOM) ** (1 / 3)
        return synth.fcc_primitive(a)
    a = (8 * VOLUME_PER_ATOM) ** (1 / 3)
    rep = {8: (1, 1, 1), 16: (2, 1, 1), 32: (2, 2, 1), 64: (2, 2, 2)}[natom]
    pos = []
    for i in range(rep[0]):
        for j in range(rep[1]):
            for k in range(rep[2]):
                for f in synth.DIAMOND8:
                    pos.append((f + np.array([i, j, k])) / np.array(rep))
    return np.diag(a * np.array(rep, dtype=float)), np.array(pos)


def grid_for(A, ecut_ry):
    """FFT grid holding the density sphere: n >= 2 G_max |a_i| / 2 pi + 1."""
    gmax = math.sqrt(ecut_ry)
    return tuple(smooth_size(int(math.ceil(2 * gmax * np.linalg.norm(a) / (2 * math.pi))) + 1) for a in A)


@dataclass
class MassQuery:
    qid: int
    natom: int
    ecut_ry: float
    nb: int
    kfracs: list
    seed: int
    arrival: float = 0.0          # seconds after the start of the run
    band: int = -1                # band index (default nb // 2)


@dataclass
class MassAnswer:
    qid: int
    inv_mass: list                # per k: 3x3 (1/m*)
    energies: list                # per k: E of the band
    latency: float
    start: float
    finish: float
    npw: list = field(default_factory=list)
    error: str | None = None


def make_stream(n, rate=50.0, seed=2024, nb_range=(16, 256), sizes=SIZES, ecuts=ECUTS, max_k=4):
    """n queries with Poisson arrivals at `rate` per second."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    out = []
    for i in range(n):
        nk = int(rng.integers(1, max_k + 1))
        out.append(MassQuery(qid=i, natom=int(rng.choice(sizes)), ecut_ry=float(rng.choice(ecuts)),
                             nb=int(rng.integers(nb_range[0], nb_range[1] + 1)),
                             kfracs=[rng.uniform(-0.5, 0.5, 3) for _ in range(nk)],
                             seed=int(rng.integers(1, 2 ** 31)), arrival=float(t[i])))
    return out


def build_workload(q: MassQuery):
    A, tau = query_cell(q.natom)
    n = grid_for(A, q.ecut_ry)
    return synth.custom(A, tau, n, q.ecut_ry, q.kfracs, q.nb, q.seed, with_psi=False)


def estimate_bytes(q: MassQuery, wl=None, prec="mixed"):
    """Device bytes a query holds while it runs: engine data (V_loc, tables,
    projectors and their k-gradients, FFT workspace for a 3nb LOBPCG block,
    split-K partials) plus the solver's blocks (FP64 X, HX, S, HS, P)."""
    wl = wl or build_workload(q)
    c = wl.cell
    N = int(np.prod(c.n))
    npw = max(k.npw for k in wl.kpts)
    nch = c.nchannels()
    S = 8 if prec == "mixed" else 16
    blk = 3 * q.nb
    planes = min(c.n[0], int(2 * math.sqrt(q.ecut_ry) * np.linalg.norm(c.A[0]) / (2 * math.pi)) + 2)
    engine = N * 12 + 4 * nch * npw * S + planes * c.n[1] * c.n[2] * S * blk + 32 * nch * blk * S
    solver = 16 * npw * (6 * blk + 4 * q.nb) + 16 * blk * blk * 4
    return int(1.25 * (engine + solver)) + (64 << 20)


def run_query(q: MassQuery, prec="mixed", device=0, wl=None, budget=None) -> MassAnswer:
    """One query, end to end, on the caller's current stream."""
    t0 = time.perf_counter()
    wl = wl or build_workload(q)
    budget = budget if budget is not None else estimate_bytes(q, wl, prec)
    H = Hamiltonian(wl.cell, device=device, hbm_budget=budget)
    band = q.band if q.band >= 0 else q.nb // 2
    inv, en, npws = [], [], []
    try:
        wdt = torch.complex64 if prec == "mixed" else torch.complex128
        tol = 2e-5 if prec == "mixed" else 1e-8
        for ik, kp in enumerate(wl.kpts):
            H.load_kpoint(ik, kp)
            kin = torch.from_numpy(kp.kin).to(f"cuda:{device}")
            nb = min(q.nb, kp.npw)
            X0 = eigensolver.random_start(nb, kp.npw, kin, seed=q.seed + ik, device=f"cuda:{device}")
            op = H.at_k(ik)
            res = eigensolver.lobpcg(lambda V, HV: op(V, HV, prec=prec, overwrite=True), X0, kin,
                                     nconv=min(nb, band + 2), tol=tol, maxiter=400, work_dtype=wdt)
            m = effmass.effective_mass(H, ik, kp, wl.cell.A, res.X.contiguous(), res.E, [band], prec=prec)
            inv.append(m.inv_mass[0])
            en.append(float(res.E[band]))
            npws.append(kp.npw)
            H.release_kpoint(ik)
        torch.cuda.current_stream(device).synchronize()
    finally:
        H.close()
    t1 = time.perf_counter()
    return MassAnswer(qid=q.qid, inv_mass=inv, energies=en, latency=t1 - t0, start=t0, finish=t1, npw=npws)


class MassServer:
    """Admits queries by HBM footprint into at most `slots` concurrent
    workers, each with its own HIP stream."""

    def __init__(self, device=0, hbm_budget=None, slots=4, prec="mixed"):
        self.device = device
        self.prec = prec
        self.slots = slots
        if hbm_budget is None:
            free, total = torch.cuda.mem_get_info(device)
            hbm_budget = int(0.8 * free)
        self.budget = hbm_budget
        self._lock = threading.Condition()
        self._in_use = 0
        self._running = 0
        self._pending = queue.Queue()
        self._answers = {}
        self._threads = [threading.Thread(target=self._worker, daemon=True) for _ in range(slots)]
        self._streams = [torch.cuda.Stream(device=device) for _ in range(slots)]
        self._stop = False
        self.max_concurrency = 0
        self.peak_bytes = 0
        for i, t in enumerate(self._threads):
            t.name = f"mass-worker-{i}"
            t.start()

    def submit(self, q: MassQuery, t_arrival: float):
        wl = build_workload(q)
        need = estimate_bytes(q, wl, self.prec)
        if need > self.budget:
            self._answers[q.qid] = MassAnswer(qid=q.qid, inv_mass=[], energies=[], latency=0.0, start=t_arrival,
                                              finish=t_arrival, error="query exceeds the server's HBM budget")
            return
        self._pending.put((q, wl, need, t_arrival))

    def _worker(self):
        idx = int(threading.current_thread().name.rsplit("-", 1)[1])
        stream = self._streams[idx]
        while True:
            item = self._pending.get()
            if item is None:
                return
            q, wl, need, t_arr = item
            with self._lock:
                # FIFO admission: wait until the footprint fits next to the running queries
                while self._in_use + need > self.budget:
                    self._lock.wait()
                self._in_use += need
                self._running += 1
                self.max_concurrency = max(self.max_concurrency, self._running)
                self.peak_bytes = max(self.peak_bytes, self._in_use)
            try:
                with torch.cuda.stream(stream):
                    ans = run_query(q, self.prec, self.device, wl, budget=need)
                ans.latency = ans.finish - t_arr
            except Exception as e:      # surfaced in the answer, the server keeps running
                now = time.perf_counter()
                ans = MassAnswer(qid=q.qid, inv_mass=[], energies=[], latency=now - t_arr, start=t_arr,
                                 finish=now, error=f"{type(e).__name__}: {e}")
            with self._lock:
                self._in_use -= need
                self._running -= 1
                self._answers[q.qid] = ans
                self._lock.notify_all()

    def close(self):
        for _ in self._threads:
            self._pending.put(None)
        for t in self._threads:
            t.join()

    def answers(self):
        return dict(self._answers)


def serve(queries, device=0, slots=4, prec="mixed", hbm_budget=None, realtime=True):
    """Replay `queries` at their arrival times (realtime) or all at once;
    return (answers by qid, stats)."""
    srv = MassServer(device=device, hbm_budget=hbm_budget, slots=slots, prec=prec)
    t0 = time.perf_counter()
    for q in sorted(queries, key=lambda x: x.arrival):
        if realtime:
            dt = t0 + q.arrival - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
        srv.submit(q, time.perf_counter())
    while len(srv.answers()) < len(queries):
        time.sleep(0.01)
    srv.close()
    t1 = time.perf_counter()
    ans = srv.answers()
    lat = np.array([a.latency for a in ans.values() if a.error is None])
    stats = {"queries": len(queries), "ok": int(len(lat)), "errors": len(queries) - int(len(lat)),
             "wall_s": t1 - t0, "queries_per_s": len(lat) / (t1 - t0) if t1 > t0 else 0.0,
             "p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
             "p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
             "max_concurrency": srv.max_concurrency, "peak_admitted_bytes": srv.peak_bytes,
             "slots": slots, "budget_bytes": srv.budget}
    return ans, stats
