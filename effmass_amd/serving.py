"""Serving loop for effective-mass queries (a stream of 2-64 atom cells).

A query is a cell of 2-64 atoms (diamond-derived, 135 bohr^3 per atom), a
cutoff, a band count nb and 1-4 k-points; its answer is the k.p
effective-mass tensor of band nb/2 at every k-point, from the sum over the
nb lowest states, every one of them converged.  The server keeps each admitted query's
data resident on the device (V_loc, projector tables and matrices, psi) for
as long as the query runs, and runs queries concurrently:

* worker processes, one per slot, each with its own HIP context and
  stream (default priority): the eigensolver's host-side Python runs in
  parallel across slots (threads would serialise on the interpreter lock)
  and the slots' kernels overlap on the device;
* admission by HBM footprint: a query starts only when a slot is free and
  its estimated device bytes fit under the server's budget next to the
  queries already running (FIFO order, so no query starves);
* continuous batching: queries that arrive while others run are admitted
  as soon as slots and memory free up;
* every query gets its own engine context with an HBM budget equal to its
  estimate, so a wrong estimate fails that query loudly instead of
  starving the others.

Results do not depend on what else runs: all kernels are deterministic
(no atomics; split-K partial sums are reduced in a fixed order), so a query
returns the same numbers alone and in the mix.
"""
from __future__ import annotations

import collections
import dataclasses
import math
import multiprocessing as mp
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


# Axis lengths with compile-time two-stage FFT plans in the engine (the
# EFFM_CT_SIZES list of effmass_amd/csrc/fft_ct.hpp).
CT_FFT_SIZES = (15, 16, 18, 20, 24, 25, 27, 32, 36, 40, 45, 48, 80, 90, 96)


def grid_for(A, ecut_ry):
    """FFT grid holding the density sphere, as PWmat's grids do (the
    reference's Si8 fixture: 36^3 at 25 Ry, a = 10.26 bohr): the wave
    functions fill |K| <= G_max = sqrt(ecut_ry) (bohr^-1), their products
    with V_loc reach 2 G_max, so n_i > 2 (2 G_max) |a_i| / 2 pi, i.e.
    n_i >= ceil(4 G_max |a_i| / 2 pi) + 1 with no wrap-around of V psi.
    Each axis takes the smallest length with a compile-time FFT plan that
    holds it (else the smallest 2,3,5-smooth length)."""
    gmax = math.sqrt(ecut_ry)
    out = []
    for a in A:
        need = int(math.ceil(4 * gmax * np.linalg.norm(a) / (2 * math.pi))) + 1
        out.append(next((m for m in CT_FFT_SIZES if m >= need), None) or smooth_size(need))
    return tuple(out)


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
    k0: int = 0                   # index of kfracs[0] in the query this one is a part of


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
    k0: int = 0                   # first k-point of the query part this answers


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


def build_workload(q: MassQuery, order: str = "plane"):
    """The query's cell, grid and k-points.  The sphere is stored plane by
    plane (synth.g_sphere order "plane"), so that each plane's scatter and
    gather touch one contiguous run of every band rather than lines all
    over it as in the |K|^2 order; the masses do not depend on the order."""
    A, tau = query_cell(q.natom)
    n = grid_for(A, q.ecut_ry)
    return synth.custom(A, tau, n, q.ecut_ry, q.kfracs, q.nb, q.seed, with_psi=False, order=order)


def buffer_rows(nb: int) -> int:
    """Extra eigenvector rows the solver carries beyond the nb states of the
    k.p sum: they speed up the convergence of the top wanted states and are
    left out of the sum."""
    return max(16, nb // 4)


def estimate_bytes(q: MassQuery, wl=None, prec="mixed"):
    """Device bytes a query holds while it runs: engine data (V_loc, tables,
    projectors and their k-gradients at the k-point and at the shifted
    k-point of the effective-mass step, FFT workspace for a 3nb block,
    split-K and overlap partials) plus the solver's blocks (FP64 X, HX and
    the filter / search directions)."""
    if wl is not None:
        c = wl.cell
        A, n, nch = c.A, c.n, c.nchannels()
        npw = max(k.npw for k in wl.kpts)
    else:
        # without building the workload: the sphere holds about
        # Omega |K|max^3 / (6 pi^2) plane waves (|K|max^2 = ecut in Ry), +10%
        A, _ = query_cell(q.natom)
        n = grid_for(A, q.ecut_ry)
        nch = 8 * q.natom
        omega = abs(np.linalg.det(A))
        npw = int(1.1 * omega * q.ecut_ry ** 1.5 / (6 * math.pi ** 2)) + 100
    N = int(np.prod(n))
    S = 8 if prec == "mixed" else 16
    blk = 3 * (q.nb + buffer_rows(q.nb))
    planes = min(n[0], int(2 * math.sqrt(q.ecut_ry) * np.linalg.norm(A[0]) / (2 * math.pi)) + 2)
    # projectors P and their k-gradients dP (4 nch x npw) for the k-point and
    # for the shifted copy the effective-mass step loads next to it
    engine = N * 12 + 2 * 4 * nch * npw * S + planes * n[1] * n[2] * S * blk + 32 * nch * blk * S
    solver = 16 * npw * (6 * blk + 4 * q.nb) + 16 * blk * blk * 4
    return int(1.25 * (engine + solver)) + (64 << 20) + (32 << 20)   # + overlap partial sums


def run_query(q: MassQuery, prec="mixed", device=0, wl=None, budget=None, stats=None,
              rr_host: bool = True, tol: float | None = None) -> MassAnswer:
    """One query, end to end, on the caller's current stream.  With a
    `stats` dict: H apply calls and band applies, eigensolver iterations,
    setup / solve / mass seconds and (per-kernel HIP-event timing on) the
    device kernel milliseconds are added to it.  rr_host: the Rayleigh-Ritz
    eigenproblems on the host (LAPACK; False: on the device).  tol: the
    residual norm every state of the sum converges to (Ha; default 2e-5
    mixed, 1e-8 FP64)."""
    t0 = time.perf_counter()
    wl = wl or build_workload(q)
    budget = budget if budget is not None else estimate_bytes(q, wl, prec)
    H = Hamiltonian(wl.cell, device=device, hbm_budget=budget)
    if stats is not None:
        H.set_timing(True)
        for key in ("apply_calls", "band_applies", "iters", "setup_s", "solve_s", "mass_s", "kernel_ms"):
            stats.setdefault(key, 0)
        stats["setup_s"] += time.perf_counter() - t0
    band = q.band if q.band >= 0 else q.nb // 2
    # the small Rayleigh-Ritz eigenproblems (nb + buffer rows square, FP64):
    # on the host by default (LAPACK, 2 threads per slot); rr_host=False runs
    # them on the device (rocSOLVER through torch, eigenvalues equal to
    # 4e-14).  Measured with 8 slots: 100 queries at 4/s p50 0.074-0.077 s,
    # p99 0.48-0.51 s on the host against 0.105-0.131 s / 0.71-0.72 s on the
    # device; at saturation (150 queries at once) 9.2-9.8 queries/s either way
    rr_eigh = eigensolver.eigh_host if rr_host else None
    inv, en, npws = [], [], []
    try:
        wdt = torch.complex64 if prec == "mixed" else torch.complex128
        if tol is None:
            tol = 2e-5 if prec == "mixed" else 1e-8
        for ik, kp in enumerate(wl.kpts):
            t1 = time.perf_counter()
            H.load_kpoint(ik, kp)
            kin = torch.from_numpy(kp.kin).to(f"cuda:{device}")
            nb = min(q.nb, kp.npw)
            rows = min(nb + buffer_rows(nb), kp.npw)
            op = H.at_k(ik)

            def apply(V, HV):
                if stats is not None:
                    stats["apply_calls"] += 1
                    stats["band_applies"] += V.shape[0]
                op(V, HV, prec=prec, overwrite=True)
            X0 = eigensolver.plane_wave_start(apply, kin, rows, work_dtype=wdt, eigh=rr_eigh)
            # Chebyshev-filtered subspace iteration: H applies are cheap next
            # to the dense eigenproblems (measured 1.3-5x faster than LOBPCG
            # on the serving query mix, same eigenvalues to 1e-7 Ha).  Every
            # state of the k.p sum is converged (the sum couples band n to
            # all nb states: with only the states up to n converged, m* moved
            # by up to 2e-2 relative with the start vectors; converged, it
            # agrees with an FP64 solve to 1e-6-1e-4); buffer rows beyond nb
            # are filtered with the largest degree a wanted row still needs.
            # Measured on 8-64 atom cells (128 bands, two k-points each): the
            # plane-wave start (a Rayleigh-Ritz step in the 1.25 rows
            # lowest-kinetic plane waves) needs 25-45% fewer band applies
            # than a random start, with 3-4 instead of 5-7 filter steps; a
            # buffer of nb/4 rows saves 7-20% of the kernel time against nb/8.
            res = eigensolver.chfsi(apply, X0, kin,
                                    nconv=nb, tol=tol, maxiter=200, work_dtype=wdt,
                                    gram=H.overlap, degree=30, eigh=rr_eigh, lincomb=H.lincomb,
                                    buffer_degree="wanted")
            if stats is not None:
                torch.cuda.current_stream(device).synchronize()
                t2 = time.perf_counter()
                stats["solve_s"] += t2 - t1
                stats["iters"] += res.iters
            m = effmass.effective_mass(H, ik, kp, wl.cell.A, res.X[:nb].contiguous(), res.E[:nb], [band],
                                       prec=prec)
            inv.append(m.inv_mass[0])
            en.append(float(res.E[band]))
            npws.append(kp.npw)
            H.release_kpoint(ik)
            if stats is not None:
                stats["mass_s"] += time.perf_counter() - t2
        torch.cuda.current_stream(device).synchronize()
        if stats is not None:
            for name in ("plane_g2r", "xpencil_vloc", "plane_r2g", "proj_gemm", "proj_reduce", "proj_apply",
                         "proj_build", "overlap", "kp_sum"):
                stats["kernel_ms"] += H.timing(name)[0]
    finally:
        H.close()
    t1 = time.perf_counter()
    return MassAnswer(qid=q.qid, inv_mass=inv, energies=en, latency=t1 - t0, start=t0, finish=t1, npw=npws,
                      k0=q.k0)


def class_breakdown(prec="mixed", device=0, nb=128, ecut_ry=40.0, reps=2, sizes=SIZES):
    """Per size class (natom), one query of one k-point alone on the device,
    `reps` times: wall seconds, its solve / mass split, H apply calls, band
    applies, eigensolver iterations, summed kernel milliseconds (HIP events
    around each kernel class) and the device idle fraction 1 - kernel / wall
    of the query.  The cost model of the serving stream."""
    out = {}
    rng = np.random.default_rng(7)
    # untimed: the solver libraries' first calls at this block size
    run_query(MassQuery(qid=-1, natom=sizes[0], ecut_ry=ecut_ry, nb=nb, kfracs=[np.zeros(3)], seed=99), prec, device)
    for natom in sizes:
        rows = []
        for r in range(reps):
            q = MassQuery(qid=r, natom=natom, ecut_ry=ecut_ry, nb=nb, kfracs=[rng.uniform(-0.5, 0.5, 3)],
                          seed=100 + r)
            st = {}
            t0 = time.perf_counter()
            run_query(q, prec, device, stats=st)
            st["wall_s"] = time.perf_counter() - t0
            rows.append(st)
        avg = {k: float(np.mean([row[k] for row in rows])) for k in rows[0]}
        avg["device_idle"] = max(0.0, 1.0 - avg["kernel_ms"] * 1e-3 / avg["wall_s"])
        out[str(natom)] = avg
    return out


# hipDeviceSchedule* flags of hipSetDeviceFlags: how a host thread waits in a
# HIP synchronisation
_SYNC_FLAGS = {"spin": 1, "yield": 2, "blocking": 4}


def _set_sync_mode(device, sync):
    """Sets how this process's host threads wait for the device: "spin"
    (busy-wait), "yield" or "blocking" (sleep until the device signals).
    Must run before anything in the process touches the device."""
    if sync == "auto":
        return
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipSetDevice(ctypes.c_int(device))
    if rc == 0:
        rc = hip.hipSetDeviceFlags(ctypes.c_uint(_SYNC_FLAGS[sync]))
    if rc != 0:
        raise RuntimeError(f"hipSetDeviceFlags({sync}) failed with HIP error {rc}")


def _worker_main(slot, device, prec, tasks, results, host_threads=2, sync="auto", rr_host=True):
    """A serving slot: its own process, HIP context and stream; its host
    LAPACK work (the small Rayleigh-Ritz eigenproblems) uses host_threads
    threads so that the slots share the host cores; `sync` sets how its
    threads wait for the device (_set_sync_mode)."""
    try:
        _set_sync_mode(device, sync)
        import threadpoolctl
        threadpoolctl.threadpool_limits(host_threads)
        torch.set_num_threads(host_threads)
        torch.cuda.set_device(device)
        run_query(MassQuery(qid=-1, natom=2, ecut_ry=30.0, nb=16, kfracs=[np.zeros(3)], seed=1),
                  prec, device)                                  # code object loading
        torch.cuda.empty_cache()
    except Exception as e:                                       # reported, the server stops
        results.put(("failed", slot, f"{type(e).__name__}: {e}"))
        return
    results.put(("ready", slot, None))
    while True:
        item = tasks.get()
        if item is None:
            return
        q, need, t_arr = item
        try:
            ans = run_query(q, prec, device, budget=need, rr_host=rr_host)
            ans.latency = ans.finish - t_arr
        except Exception as e:                                   # surfaced in the answer
            now = time.perf_counter()
            ans = MassAnswer(qid=q.qid, inv_mass=[], energies=[], latency=now - t_arr, start=t_arr,
                             finish=now, error=f"{type(e).__name__}: {e}", k0=q.k0)
        # torch keeps the query's blocks cached for the next query of this
        # slot (no device-wide free between queries)
        results.put(("done", slot, ans))


class MassServer:
    """Admits queries by HBM footprint into at most `slots` concurrent
    worker processes (at most 15, next to the caller's own process).
    A query with several k-points is admitted as one part per k-point (the
    k-points are independent eigenproblems on the same cell), so idle slots
    shorten its latency; the parts' results are merged in k order.  Each
    slot's host-side LAPACK and torch work uses host_threads threads.
    Timestamps are time.perf_counter() (CLOCK_MONOTONIC, the same clock in
    every process)."""

    MAX_SLOTS = 15

    def __init__(self, device=0, hbm_budget=None, slots=4, prec="mixed", start_timeout=600.0,
                 split_kpoints=True, host_threads=2, sync="auto", rr_host=True):
        if not 1 <= slots <= self.MAX_SLOTS:
            raise ValueError(f"slots must be 1..{self.MAX_SLOTS}")
        if sync not in ("auto",) + tuple(_SYNC_FLAGS):
            raise ValueError(f"sync must be one of auto, {', '.join(_SYNC_FLAGS)}")
        self.device = device
        self.prec = prec
        self.slots = slots
        self.split_kpoints = split_kpoints
        self._parts = {}                  # qid -> [parts expected, {k0: answer}, arrival]
        ctx = mp.get_context("spawn")
        self._tasks = [ctx.Queue() for _ in range(slots)]
        self._results = ctx.Queue()
        self._procs = [ctx.Process(target=_worker_main,
                                   args=(i, device, prec, self._tasks[i], self._results, host_threads, sync, rr_host),
                                   daemon=True, name=f"mass-slot-{i}") for i in range(slots)]
        for p in self._procs:
            p.start()
        ready = 0
        t_end = time.monotonic() + start_timeout
        while ready < slots:
            try:
                kind, slot, msg = self._results.get(timeout=5.0)
            except Exception:
                dead = [p.name for p in self._procs if not p.is_alive()]
                if dead or time.monotonic() > t_end:
                    self._kill()
                    raise RuntimeError("serving slots did not start" + (f": {dead} exited" if dead else " in time")
                                       + " (a script that starts a server needs an "
                                         "if __name__ == '__main__' guard: slots are spawned processes)")
                continue
            if kind == "failed":
                self._kill()
                raise RuntimeError(f"serving slot {slot} failed to start: {msg}")
            ready += 1
        if hbm_budget is None:
            free, total = torch.cuda.mem_get_info(device)
            hbm_budget = int(0.8 * free)
        self.budget = hbm_budget
        self._lock = threading.Condition()
        self._pending = collections.deque()
        self._free = list(range(slots))
        self._need = {}
        self._inflight = {}
        self._in_use = 0
        self._running = 0
        self._answers = {}
        self.max_concurrency = 0
        self.peak_bytes = 0
        self._collector = threading.Thread(target=self._collect, daemon=True, name="mass-collector")
        self._collector.start()

    def submit(self, q: MassQuery, t_arrival: float):
        need = estimate_bytes(q, None, self.prec)
        if need > self.budget:
            with self._lock:
                self._answers[q.qid] = MassAnswer(qid=q.qid, inv_mass=[], energies=[], latency=0.0,
                                                  start=t_arrival, finish=t_arrival,
                                                  error="query exceeds the server's HBM budget")
            return
        with self._lock:
            nk = len(q.kfracs)
            if self.split_kpoints and nk > 1:
                self._parts[q.qid] = [nk, {}, t_arrival]
                for ik in range(nk):
                    part = dataclasses.replace(q, kfracs=[q.kfracs[ik]], k0=q.k0 + ik)
                    self._pending.append((part, need, t_arrival))
            else:
                self._pending.append((q, need, t_arrival))
            self._admit()

    def _answer(self, ans: MassAnswer):
        """Record an answer (lock held); parts of a split query are merged
        once all of them are in."""
        rec = self._parts.get(ans.qid)
        if rec is None:
            self._answers[ans.qid] = ans
            return
        nk, got, t_arr = rec
        got[ans.k0] = ans
        if len(got) < nk:
            return
        del self._parts[ans.qid]
        parts = [got[k] for k in sorted(got)]
        errs = [a.error for a in parts if a.error]
        finish = max(a.finish for a in parts)
        merged = MassAnswer(qid=ans.qid, inv_mass=[m for a in parts for m in a.inv_mass],
                            energies=[e for a in parts for e in a.energies], latency=finish - t_arr,
                            start=min(a.start for a in parts), finish=finish,
                            npw=[n for a in parts for n in a.npw], error=errs[0] if errs else None,
                            k0=parts[0].k0)
        if errs:
            merged.inv_mass, merged.energies = [], []
        self._answers[ans.qid] = merged

    def _admit(self):
        # FIFO: the head waits until a slot is free and its footprint fits
        while self._pending and self._free and self._in_use + self._pending[0][1] <= self.budget:
            q, need, t_arr = self._pending.popleft()
            slot = self._free.pop()
            self._need[slot] = need
            self._inflight[slot] = q
            self._in_use += need
            self._running += 1
            self.max_concurrency = max(self.max_concurrency, self._running)
            self.peak_bytes = max(self.peak_bytes, self._in_use)
            self._tasks[slot].put((q, need, t_arr))

    def _collect(self):
        while True:
            try:
                kind, slot, ans = self._results.get(timeout=5.0)
            except Exception:
                self._reap_dead()
                continue
            if kind == "stop":
                return
            with self._lock:
                self._in_use -= self._need.pop(slot)
                self._inflight.pop(slot, None)
                self._running -= 1
                self._free.append(slot)
                self._answer(ans)
                self._admit()
                self._lock.notify_all()

    def _reap_dead(self):
        """A slot whose process died answers its query with an error and is
        not used again."""
        with self._lock:
            for slot, p in enumerate(self._procs):
                if p.is_alive() or slot not in self._need:
                    continue
                self._in_use -= self._need.pop(slot)
                self._running -= 1
                q = self._inflight.pop(slot, None)
                if q is not None:
                    now = time.perf_counter()
                    self._answer(MassAnswer(qid=q.qid, inv_mass=[], energies=[], latency=0.0, start=now,
                                            finish=now, error=f"serving slot {slot} exited (code {p.exitcode})",
                                            k0=q.k0))
                self._lock.notify_all()

    def wait(self, n, timeout=None, progress=None):
        """Block until n answers are in; progress(done, n) is called about
        once a second.  On timeout the error names the slots still busy."""
        t_end = None if timeout is None else time.monotonic() + timeout
        with self._lock:
            while len(self._answers) < n:
                left = None if t_end is None else t_end - time.monotonic()
                if left is not None and left <= 0:
                    dead = [i for i, p in enumerate(self._procs) if not p.is_alive()]
                    raise TimeoutError(f"serving did not finish in time: {len(self._answers)}/{n} answered, "
                                       f"{self._running} running, {len(self._pending)} pending, "
                                       f"slots not alive: {dead}")
                self._lock.wait(timeout=min(1.0, left) if left is not None else 1.0)
                if progress is not None:
                    progress(len(self._answers), n)

    def _kill(self):
        for p in self._procs:
            if p.is_alive():
                p.terminate()
        for p in self._procs:
            p.join(timeout=10)

    def close(self):
        for t in self._tasks:
            t.put(None)
        for p in self._procs:
            p.join(timeout=60)
        self._kill()
        self._results.put(("stop", -1, None))
        self._collector.join(timeout=10)

    def answers(self):
        with self._lock:
            return dict(self._answers)


def serve(queries, device=0, slots=4, prec="mixed", hbm_budget=None, realtime=True, timeout=None, progress=None,
          split_kpoints=True, host_threads=2, sync="auto", rr_host=True):
    """Replay `queries` at their arrival times (realtime) or all at once;
    return (answers by qid, stats).  progress(done, total) is called about
    once a second while waiting."""
    srv = MassServer(device=device, hbm_budget=hbm_budget, slots=slots, prec=prec, split_kpoints=split_kpoints,
                     host_threads=host_threads, sync=sync, rr_host=rr_host)
    try:
        t0 = time.perf_counter()
        for q in sorted(queries, key=lambda x: x.arrival):
            if realtime:
                dt = t0 + q.arrival - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
            srv.submit(q, time.perf_counter())
        srv.wait(len(queries), timeout=timeout, progress=progress)
        t1 = time.perf_counter()
    finally:
        srv.close()
    ans = srv.answers()
    lat = np.array([a.latency for a in ans.values() if a.error is None])
    stats = {"queries": len(queries), "ok": int(len(lat)), "errors": len(queries) - int(len(lat)),
             "wall_s": t1 - t0, "queries_per_s": len(lat) / (t1 - t0) if t1 > t0 else 0.0,
             "p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
             "p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
             "max_concurrency": srv.max_concurrency, "peak_admitted_bytes": srv.peak_bytes,
             "slots": slots, "budget_bytes": srv.budget}
    return ans, stats
