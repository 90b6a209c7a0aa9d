"""A solve server: continuous batching of eigensolve requests on one device.

Requests (a system, a momentum sector, how many eigenvalues) arrive over time. The server
keeps a bounded FIFO queue; a request is admitted when the HBM budget of the server's context
holds its basis, term table and Krylov workspace, and its problem then joins the running
batch (:class:`momentumed.batch.BatchSolver`) between two Lanczos steps while the problems
already on the device keep going. Finished problems are collected after every step and their
device memory returns to the budget. Host preparation of queued requests (term lists) runs on
a worker thread while the device steps.

Because every batched problem runs exactly its stand-alone iteration, a request's result is
bitwise the one a lone EDsolve of the same request gives, whatever it shared the device
with. The reference has no serving layer (its solves are single calls of EDsolve,
MomentumED/src/MomentumED.jl:375-480); ``mode="plain"`` serves the same queue with one such
solve at a time, the baseline the batched server is measured against.
"""
from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass, field
from fractions import Fraction

import numpy as np

from . import _lib
from .batch import BatchSolver, problem_bytes
from .core import MBOperator
from .device import Context, DeviceBasis, DeviceOperator, ROCmLinearMap, krylov_map_solve, \
    system_struct
from .preparation import EDPara, ED_scatterlist_twobody
from .workloads import CoulombGaussian, mesh

__all__ = ["Request", "served_stream", "request_operator", "SolveServer", "ServeReport",
           "MESHES", "FILLINGS"]

MESHES = ((3, 5), (4, 5), (4, 6), (5, 6), (6, 6))
FILLINGS = (Fraction(1, 3), Fraction(2, 5))


@dataclass
class Request:
    rid: int
    arrival: float           # seconds after the stream starts
    mesh: tuple
    filling: Fraction
    N: int
    sector: tuple            # total momentum (k1, k2)
    screening: float         # screening length of V(q)
    nev: int = 4

    @property
    def label(self) -> str:
        return f"{self.mesh[0]}x{self.mesh[1]}@{self.filling}"


def served_stream(rate: float = 50.0, duration: float = 120.0, seed: int = 7,
                  meshes=MESHES, fillings=FILLINGS, nev: int = 4) -> list[Request]:
    """Poisson arrivals at `rate` per second for `duration` seconds. Each request draws its
    mesh and filling uniformly; a pair whose particle number Nk * filling is not an integer
    (4x5@1/3, 4x6@2/5, 6x6@2/5) is redrawn, mesh and filling together. Then a sector uniform
    over the mesh's momenta and a screening length ~ U(1, 4)."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        while True:
            m = meshes[int(rng.integers(len(meshes)))]
            nu = fillings[int(rng.integers(len(fillings)))]
            n = m[0] * m[1] * nu
            if n.denominator == 1:
                break
        k = int(rng.integers(m[0] * m[1]))
        out.append(Request(len(out), t, m, nu, int(n), (k % m[0], k // m[0]),
                           float(rng.uniform(1.0, 4.0)), nev))
    return out


def request_operator(req: Request) -> MBOperator:
    """Upper-Hermitian two-body operator of a request: spinless fermions on the mesh with
    V(q) = 2 pi / sqrt(|q|^2 + lambda^-2) * exp(-|q|^2 / 4)."""
    k = mesh(*req.mesh)
    para = EDPara(k_list=k, Gk=req.mesh,
                  H_two=CoulombGaussian(k, req.mesh, c=0.25, screening=req.screening),
                  two_momentum_coordinate=False)
    return MBOperator(ED_scatterlist_twobody(para), upper_hermitian=True)


@dataclass
class _Job:
    req: Request
    arrays: tuple | None = None          # (amp, out, in) of the upper term list
    basis: DeviceBasis | None = None
    op: DeviceOperator | None = None
    pid: int = -1
    t_admit: float = 0.0
    big: bool = False                    # counted against big_queue_limit


@dataclass
class ServeReport:
    mode: str
    offered: int = 0
    admitted: int = 0
    completed: int = 0
    rejected: int = 0
    unfinished: int = 0
    elapsed_s: float = 0.0
    latencies: list = field(default_factory=list)     # (label, seconds)
    results: dict = field(default_factory=dict)       # rid -> eigenvalues
    peak_hbm_bytes: int = 0
    steps: int = 0
    rejected_labels: list = field(default_factory=list)
    # (label, seconds since arrival when the run ended) of requests still queued or running
    unfinished_ages: list = field(default_factory=list)

    def summary(self) -> dict:
        """Counts, throughput and latency percentiles, overall and per request class.
        p50_s / p99_s are over completed requests; p50_all_s / p99_all_s count every
        unfinished request too, at the time it had waited when the run ended (a lower bound
        on its latency), so an over-capacity class cannot hide behind its completions.
        Rejected requests (queue full on arrival, or larger than the budget) are counted
        apart; they have no latency."""
        lat = np.array([s for _, s in self.latencies]) if self.latencies else np.zeros(0)
        cens = np.array([s for _, s in self.unfinished_ages]) if self.unfinished_ages else \
            np.zeros(0)
        both = np.concatenate([lat, cens])

        def pct(a, q):
            return round(float(np.percentile(a, q)), 4) if a.size else None
        by = {}
        labels = {lab for lab, _ in self.latencies} | {lab for lab, _ in self.unfinished_ages} | \
            set(self.rejected_labels)
        for lab in sorted(labels):
            a = np.array([s for l2, s in self.latencies if l2 == lab])
            u = np.array([s for l2, s in self.unfinished_ages if l2 == lab])
            ab = np.concatenate([a, u])
            by[lab] = {"completed": int(a.size), "unfinished": int(u.size),
                       "rejected": sum(1 for l2 in self.rejected_labels if l2 == lab),
                       "p50_s": pct(a, 50), "p99_s": pct(a, 99),
                       "p50_all_s": pct(ab, 50), "p99_all_s": pct(ab, 99)}
        return {
            "mode": self.mode, "offered": self.offered, "admitted": self.admitted,
            "completed": self.completed, "rejected": self.rejected,
            "unfinished": self.unfinished, "elapsed_s": round(self.elapsed_s, 3),
            "solves_per_s": round(self.completed / self.elapsed_s, 3) if self.elapsed_s else None,
            "p50_s": pct(lat, 50), "p99_s": pct(lat, 99),
            "p50_all_s": pct(both, 50), "p99_all_s": pct(both, 99),
            "max_s": round(float(lat.max()), 4) if lat.size else None,
            "by_class": by, "peak_hbm_bytes": self.peak_hbm_bytes, "steps": self.steps,
        }


class SolveServer:
    """Serve a list of timed requests, batched (default) or one solve at a time."""

    def __init__(self, hbm_budget_bytes: int = 200 * 10 ** 9, max_queue: int = 256,
                 krylovdim: int = 30, tol: float = 1e-12, maxiter: int = 100, seed: int = 0x5EED,
                 mode: str = "batched", prepare_ahead: int = 64, device: int = 0,
                 quantum_rows: int = 1 << 15, big_share: float = 0.5,
                 head_wait_s: float = 60.0, big_queue_limit: int = 128):
        """quantum_rows: the step quantum of the batch's round robin (a problem of D rows
        steps about once every D / quantum_rows steps). big_share: the share of the rows
        stepped that the oldest running problem larger than the quantum receives (its
        round-robin weight is set so), so a large request keeps moving however many small
        ones run beside it (0 turns this off). head_wait_s: once the head of the queue has
        waited this long, requests behind it are no longer admitted ahead of it, so the
        budget frees up for it. big_queue_limit: at most
        this many requests larger than the quantum wait in the queue (0: no limit beside
        max_queue); a large request arriving beyond it is rejected, so that under overload the
        queue keeps room for small ones (the large ones are bound by the device either way)."""
        if mode not in ("batched", "plain"):
            raise ValueError("mode must be 'batched' or 'plain'")
        if not 0.0 <= big_share < 1.0:
            raise ValueError("big_share must be in [0, 1)")
        self.mode = mode
        self.quantum = int(quantum_rows)
        self.big_share = float(big_share)
        self.head_wait_s = float(head_wait_s)
        self.big_queue_limit = int(big_queue_limit)
        self._weights: dict = {}
        self.ctx = Context(device, hbm_budget_bytes)
        self.budget = int(hbm_budget_bytes)
        self.max_queue = max_queue
        self.kw = dict(krylovdim=krylovdim, tol=tol, maxiter=maxiter, seed=seed)
        self.prepare_ahead = prepare_ahead
        self.batch = BatchSolver(self.ctx) if mode == "batched" else None
        if self.batch is not None:
            # small problems step every time, a problem of D rows about every D/quantum steps
            self.batch.set_quantum(quantum_rows)
        self._sizes: dict = {}

    # ------------------------------------------------------------------ pieces
    def sector_size(self, req: Request) -> int:
        key = (req.mesh, req.N)
        s = self._sizes.get(key)
        if s is None:
            import ctypes
            sysd = system_struct(mesh(*req.mesh), req.mesh, 1, 1, (req.N,))
            n = req.mesh[0] * req.mesh[1]
            s = np.zeros(n, dtype=np.int64)
            _lib.check(self.ctx.L.edm_sector_sizes(self.ctx.h, ctypes.byref(sysd), s.ctypes.data),
                       self.ctx.h)
            self._sizes[key] = s
        return int(s[req.sector[0] + req.mesh[0] * req.sector[1]])

    def estimate_bytes(self, req: Request) -> int:
        """Device bytes the request holds while it runs: Krylov workspace, basis (states and
        rank tables) and term table."""
        D = self.sector_size(req)
        nbits = req.mesh[0] * req.mesh[1]
        nhi = max(0, nbits - (nbits + 1) // 2)
        rank = (1 << ((nbits + 1) // 2)) * 6 + (1 << nhi) * 4
        if 14 < nbits <= 38:   # the split at bit 14 the LDS rank-table matvec reads
            rank += (4 << (nbits - 14)) + (2 << 14)
        elif 38 < nbits <= 50 and D >= 1 << 18:
            # the same split with J_hi in two levels: a 24-bit top table and at most 128 MB
            # of rows of 2^(nbits - 38) words, one per distinct top
            rank += (4 << 24) + (2 << 14) + min(128 << 20, D * (4 << (nbits - 38)))
        ws = problem_bytes(D, req.nev, self.kw["krylovdim"]) if self.mode == "batched" else \
            (2 * min(max(self.kw["krylovdim"], req.nev + 2), D) + 1) * D * 16
        return ws + D * 8 + rank + (4 << 20)

    def _upload(self, job: _Job):
        req = job.req
        sysd = system_struct(mesh(*req.mesh), req.mesh, 1, 1, (req.N,))
        sector = req.sector[0] + req.mesh[0] * req.sector[1]
        nbits = req.mesh[0] * req.mesh[1]
        job.basis = DeviceBasis.enumerate(self.ctx, sysd, sector)
        amp, out, inn = job.arrays
        job.op = DeviceOperator(self.ctx, nbits, amp, out, inn, True)

    def _release(self, job: _Job):
        if job.op is not None:
            job.op.close()
        if job.basis is not None:
            job.basis.close()
        job.op = job.basis = None

    # ------------------------------------------------------------------ loop
    def run(self, requests: list[Request], drain_s: float = 30.0,
            on_result=None) -> ServeReport:
        """Replay `requests` in real time (arrival offsets from now). Arrivals stop with the
        last request; the server then drains for at most drain_s seconds. A request that
        arrives to a full queue is rejected."""
        rep = ServeReport(self.mode, offered=len(requests))
        reqs = sorted(requests, key=lambda r: r.arrival)
        waiting: list[_Job] = []          # FIFO, arrived, not admitted
        running: dict[int, _Job] = {}     # pid -> job (batched)
        prep_in: queue.Queue = queue.Queue()
        prep_lock = threading.Lock()
        stop = threading.Event()

        def prep_worker():
            while not stop.is_set():
                try:
                    job = prep_in.get(timeout=0.05)
                except queue.Empty:
                    continue
                arrays = request_operator(job.req).arrays()
                with prep_lock:
                    job.arrays = arrays

        worker = threading.Thread(target=prep_worker, daemon=True)
        worker.start()
        queued_for_prep: set = set()
        nbig_waiting = 0                  # queued large requests (big_queue_limit)
        t0 = time.perf_counter()
        i = 0
        last_arrival = reqs[-1].arrival if reqs else 0.0
        try:
            while True:
                now = time.perf_counter() - t0
                while i < len(reqs) and reqs[i].arrival <= now:
                    big = self.big_queue_limit and self.sector_size(reqs[i]) > self.quantum
                    if len(waiting) >= self.max_queue or \
                            (big and nbig_waiting >= self.big_queue_limit):
                        rep.rejected += 1
                        rep.rejected_labels.append(reqs[i].label)
                    else:
                        waiting.append(_Job(reqs[i], big=bool(big)))
                        nbig_waiting += 1 if big else 0
                    i += 1
                # keep prepare_ahead requests of the queue in preparation on the worker thread,
                # front first; requests already prepared (say large ones that wait for budget
                # at the front) do not count, so those behind them get prepared too
                pending = prep_in.qsize()
                for job in waiting:
                    if pending >= self.prepare_ahead:
                        break
                    if job.req.rid not in queued_for_prep:
                        queued_for_prep.add(job.req.rid)
                        prep_in.put(job)
                        pending += 1
                if self.mode == "batched":
                    self._admit_batched(waiting, running, rep, prep_lock, now)
                    nbig_waiting = sum(1 for j in waiting if j.big)
                    if running:
                        self._rebalance(running)
                        _, nfin = self.batch.step(1)
                        rep.steps += 1
                        if nfin:
                            self._collect(running, rep, t0, on_result)
                    elif i < len(reqs) and not waiting:
                        time.sleep(min(0.005, max(0.0, reqs[i].arrival - now)))
                    elif waiting:
                        time.sleep(0.0005)          # head not prepared yet
                else:
                    if waiting and self._ready(waiting[0], prep_lock):
                        job = waiting.pop(0)
                        nbig_waiting -= 1 if job.big else 0
                        self._solve_plain(job, rep, t0, on_result)
                    elif i < len(reqs) and not waiting:
                        time.sleep(min(0.005, max(0.0, reqs[i].arrival - now)))
                    elif waiting:
                        time.sleep(0.0005)
                used, _ = self.ctx.memory()
                rep.peak_hbm_bytes = max(rep.peak_hbm_bytes, used)
                now = time.perf_counter() - t0
                if i >= len(reqs) and not waiting and not running:
                    break
                if i >= len(reqs) and now > last_arrival + drain_s:
                    break
        finally:
            stop.set()
            worker.join()
        rep.elapsed_s = time.perf_counter() - t0
        rep.unfinished = len(waiting) + len(running)
        for job in list(waiting) + list(running.values()):
            rep.unfinished_ages.append((job.req.label, rep.elapsed_s - job.req.arrival))
        for pid, job in list(running.items()):
            self.batch.remove(pid)
            self._release(job)
        return rep

    def _ready(self, job: _Job, lock) -> bool:
        with lock:
            return job.arrays is not None

    def _admit_batched(self, waiting, running, rep, lock, now=0.0):
        """FIFO admission with backfilling: when the head of the queue does not fit the free
        budget, prepared requests behind it that do fit (so are smaller) are admitted, until
        the head has waited head_wait_s; from then on nothing passes it."""
        used, _ = self.ctx.memory()
        free = self.budget - used
        k = 0
        while k < len(waiting):
            job = waiting[k]
            need = self.estimate_bytes(job.req)
            if not self._ready(job, lock):
                if k == 0 and need > free and now - job.req.arrival > self.head_wait_s:
                    break                      # owed the budget, even while in preparation
                k += 1                         # still in preparation: the ones behind may go
                continue
            if need > self.budget:
                waiting.pop(k)                 # can never fit
                rep.rejected += 1
                rep.rejected_labels.append(job.req.label)
                continue
            if need > free:
                if k == 0 and now - job.req.arrival > self.head_wait_s:
                    break                      # the head is owed the budget that frees up
                k += 1
                continue
            try:
                self._upload(job)
                job.pid = self.batch.add(
                    ROCmLinearMap(job.op, job.basis), job.req.nev, **self.kw)
            except _lib.EDMError as e:
                self._release(job)
                if e.code == _lib.EDM_ENOMEM:
                    break
                raise
            job.t_admit = time.perf_counter()
            running[job.pid] = job
            waiting.pop(k)
            rep.admitted += 1
            used, _ = self.ctx.memory()
            free = self.budget - used

    def _rebalance(self, running):
        """Round-robin weights: the oldest running problem larger than the quantum gets the
        weight that gives it big_share of the rows stepped (the rows the others step per
        step are sum(min(D, quantum))); every other problem weight 1."""
        want = {}
        if self.big_share > 0 and running:
            big = next((p for p, j in running.items() if j.basis.D > self.quantum), None)
            if big is not None:
                rest = sum(min(j.basis.D, self.quantum) for p, j in running.items() if p != big)
                f = self.big_share
                want[big] = max(1.0, f * rest / ((1.0 - f) * self.quantum))
        for pid in running:
            w = want.get(pid, 1.0)
            if self._weights.get(pid, 1.0) != w:
                self.batch.set_weight(pid, w)
                self._weights[pid] = w

    def _collect(self, running, rep, t0, on_result):
        for pid in self.batch.finished():
            job = running.pop(pid)
            self._weights.pop(pid, None)
            r = self.batch.result(pid)
            t = time.perf_counter() - t0
            self._release(job)
            rep.completed += 1
            rep.latencies.append((job.req.label, t - job.req.arrival))
            rep.results[job.req.rid] = r.values
            if on_result:
                on_result(job.req, r.values, r.info)

    def _solve_plain(self, job, rep, t0, on_result):
        try:
            self._upload(job)
            vals, _, info = krylov_map_solve(ROCmLinearMap(job.op, job.basis), job.req.nev,
                                             want_vectors=False, **self.kw)
        finally:
            self._release(job)
        rep.admitted += 1
        rep.completed += 1
        t = time.perf_counter() - t0
        rep.latencies.append((job.req.label, t - job.req.arrival))
        rep.results[job.req.rid] = vals
        if on_result:
            on_result(job.req, vals, info)

    def close(self):
        if self.batch is not None:
            self.batch.close()
        self.ctx.close()


def solve_alone(req: Request, ctx: Context, krylovdim=30, tol=1e-12, maxiter=100,
                seed=0x5EED, want_vectors: bool = False):
    """The request solved by itself (the parity reference for the served results): its
    eigenvalues, or with want_vectors (eigenvalues, eigenvectors, basis states, info)."""
    sysd = system_struct(mesh(*req.mesh), req.mesh, 1, 1, (req.N,))
    sector = req.sector[0] + req.mesh[0] * req.sector[1]
    basis = DeviceBasis.enumerate(ctx, sysd, sector)
    amp, out, inn = request_operator(req).arrays()
    op = DeviceOperator(ctx, req.mesh[0] * req.mesh[1], amp, out, inn, True)
    try:
        vals, vecs, info = krylov_map_solve(ROCmLinearMap(op, basis), req.nev,
                                            want_vectors=want_vectors, krylovdim=krylovdim,
                                            tol=tol, maxiter=maxiter, seed=seed)
        states = basis.download() if want_vectors else None
    finally:
        op.close()
        basis.close()
    return (vals, vecs, states, info) if want_vectors else vals

