"""Serving loop for many forward rupture models (continuous batching).

A dynamic inversion evaluates many forward models of FD3D_TSN, each one a
run of the reference's whole program (dynamicsolver.f90:66 -> fd3d()).
Here requests of different grid sizes, friction laws and fault fields
arrive as a stream and share one GPU:

* a request is admitted when a batch slot is free and its device
  footprint (fd3d_footprint) fits in free HBM (fd3d_device_memory) minus a
  reserve; its wavefield then stays resident until it finishes;
* admitted problems are advanced together by a :class:`capi.Batch` (one
  launch per stage covering every problem, the step replayed from a HIP
  graph), in chunks of steps no longer than the shortest remaining run;
* while a chunk runs on the GPU, the host prepares the next requests
  (inputs, device allocation, upload, prologue) on their own streams, for
  every slot that is free or is freed by a problem finishing with that
  chunk; they join at the next chunk boundary, when finished problems
  leave; the results of the problems that left are read, and their
  memory freed, while the following chunk runs;
* each result (the final 2D arrays of fd3d(), fd3d_theo.f90:604-658, and
  on request the inversion's patch slip rates and moment rate,
  fd3d_theo.f90:560-578) is bit-identical to the same request solved alone.

:func:`run_plain` serves the same stream back to back, one request at a
time through fd3d_step, as the baseline.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import autotune
from . import capi
from . import cases as C


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the stream
    case: C.Case
    # inversion products to accumulate on the device: (NL, NW, dL, dW, dtseis)
    # for fd3d_set_source_bins, or None for the final arrays only
    bins: Optional[tuple] = None

    @property
    def steps(self) -> int:
        return self.case.ntfd


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    steps: int
    cells: int
    final: Dict[str, np.ndarray] = field(repr=False, default_factory=dict)
    bins: Dict[str, np.ndarray] = field(repr=False, default_factory=dict)
    # per-step slabs and stations of the last steps (Server(keep_outputs=True))
    outputs: Dict[str, np.ndarray] = field(repr=False, default_factory=dict)
    # why the request failed (FD3D_E_NEWTON, FD3D_E_NONFINITE, ...), or None
    error: Optional[str] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def c5_requests(n: int = 2000, rate: float = 20.0, seed: int = 5) -> List[Request]:
    """Poisson stream of rupture requests (mean `rate` per second): 70% slip
    weakening, 30% rate-and-state; interior grids from 64^3 to 384x192x192,
    dh 100-200 m, 800-3000 steps, dt at CFL 0.24; per-request seeded
    prestress and friction fields.  Chosen where the workload leaves it
    open: 10-node PML; slip weakening in a full space with the batched
    workload's fault recipe; rate-and-state with a free surface (the TPV104
    layout) and the nucleation at a seeded point of the fault box."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for q in range(n):
        t += float(rng.exponential(1.0 / rate))
        fvw = bool(rng.random() < 0.3)
        nx = int(rng.integers(64, 385))
        ny, nz = (int(v) for v in rng.integers(64, 193, size=2))
        dh = float(round(rng.uniform(100.0, 200.0)))
        steps = int(rng.integers(800, 3001))
        dt = float(np.float32(0.24 * dh / 6000.0))
        s = seed * 100000 + q
        flen, fwid = 0.8 * nx * dh, 0.8 * nz * dh
        if fvw:
            case = C.Case(f"req{q:04d}_fvw_{nx}x{ny}x{nz}", nx, ny, nz, dh=dh, dt=dt, ntfd=steps, nabc=10,
                          fspace=False, friction="fvw", seed=s, fault_len=flen, fault_wid=fwid,
                          trans=min(3e3, 0.1 * flen), nuc_radius=min(3e3, 0.25 * fwid))
        else:
            mus = float(rng.uniform(0.6, 0.75))
            dc = float(rng.uniform(0.2, 0.8))
            fx, fz = rng.uniform(-0.25, 0.25, size=2)
            sn, mud = 120e6, 0.525
            xc, zc = 0.5 * (nx - 1) * dh, 0.5 * (nz - 1) * dh
            case = C.Case(f"req{q:04d}_sw_{nx}x{ny}x{nz}", nx, ny, nz, dh=dh, dt=dt, ntfd=steps, nabc=10,
                          fspace=True, seed=s, fault_len=flen, fault_wid=fwid, mus=mus, mud=mud, dc=dc, sn=sn,
                          t0=sn * (mud + 0.6 * (mus - mud)), t0_std=2e6, t0_corr=2e3,
                          nuc_size=min(3e3, 0.3 * fwid), nuc_stress=sn * mus + 1.5e6,
                          nuc_xc=xc + float(fx) * flen, nuc_zc=zc + float(fz) * fwid)
        out.append(Request(q, t, case))
    return out


def percentile(x, p):
    return float(np.percentile(np.asarray(x, dtype=np.float64), p)) if len(x) else float("nan")


def summary(results: List[Result], makespan: float) -> Dict[str, float]:
    failed = sum(1 for r in results if r.error)
    results = [r for r in results if not r.error]
    lat = [r.latency for r in results]
    upd = sum(r.cells * r.steps for r in results)
    return {"requests": len(results), "failed": failed, "makespan_s": makespan,
            "ruptures_per_s": len(results) / makespan if makespan > 0 else float("nan"),
            "Gcell_updates_per_s": upd / makespan / 1e9 if makespan > 0 else float("nan"),
            "latency_p50_s": percentile(lat, 50), "latency_p99_s": percentile(lat, 99),
            "latency_mean_s": float(np.mean(lat)) if lat else float("nan")}


class Server:
    """Continuous-batching server on one GPU."""

    def __init__(self, max_slots: int = 32, chunk: int = 16, tz: int = 16, mode: int = capi.STAGED,
                 reserve_bytes: int = 4 << 30, keep_final: bool = True,
                 log: Optional[Callable[[str], None]] = None, fused: bool = True,
                 keep_outputs: bool = False, solver_args: Optional[Callable[[Request], Dict]] = None,
                 prefetch: int = 0):
        self.max_slots = max_slots
        # requests prepared ahead of a free slot, so that a slot freed after a
        # chunk shorter than a preparation fills at once (on the 2000-request
        # stream 2 measured no faster than 0: 739.0 against 737.6 s)
        self.prefetch = prefetch
        # keep_outputs: each result also holds the output slabs and station
        # samples of its last min(steps, ring) steps (the device ring)
        self.keep_outputs = keep_outputs
        # extra capi.Solver arguments per request (e.g. newton_max)
        self.solver_args = solver_args or (lambda r: {})
        self.fused = fused and mode == capi.STAGED   # members ask for the fused sweep
        self.chunk = chunk
        self.reserve = reserve_bytes
        self.keep_final = keep_final
        self.tz = tz
        self.batch = capi.Batch(max_slots, tz=tz, mode=mode)
        self.log = log or (lambda m: None)
        self.gpu_busy_s = 0.0      # host time spent waiting for chunks
        self.gpu_step_s = 0.0      # GPU time of the batch steps (HIP events)
        self.chunks = 0
        # host time by phase: queueing a chunk (with the batch's graph
        # re-capture after a membership change), preparing admissions,
        # reading the results of finished members and freeing them (while
        # the next chunk runs; freeing device memory may wait for it)
        self.host_s = {"submit": 0.0, "admit": 0.0, "retire": 0.0}

    def close(self):
        self.batch.close()

    def _fits(self, req: Request, pending_bytes: int) -> bool:
        free, _total = capi.device_memory()
        return capi.footprint(req.case) + pending_bytes + self.reserve <= free

    def run(self, requests: List[Request], clock=time.perf_counter) -> List[Result]:
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue: collections.deque = collections.deque()
        active: List[dict] = []
        ready: List[dict] = []          # prepared, join at the first boundary with a free slot
        leaving: List[dict] = []        # left the batch at the last boundary, read during the next chunk
        results: List[Result] = []
        t0 = clock()
        while pending or queue or active or ready or leaving:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            while ready and len(active) < self.max_slots:
                a = ready.pop(0)
                self.batch.add(a["solver"])
                active.append(a)
            n = min([self.chunk] + [a["left"] for a in active]) if active else 0
            t1 = clock()
            if n:
                self.batch.step(n)          # asynchronous: the host prepares admissions meanwhile
            t2 = clock()
            self.host_s["submit"] += t2 - t1
            # admission: a free slot (members whose run ends with this chunk
            # leave at its end) or room in the prefetch, and the footprint
            # in free HBM
            staying = sum(1 for a in active if a["left"] > n)
            while queue and staying + len(ready) < self.max_slots + self.prefetch and self._fits(queue[0], 0):
                r = queue.popleft()
                s = capi.Solver(r.case, **self.solver_args(r))
                if r.bins is not None:
                    s.set_source_bins(*r.bins, nbins=r.steps)
                # rows per tile for this size class; the fused sweep when
                # the batch holds enough fused work
                t = autotune.batch_member(r.case, autotune.choose(r.case, self.tz), self.tz, self.fused)
                if t is not None:
                    s.set_tuning(*t)
                ready.append({"req": r, "solver": s, "left": r.steps, "admitted": clock() - t0})
            self.host_s["admit"] += clock() - t2
            if queue and not active and not ready and not leaving:
                raise RuntimeError(f"request {queue[0].rid} does not fit in device memory")
            # results of the members that left at the last boundary (they
            # are out of the batch, so the chunk in flight does not touch them)
            t3 = clock()
            for a in leaving:
                results.append(self._finish(a, lambda: clock() - t0))
                what = f"failed ({a['error']})" if a["error"] else f"done: latency {results[-1].latency:.3f} s"
                self.log(f"[serve] request {a['req'].rid} {what}, {len(active)} active, {len(queue)} queued")
            leaving = []
            self.host_s["retire"] += clock() - t3
            if n:
                t1 = clock()
                failed = self._sync()
                self.gpu_busy_s += clock() - t1
                self.gpu_step_s += self.batch.elapsed()[0] * 1e-3
                self.chunks += 1
                for a in list(active):
                    a["left"] -= n
                    if a["left"] == 0 or id(a["solver"]) in failed:
                        self.batch.remove(a["solver"])
                        a["error"] = failed.get(id(a["solver"]))
                        active.remove(a)
                        leaving.append(a)
            elif not ready and pending:
                time.sleep(max(0.0, pending[0].arrival - (clock() - t0)))
        self.makespan = clock() - t0
        return results

    def _finish(self, a: dict, now: Callable[[], float]) -> Result:
        """Read the result of a member that has left the batch and free it;
        the request is finished once its arrays are on the host."""
        s, r, err = a["solver"], a["req"], a["error"]
        try:
            fin = s.final() if self.keep_final and not err else {}
            bins = s.source_bins() if r.bins is not None and not err else {}
            outs = s.read_outputs(min(r.steps, s.ring_steps)) if self.keep_outputs and not err else {}
        finally:
            s.close()
        return Result(r.rid, r.arrival, a["admitted"], now(), r.steps, r.case.cells, fin, bins, outs, err)

    def _sync(self) -> Dict[int, str]:
        """Wait for the chunk.  A member whose rate-and-state Newton solve hit
        its cap, or whose outputs went non-finite, fails alone: the batch's
        error names a member, each member's own flags (fd3d_sync) say which,
        and the others go on (the members do not share state).  Returns
        {id(solver): error} of the failed members."""
        try:
            self.batch.sync()
            return {}
        except capi.FD3DError as e:
            if e.code not in (-6, -7):       # FD3D_E_NEWTON, FD3D_E_NONFINITE
                raise
        failed = {}
        for s in self.batch.members:
            try:
                s.sync()
            except capi.FD3DError as e:
                if e.code not in (-6, -7):
                    raise
                failed[id(s)] = str(e)
        if not failed:
            raise RuntimeError("the batch reported a failed member, but no member's flags are set")
        return failed


def run_plain(requests: List[Request], keep_final: bool = True, clock=time.perf_counter,
              log: Optional[Callable[[str], None]] = None,
              solver_args: Optional[Callable[[Request], Dict]] = None) -> List[Result]:
    """Back-to-back plain solves in arrival order (one request at a time).
    Service times are measured; the queueing timeline of a FIFO single
    server then follows from the arrival times, so the run does not wait
    for arrivals in real time."""
    results = []
    free_at = 0.0
    for r in sorted(requests, key=lambda q: q.arrival):
        t1 = clock()
        err = None
        fin, bins = {}, {}
        with capi.Solver(r.case, **(solver_args(r) if solver_args else {})) as s:
            if r.bins is not None:
                s.set_source_bins(*r.bins, nbins=r.steps)
            autotune.apply(s)
            s.step(r.steps, outputs=False)
            try:
                s.sync()
            except capi.FD3DError as e:
                # as in Server: a Newton cap or non-finite outputs fail this
                # request alone
                if e.code not in (-6, -7):
                    raise
                err = str(e)
            if err is None:
                fin = s.final() if keep_final else {}
                bins = s.source_bins() if r.bins is not None else {}
        service = clock() - t1
        start = max(free_at, r.arrival)
        free_at = start + service
        results.append(Result(r.rid, r.arrival, start, free_at, r.steps, r.case.cells, fin, bins, error=err))
        if log:
            log(f"[plain] request {r.rid}: {service:.3f} s" + (f", failed ({err})" if err else ""))
    return results


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    """||a - b|| / ||b||; infinite when either side holds a NaN or Inf."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if not (np.all(np.isfinite(a)) and np.all(np.isfinite(b))):
        return float("inf")
    den = np.sqrt(np.sum(b ** 2))
    num = np.sqrt(np.sum((a - b) ** 2))
    return float(num / den) if den > 0 else float(num)
