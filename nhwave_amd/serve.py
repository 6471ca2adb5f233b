"""Serving loop: scenario runs arrive as requests, are admitted into a
running batch while their device footprint fits the HBM budget, advance
together with the runs already on the GPU (one launch per kernel for all of
them), and leave the batch when they reach their step count, which frees
their memory for the next arrivals (continuous batching).  Reports
time-to-result percentiles beside throughput.

Size classes (optional) keep small runs from waiting on large ones: each
class is its own batch on its own stream, stepped by its own host thread.

The reference runs one scenario per process to its SIM_STEPS
(src/nhwave.F:119-300); a request here is the same run: a scenario, a step
count, and the fields that the run returns at its end.
"""
from __future__ import annotations

import ctypes as C
import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .capi import check
from .model import Model, make_config
from .scenario import Scenario, bumps_beach


@dataclass
class Request:
    rid: int
    scenario: Scenario
    steps: int
    arrival: float                 # seconds after the server starts
    fields: tuple = ("Eta", "U", "V", "W", "Mask")
    total_time: float = math.inf   # TOTAL_TIME: the run also ends when TIME reaches it
    footprint: int = 0
    admitted: float = math.nan
    finished: float = math.nan
    result: dict = field(default_factory=dict)
    stats: dict = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def cell_stages(self) -> int:
        s = self.scenario
        return s.mglob * s.nglob * s.kglob * s.time_order * self.steps

    @classmethod
    def from_run_dir(cls, rid: int, run_dir, arrival: float = 0.0, steps: int | None = None,
                     fields: tuple | None = None) -> "Request":
        """A request for a reference run directory (input.txt and the
        depth/eta0/uvw0 files it names), run to its SIM_STEPS unless `steps`
        says otherwise, or until TIME reaches its TOTAL_TIME, as the
        reference's loop ends (src/nhwave.F:119); ingest.InputError for a
        directory off this path."""
        from .ingest import scenario_from_run_dir
        s = scenario_from_run_dir(run_dir)
        r = cls(rid, s, s.steps if steps is None else steps, arrival, total_time=s.total_time)
        if fields is not None:
            r.fields = tuple(fields)
        return r


class _Lane:
    """One size class: its own batch (and so its own HIP stream), its active
    runs, and the host thread that steps it."""

    def __init__(self, srv: "Server", bound: float):
        self.srv = srv
        self.bound = bound
        self.h = C.c_void_p()
        check(srv.lib.nhw_batch_create(None, 0, C.byref(self.h)), "nhw_batch_create")
        self.active: list[tuple[Request, Model]] = []
        self.error: BaseException | None = None
        self.steps_run = 0

    def _admit(self, r: Request, now: float) -> None:
        srv = self.srv
        cfg = make_config(r.scenario, srv.precision)
        ctx = C.c_void_p()
        check(srv.lib.nhw_batch_add(self.h, C.byref(cfg), C.byref(ctx)), "nhw_batch_add")
        m = Model(r.scenario, handle=ctx)
        r.admitted = now
        srv.used += r.footprint
        self.active.append((r, m))
        srv.n_active += 1
        srv.peak_used = max(srv.peak_used, srv.used)
        srv.peak_active = max(srv.peak_active, srv.n_active)

    def _retire(self, idx: int, now: float) -> None:
        srv = self.srv
        r, m = self.active[idx]
        r.result = m.download(*r.fields)
        st = m.stats()
        r.stats = dict(solves=st.solves, iters=st.total_iters, not_converged=st.not_converged,
                       last_relres=st.last_relres, time=m.time())
        m.ctx = C.c_void_p()  # the batch owns the context
        check(srv.lib.nhw_batch_remove(self.h, idx), "nhw_batch_remove")
        srv.used -= r.footprint
        srv.n_active -= 1
        r.finished = now
        srv.done.append(r)
        del self.active[idx]

    def run(self, pending: list[Request], t0: float) -> None:
        try:
            self._run(pending, t0)
        except BaseException as e:  # re-raised by Server.run
            self.error = e

    def _run(self, pending: list[Request], t0: float) -> None:
        srv = self.srv
        progress: dict[int, int] = {}
        dirty = False
        fresh = False
        while pending or self.active:
            now = time.perf_counter() - t0
            # admit, in arrival order, while slots and memory allow (shared
            # by every class); allocation, release and graph capture hold
            # the server lock, stepping does not
            with srv.lock:
                while (pending and pending[0].arrival <= now and srv.n_active < srv.max_slots
                       and srv.used + pending[0].footprint <= srv.hbm_budget):
                    r = pending.pop(0)
                    self._admit(r, now)
                    progress[r.rid] = 0
                    dirty = True
                if self.active and dirty:
                    check(srv.lib.nhw_batch_prepare(self.h, 0), "nhw_batch_prepare")
                    dirty = False
                    fresh = True
            if not self.active:
                # idle until the next arrival (or until another class frees
                # memory or a slot)
                time.sleep(min(max(pending[0].arrival - now, 0.001), 0.01))
                continue
            n = min([srv.chunk] + [r.steps - progress[r.rid] for r, _ in self.active])
            for r, m in self.active:
                if r.total_time < math.inf:
                    n = min(n, steps_before_total_time(*m.time()[:2], r.total_time, r.scenario.dt_max, n))
            if fresh:
                # the first step after a prepare captures the batch's graphs:
                # no other class allocates or frees meanwhile
                with srv.lock:
                    check(srv.lib.nhw_batch_step(self.h, n), "nhw_batch_step")
                fresh = False
            else:
                check(srv.lib.nhw_batch_step(self.h, n), "nhw_batch_step")
            self.steps_run += n
            for r, _ in self.active:
                progress[r.rid] += n
            now = time.perf_counter() - t0
            with srv.lock:
                for idx in reversed(range(len(self.active))):
                    r, m = self.active[idx]
                    if progress[r.rid] >= r.steps or (r.total_time < math.inf and m.time()[0] >= r.total_time):
                        self._retire(idx, now)
                        dirty = True

    def close(self) -> None:
        for _, m in self.active:
            m.ctx = C.c_void_p()
        if self.h:
            self.srv.lib.nhw_batch_destroy(self.h)
            self.h = C.c_void_p()


class Server:
    """One GPU.  `hbm_budget` bounds the sum of the admitted runs' device
    footprints (nhw_footprint); `max_slots` bounds their count; `chunk` is
    the number of steps between admission/retirement points.

    `classes`: ascending upper bounds on a request's cell count
    (Mglob*Nglob*Kglob).  Each class is a batch of its own, on its own HIP
    stream (default priority), stepped by its own host thread, so a small
    run is not held to the step time of the largest run on the GPU; the
    classes share the HBM budget and the slots.  None: one batch for all."""

    def __init__(self, precision: int = capi.NHW_FP64, hbm_budget: float = 200e9, max_slots: int = 64,
                 chunk: int = 10, classes=None):
        self.lib = capi.load()
        self.precision = precision
        self.hbm_budget = hbm_budget
        self.max_slots = max_slots
        self.chunk = chunk
        bounds = sorted(float(b) for b in (classes or []))
        self.lanes = [_Lane(self, b) for b in bounds + [math.inf]]
        self.lock = threading.Lock()
        self.done: list[Request] = []
        self.used = 0
        self.n_active = 0
        self.peak_used = 0
        self.peak_active = 0

    @property
    def steps_run(self) -> int:
        """Batch steps taken, summed over the classes."""
        return sum(ln.steps_run for ln in self.lanes)

    def footprint(self, s: Scenario) -> int:
        b = C.c_int64()
        cfg = make_config(s, self.precision)
        check(self.lib.nhw_footprint(C.byref(cfg), C.byref(b)), "nhw_footprint")
        return int(b.value)

    def lane_of(self, r: Request) -> int:
        cells = r.scenario.mglob * r.scenario.nglob * r.scenario.kglob
        return next(q for q, ln in enumerate(self.lanes) if cells <= ln.bound)

    def run(self, requests: list[Request]) -> list[Request]:
        """Serves `requests` (arrival times relative to the call) to the end;
        returns them in completion order."""
        for r in requests:
            r.footprint = self.footprint(r.scenario)
            if r.footprint > self.hbm_budget:
                raise ValueError(f"request {r.rid}: footprint {r.footprint} exceeds the HBM budget")
        per_lane = [[] for _ in self.lanes]
        for r in sorted(requests, key=lambda r: r.arrival):
            per_lane[self.lane_of(r)].append(r)
        t0 = time.perf_counter()
        busy = [(ln, rs) for ln, rs in zip(self.lanes, per_lane) if rs]
        if len(busy) == 1:
            busy[0][0].run(busy[0][1], t0)
        else:
            threads = [threading.Thread(target=ln.run, args=(rs, t0), daemon=True) for ln, rs in busy]
            for th in threads:
                th.start()
            for th in threads:
                th.join()
        for ln, _ in busy:
            if ln.error is not None:
                raise ln.error
        return self.done

    def close(self) -> None:
        for ln in self.lanes:
            ln.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# estimate_dt's growth limit, dt <= 1.05 dt_old (src/nhwave.F:3715, 3838);
# a bound from above is all that is needed here
DT_GROWTH = 1.05


def steps_before_total_time(t: float, dt_old: float, total_time: float, dt_max: float, cap: int) -> int:
    """How many more steps, up to `cap`, a run at TIME t is sure to take
    before the reference's loop test TIME < TOTAL_TIME stops it: step 1
    runs (t < total_time), and step n+1 runs if TIME after n steps, at most
    t + sum_{i<=n} min(dt_old 1.05^i, DT_MAX), is below total_time.  A
    batch stepped that many steps leaves no member past its end."""
    if not t < total_time:
        return 0
    n, bound, d = 1, t, dt_old
    while n < cap:
        d = min(d * DT_GROWTH, dt_max)
        bound += d
        if not bound * (1.0 + 1e-12) < total_time:
            break
        n += 1
    return n


def poisson_stream(lam: float, count: int, seed: int = 7, n_range=(64, 1024), k_range=(3, 10), steps: int = 1000,
                   tol: float = 1e-10) -> list[Request]:
    """Requests with exponential inter-arrival times at rate `lam` (1/s):
    grids log-uniform between n_range[0]^2 and n_range[1]^2 columns, Kglob
    uniform in k_range, seeded bump bathymetry with a drying beach, a hump
    source of amplitude 0.5-5 m and radius 200-2000 m."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(count):
        t += rng.exponential(1.0 / lam) if lam > 0 else 0.0
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        k = int(rng.integers(k_range[0], k_range[1] + 1))
        amp = float(rng.uniform(0.5, 5.0))
        radius = float(rng.uniform(200.0, 2000.0))
        s = bumps_beach(n=n, kglob=k, dx=10.0, seed=int(rng.integers(1 << 30)), amp=amp, radius=radius,
                        steps=steps, tol=tol)
        out.append(Request(rid=rid, scenario=s, steps=steps, arrival=t))
    return out


def summary(done: list[Request], wall: float) -> dict:
    lat = np.array([r.latency for r in done])
    cs = sum(r.cell_stages for r in done)
    return dict(requests=len(done), p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_s=float(lat.mean()), max_s=float(lat.max()), wall_s=wall, cell_stages=cs,
                value=cs / wall, not_converged=int(sum(r.stats.get("not_converged", 0) for r in done)))
