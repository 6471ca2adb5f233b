"""Serving loop: scenario runs arrive as requests, are admitted into a
running batch while their device footprint fits the HBM budget, advance
together with the runs already on the GPU (one launch per kernel for all of
them), and leave the batch when they reach their step count, which frees
their memory for the next arrivals (continuous batching).  Reports
time-to-result percentiles beside throughput.

The reference runs one scenario per process to its SIM_STEPS
(src/nhwave.F:119-300); a request here is the same run: a scenario, a step
count, and the fields that the run returns at its end.
"""
from __future__ import annotations

import ctypes as C
import math
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


class Server:
    """One GPU, one batch.  `hbm_budget` bounds the sum of the admitted
    runs' device footprints (nhw_footprint); `max_slots` bounds their count;
    `chunk` is the number of steps between admission/retirement points."""

    def __init__(self, precision: int = capi.NHW_FP64, hbm_budget: float = 200e9, max_slots: int = 64,
                 chunk: int = 10):
        self.lib = capi.load()
        self.precision = precision
        self.hbm_budget = hbm_budget
        self.max_slots = max_slots
        self.chunk = chunk
        self.h = C.c_void_p()
        check(self.lib.nhw_batch_create(None, 0, C.byref(self.h)), "nhw_batch_create")
        self.active: list[tuple[Request, Model]] = []
        self.done: list[Request] = []
        self.used = 0
        self.peak_used = 0
        self.peak_active = 0
        self.steps_run = 0

    def footprint(self, s: Scenario) -> int:
        b = C.c_int64()
        cfg = make_config(s, self.precision)
        check(self.lib.nhw_footprint(C.byref(cfg), C.byref(b)), "nhw_footprint")
        return int(b.value)

    def _admit(self, r: Request, now: float) -> None:
        cfg = make_config(r.scenario, self.precision)
        ctx = C.c_void_p()
        check(self.lib.nhw_batch_add(self.h, C.byref(cfg), C.byref(ctx)), "nhw_batch_add")
        m = Model(r.scenario, handle=ctx)
        r.admitted = now
        self.used += r.footprint
        self.active.append((r, m))
        self.peak_used = max(self.peak_used, self.used)
        self.peak_active = max(self.peak_active, len(self.active))

    def _retire(self, idx: int, now: float) -> None:
        r, m = self.active[idx]
        r.result = m.download(*r.fields)
        st = m.stats()
        r.stats = dict(solves=st.solves, iters=st.total_iters, not_converged=st.not_converged,
                       last_relres=st.last_relres, time=m.time())
        m.ctx = C.c_void_p()  # the batch owns the context
        check(self.lib.nhw_batch_remove(self.h, idx), "nhw_batch_remove")
        self.used -= r.footprint
        r.finished = now
        self.done.append(r)
        del self.active[idx]

    def run(self, requests: list[Request]) -> list[Request]:
        """Serves `requests` (arrival times relative to the call) to the end;
        returns them in completion order."""
        for r in requests:
            r.footprint = self.footprint(r.scenario)
            if r.footprint > self.hbm_budget:
                raise ValueError(f"request {r.rid}: footprint {r.footprint} exceeds the HBM budget")
        pending = sorted(requests, key=lambda r: r.arrival)
        progress: dict[int, int] = {}
        t0 = time.perf_counter()
        dirty = False
        while pending or self.active:
            now = time.perf_counter() - t0
            # admit, in arrival order, while slots and memory allow
            while (pending and pending[0].arrival <= now and len(self.active) < self.max_slots
                   and self.used + pending[0].footprint <= self.hbm_budget):
                r = pending.pop(0)
                self._admit(r, now)
                progress[r.rid] = 0
                dirty = True
            if not self.active:
                # idle until the next arrival
                time.sleep(max(0.0, min(pending[0].arrival - now, 0.01)))
                continue
            if dirty:
                check(self.lib.nhw_batch_prepare(self.h, 0), "nhw_batch_prepare")
                dirty = False
            n = min([self.chunk] + [r.steps - progress[r.rid] for r, _ in self.active])
            for r, m in self.active:
                if r.total_time < math.inf:
                    n = min(n, steps_before_total_time(*m.time()[:2], r.total_time, r.scenario.dt_max, n))
            check(self.lib.nhw_batch_step(self.h, n), "nhw_batch_step")
            self.steps_run += n
            for r, _ in self.active:
                progress[r.rid] += n
            now = time.perf_counter() - t0
            for idx in reversed(range(len(self.active))):
                r, m = self.active[idx]
                if progress[r.rid] >= r.steps or (r.total_time < math.inf and m.time()[0] >= r.total_time):
                    self._retire(idx, now)
                    dirty = True
        return self.done

    def close(self) -> None:
        for _, m in self.active:
            m.ctx = C.c_void_p()
        if self.h:
            self.lib.nhw_batch_destroy(self.h)
            self.h = C.c_void_p()

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
