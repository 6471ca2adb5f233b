"""Serving loop (continuous batching): requests that arrive while others run
join the same batch, advance in the same launches, and leave at their step
count; each result equals the run of the same request alone (bit-identical
on the hydrostatic path, within the Poisson tolerance with the solve)."""
import numpy as np
import pytest

from nhwave_amd import capi
from nhwave_amd import scenario as S
from nhwave_amd.model import Model
from nhwave_amd.serve import Request, Server, poisson_stream, summary

pytestmark = pytest.mark.gpu

G = S.NGHOST


def rel(a, b):
    return float(np.abs(a - b).max()) / max(float(np.abs(b).max()), 1e-300)


def requests(non_hydro):
    tol = 1e-11
    return [
        Request(0, S.bumps_beach(n=48, kglob=4, dx=100.0, non_hydro=non_hydro, tol=tol), 12, 0.0),
        Request(1, S.beach_runup(mglob=64, nglob=6, kglob=3, non_hydro=non_hydro, tol=tol), 20, 0.0),
        Request(2, S.bumps_beach(n=70, m=40, kglob=6, dx=50.0, seed=3, non_hydro=non_hydro, tol=tol), 8, 0.3),
        Request(3, S.random_scenario(5, mglob=30, nglob=22, kglob=5, non_hydro=non_hydro, tol=tol), 16, 0.6),
    ]


def alone(r):
    with Model(r.scenario) as m:
        m.step(r.steps)
        return m.download(*r.fields), m.time()


@pytest.mark.parametrize("non_hydro", [False, True])
def test_serving_results_match_alone(gpu, non_hydro):
    reqs = requests(non_hydro)
    with Server(chunk=4) as srv:
        done = srv.run(reqs)
    assert sorted(r.rid for r in done) == [0, 1, 2, 3]
    # later arrivals joined while earlier runs were still on the GPU
    by = {r.rid: r for r in done}
    assert any(by[a].admitted < by[b].finished and by[b].admitted < by[a].finished
               for a in by for b in by if a < b)
    for r in done:
        ref, t = alone(r)
        assert r.stats["time"] == t or non_hydro, r.rid
        assert np.array_equal(r.result["Mask"], ref["Mask"]), r.rid
        if not non_hydro:
            for n in r.fields:
                assert np.array_equal(r.result[n], ref[n]), (r.rid, n)
        else:
            assert r.stats["not_converged"] == 0
            g = np.s_[..., G:-G, G:-G]
            assert rel(r.result["Eta"][g], ref["Eta"][g]) < 1e-11, r.rid
            for n in ("U", "W"):
                assert rel(r.result[n][G:-G, G:-G, G:-G], ref[n][G:-G, G:-G, G:-G]) < 1e-7, (r.rid, n)


def test_serving_admission_respects_slots_and_budget(gpu):
    """One slot: the requests are served one after another, in arrival
    order.  A budget of the largest footprint: the admitted footprints never
    exceed it, and every request still completes."""
    reqs = requests(True)
    with Server(chunk=4, max_slots=1) as srv:
        done = srv.run(reqs)
        assert srv.peak_active == 1
    assert [r.rid for r in done] == [0, 1, 2, 3]
    for a, b in zip(done, done[1:]):
        assert b.admitted >= a.finished
    reqs = requests(True)
    with Server(chunk=4) as srv:
        budget = max(srv.footprint(r.scenario) for r in reqs)
        srv.hbm_budget = budget
        done = srv.run(reqs)
        assert 0 < srv.peak_used <= budget
    assert sorted(r.rid for r in done) == [0, 1, 2, 3]


def test_poisson_stream_summary(gpu):
    reqs = poisson_stream(lam=50.0, count=5, n_range=(32, 96), k_range=(3, 5), steps=10)
    with Server(precision=capi.NHW_MIXED, chunk=5) as srv:
        import time
        t0 = time.perf_counter()
        done = srv.run(reqs)
        wall = time.perf_counter() - t0
    sm = summary(done, wall)
    assert sm["requests"] == 5 and sm["not_converged"] == 0
    assert 0 < sm["p50_s"] <= sm["p99_s"] <= sm["max_s"]


def test_run_stops_where_the_reference_loop_stops(gpu):
    """Model.run (nhw_run) against the TIME the reference printed after each
    step (tests/golden/stats_beach_hydro.json): with TOTAL_TIME between the
    printed TIME of steps n and n+1 the reference's loop takes n+1 steps."""
    import json
    from pathlib import Path
    rec = json.loads((Path(__file__).resolve().parent / "golden" / "stats_beach_hydro.json").read_text())
    times = [st["time"] for st in rec["steps"]]
    s = getattr(S, rec["generator"])(**rec["kwargs"])
    for n in (0, 9, 25):
        with Model(s) as m:
            taken = m.run(total_time=0.5 * (times[n] + times[n + 1]), sim_steps=10 ** 9)
            assert taken == n + 2 and m.time()[2] == n + 2
    with Model(s) as m:
        assert m.run(total_time=1e9, sim_steps=7) == 7
        assert m.run(total_time=1e9, sim_steps=7) == 0   # RUN_STEP already at SIM_STEPS


@pytest.mark.parametrize("non_hydro", [False, True])
def test_serving_run_dirs_to_total_time(gpu, tmp_path, non_hydro):
    """Requests read from run directories (the reference's input.txt and
    array files) and bounded by TOTAL_TIME end on the step the reference's
    loop ends on, with the state of the same run alone (a wavemaker flume
    among them); the chunk bound keeps every member from passing its end
    inside a chunk."""
    tol = 1e-11
    scen = [S.bumps_beach(n=40, kglob=4, dx=100.0, non_hydro=non_hydro, tol=tol, steps=10 ** 6),
            S.beach_runup(mglob=64, nglob=6, kglob=3, non_hydro=non_hydro, tol=tol, steps=10 ** 6),
            S.random_scenario(5, mglob=30, nglob=22, kglob=5, non_hydro=non_hydro, tol=tol, steps=14),
            S.wave_flume(mglob=100, nglob=3, wavemaker="LEF_LIN", amp=0.02, per=1.5, bottom="bar",
                         sponge=(0.0, 0.5, 0.0, 0.0), non_hydro=non_hydro, tol=tol, steps=10 ** 6)]
    reqs = []
    for i, s in enumerate(scen):
        with Model(s) as m:          # a TOTAL_TIME that lands between steps
            m.step(9 + 4 * i)
            t1 = m.time()[0]
            m.step(1)
            s.total_time = 0.5 * (t1 + m.time()[0])
        S.write_reference_inputs(s, tmp_path / f"run{i}")
        reqs.append(Request.from_run_dir(i, tmp_path / f"run{i}", arrival=0.1 * i))
    with Server(chunk=8) as srv:
        done = srv.run(reqs)
    by = {r.rid: r for r in done}
    for i, s in enumerate(scen):
        r = by[i]
        with Model(r.scenario) as m:
            taken = m.run()
            ref = m.download(*r.fields)
            t = m.time()
        assert taken == min(10 + 4 * i, s.steps) and r.stats["time"][2] == t[2], (i, taken, r.stats["time"])
        assert np.array_equal(r.result["Mask"], ref["Mask"]), i
        if not non_hydro:
            assert r.stats["time"] == t
            for n in r.fields:
                assert np.array_equal(r.result[n], ref[n]), (i, n)
        else:
            assert rel(r.result["Eta"][G:-G, G:-G], ref["Eta"][G:-G, G:-G]) < 1e-11, i


def test_serving_size_classes(gpu):
    """Size classes: small runs in their own batch, on their own stream and
    host thread, finish while a large run that arrived first is still on
    the GPU, and every result equals the run alone (hydrostatic: bit for
    bit)."""
    reqs = [Request(0, S.bumps_beach(n=256, kglob=6, dx=100.0, non_hydro=False), 300, 0.0)]
    for q in range(1, 4):
        reqs.append(Request(q, S.random_scenario(q, mglob=24, nglob=18, kglob=3, non_hydro=False), 10,
                            0.01 * q))
    with Server(chunk=5, classes=[5000]) as srv:
        done = srv.run(reqs)
        assert srv.peak_active >= 2
    by = {r.rid: r for r in done}
    big = by[0]
    for q in range(1, 4):
        assert big.admitted <= by[q].admitted and by[q].finished < big.finished, q
    for r in done:
        ref, t = alone(r)
        assert r.stats["time"] == t, r.rid
        for n in r.fields:
            assert np.array_equal(r.result[n], ref[n]), (r.rid, n)
