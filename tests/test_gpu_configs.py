"""The mixed, batched and serving forms at their full configurations, and
batch and serving members checked against the reference's golden dumps.

* mixed precision on 2048 x 2048 x 8 (outer Poisson TOL 1e-8) against the
  plain FP64 form on the same input, the FP64 oracle at a size the CPU
  cannot hold;
* the 64-scenario ensemble (grids 128^2 .. 768^2, Kglob 3 .. 10) in one
  mixed-precision batch, with members re-run alone;
* a serving stream whose requests span 64^2 x 3 .. 1024^2 x 10 at 1000 steps
  each, its largest and smallest results against the same runs alone;
* a batch and a serving run made of golden scenarios, every member against
  the reference build's dumps."""
import numpy as np
import pytest

from nhw_testutil import golden_case, golden_state, interior_state, rel_errors
from nhwave_amd import capi
from nhwave_amd import scenario as S
from nhwave_amd.batch import Batch
from nhwave_amd.model import Model
from nhwave_amd.serve import Request, Server, poisson_stream

pytestmark = pytest.mark.gpu

G = S.NGHOST


def rel(a, b):
    return float(np.abs(a - b).max()) / max(float(np.abs(b).max()), 1e-300)


def wet_rel(a, b, ok):
    """max relative error of 3-D fields over the interior sigma levels of the
    columns selected by `ok`."""
    x = a[G:-G, G:-G, G:-G][:, ok]
    y = b[G:-G, G:-G, G:-G][:, ok]
    return float(np.abs(x - y).max()) / max(float(np.abs(y).max()), 1e-300)


def run_alone(s, steps, precision, fields=("Eta", "D", "Mask", "U", "V", "W")):
    with Model(s, precision=precision) as m:
        m.step(steps)
        out = m.download(*fields)
        st = m.stats()
    return out, st


def test_config3_mixed_2048x2048x8_vs_fp64(gpu):
    """2048 x 2048 x 8, bump bathymetry with a drying beach, TOL 1e-8: the
    mixed form stays within 1e-5 (Eta) and 1e-4 (U, V, W where the water is
    deeper than 0.5 m) of the FP64 form after 3 steps, every solve
    converges, the state is finite and the masks agree."""
    s = S.bumps_beach(n=2048, kglob=8, dx=10.0, steps=2000, tol=1e-8)
    a, sa = run_alone(s, 3, capi.NHW_MIXED)
    assert sa.not_converged == 0 and sa.solves == 6 and sa.last_relres <= s.tol, sa.last_relres
    for n, v in a.items():
        assert np.isfinite(v).all(), n
    b, sb = run_alone(s, 3, capi.NHW_FP64)
    assert sb.not_converged == 0
    ok = (a["Mask"][G:-G, G:-G] == 1) & (b["Mask"][G:-G, G:-G] == 1) & (b["D"][G:-G, G:-G] > 0.5)
    err = {"Eta": rel(a["Eta"][G:-G, G:-G], b["Eta"][G:-G, G:-G])}
    for n in ("U", "V", "W"):
        err[n] = wet_rel(a[n], b[n], ok)
    flips = int((a["Mask"] != b["Mask"]).sum())
    print(f"config 3 mixed vs FP64 after 3 steps: {err}, mask flips {flips}")
    assert err["Eta"] < 1e-5, err
    assert max(err["U"], err["V"], err["W"]) < 1e-4, err
    assert flips == 0


def test_config4_ensemble64_mixed(gpu):
    """The 64-member ensemble in one mixed-precision batch for 2 steps: every
    member's solves converge to TOL 1e-8 and its state is finite; the
    smallest, the largest and a middle member re-run alone agree with their
    batch results to the solve tolerance, masks exactly."""
    sc = S.ensemble(64, 11)
    cells = [s.mglob * s.nglob * s.kglob for s in sc]
    order = np.argsort(cells)
    pick = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    got = {}
    with Batch(sc, precision=capi.NHW_MIXED, autotune=False) as b:
        b.step(2)
        for i, m in enumerate(b.members):
            st = m.stats()
            assert st.not_converged == 0 and st.solves == 4 and st.last_relres <= sc[i].tol, (i, st.last_relres)
            eta = m.download("Eta")["Eta"]
            assert np.isfinite(eta).all(), i
            if i in pick:
                got[i] = m.download("Eta", "D", "Mask", "U", "V", "W")
    for i in pick:
        want, st = run_alone(sc[i], 2, capi.NHW_MIXED)
        ok = (want["Mask"][G:-G, G:-G] == 1) & (want["D"][G:-G, G:-G] > 0.5)
        e_eta = rel(got[i]["Eta"][G:-G, G:-G], want["Eta"][G:-G, G:-G])
        e_vel = max(wet_rel(got[i][n], want[n], ok) for n in ("U", "V", "W"))
        print(f"member {i} ({sc[i].mglob}x{sc[i].nglob}x{sc[i].kglob}): batch vs alone Eta {e_eta:.2e}, "
              f"U/V/W {e_vel:.2e}")
        assert np.array_equal(got[i]["Mask"], want["Mask"]), i
        assert e_eta < 1e-7 and e_vel < 1e-6, (i, e_eta, e_vel)


def test_config5_serving_stream_extremes(gpu):
    """A Poisson-arrival stream (8 requests/s, mixed precision, TOL 1e-10,
    1000 steps each) whose first request is the largest grid of the range,
    1024^2 x 10, and whose last is the smallest, 64^2 x 3: both results,
    computed while other requests shared the GPU, equal their runs alone to
    the solve tolerance; every request converges."""
    reqs = poisson_stream(8.0, 6, seed=5, steps=1000, tol=1e-10)
    first, last = reqs[0], reqs[-1]
    first.scenario = S.bumps_beach(n=1024, kglob=10, dx=10.0, seed=77, amp=3.0, radius=800.0, steps=1000,
                                   tol=1e-10)
    last.scenario = S.bumps_beach(n=64, kglob=3, dx=10.0, seed=78, amp=1.0, radius=300.0, steps=1000,
                                  tol=1e-10)
    with Server(precision=capi.NHW_MIXED, chunk=50, classes=[131072, 1048576]) as srv:
        done = srv.run(reqs)
    assert len(done) == 6 and all(r.stats["not_converged"] == 0 for r in done)
    by = {r.rid: r for r in done}
    # the large request shared the GPU with later arrivals
    big = by[first.rid]
    assert any(by[q].admitted < big.finished for q in by if q != first.rid)
    for r in (big, by[last.rid]):
        want, _ = run_alone(r.scenario, r.steps, capi.NHW_MIXED, fields=r.fields + ("D",))
        ok = (want["Mask"][G:-G, G:-G] == 1) & (want["D"][G:-G, G:-G] > 0.5)
        e_eta = rel(r.result["Eta"][G:-G, G:-G], want["Eta"][G:-G, G:-G])
        e_vel = max(wet_rel(r.result[n], want[n], ok) for n in ("U", "V", "W"))
        print(f"request {r.rid} ({r.scenario.mglob}^2 x {r.scenario.kglob}): served vs alone Eta {e_eta:.2e}, "
              f"U/V/W {e_vel:.2e}, latency {r.latency:.2f} s")
        assert np.array_equal(r.result["Mask"], want["Mask"]), r.rid
        assert e_eta < 1e-8 and e_vel < 1e-6, (r.rid, e_eta, e_vel)


GOLDEN_MIX = ("beach_hydro", "bumps48_hydro", "flume_sol_beach_hydro")
# the wavemaker's functions of TIME use the device libm (see the parity tests)
LIBM_TOL = {"flume_sol_beach_hydro": 1e-12}


def _check_golden(name, s, st, got, data):
    ours = interior_state(s, got.__getitem__)
    ref = golden_state(data, st)
    assert np.array_equal(ours["Mask"], ref["Mask"]), (name, st)
    tol = LIBM_TOL.get(name)
    if tol is None:
        bad = [n for n in ("Eta", "D", "U", "V", "W", "DU", "DV", "DW", "Omega") if not np.array_equal(ours[n], ref[n])]
        assert bad == [], (name, st, bad)
    else:
        err = rel_errors(ours, ref)
        assert max(err[n] for n in ("Eta", "D", "U", "V", "W")) < tol, (name, st, err)


def test_batch_of_golden_scenarios_vs_reference(gpu):
    """Three golden scenarios (a run-up beach, the bump basin, a solitary-wave
    flume with its inflow boundary) in one FP64 batch: at every dump step of
    its fixture each member equals the reference build's state."""
    cases = {n: golden_case(n) for n in GOLDEN_MIX}
    scen = [cases[n][0] for n in GOLDEN_MIX]
    dumps = sorted({st for n in GOLDEN_MIX for st in cases[n][1]["steps"]})
    fields = ("Eta", "D", "Mask", "U", "V", "W", "DU", "DV", "DW", "P", "Omega")
    with Batch(scen, precision=capi.NHW_FP64, autotune=False) as b:
        done = 0
        for st in dumps:
            b.step(st - done)
            done = st
            for q, n in enumerate(GOLDEN_MIX):
                s, meta, data = cases[n]
                if st in meta["steps"]:
                    _check_golden(n, s, st, b.members[q].download(*fields), data)
                    t, dt, rs = b.members[q].time()
                    assert rs == st and abs(t - meta["time"][str(st)]) <= 1e-13 * t, (n, st, t)


def test_serving_golden_scenarios_vs_reference(gpu):
    """The same three golden scenarios as requests of the serving loop, each
    run to its fixture's last dump step while the others share the GPU: each
    result equals the reference build's state at that step."""
    cases = {n: golden_case(n) for n in GOLDEN_MIX}
    fields = ("Eta", "D", "Mask", "U", "V", "W", "DU", "DV", "DW", "P", "Omega")
    reqs = [Request(q, cases[n][0], cases[n][1]["steps"][-1], 0.01 * q, fields=fields)
            for q, n in enumerate(GOLDEN_MIX)]
    with Server(precision=capi.NHW_FP64, chunk=10) as srv:
        done = srv.run(reqs)
    by = {r.rid: r for r in done}
    assert any(by[a].admitted < by[b].finished and by[b].admitted < by[a].finished
               for a in by for b in by if a < b)
    for q, n in enumerate(GOLDEN_MIX):
        s, meta, data = cases[n]
        _check_golden(n, s, meta["steps"][-1], by[q].result, data)
