"""Batched form: an ensemble of uneven scenarios advanced together (one
launch per step for all members) gives each member the state it gets alone:
bit-identical on the explicit (hydrostatic) path, which runs the same
per-cell code; within the Poisson tolerance with the solve (the batched
reductions sum in a different order)."""
import numpy as np
import pytest

from nhwave_amd import capi
from nhwave_amd import scenario as S
from nhwave_amd.batch import Batch
from nhwave_amd.model import Model

pytestmark = pytest.mark.gpu

G = S.NGHOST
FIELDS = ("Eta", "D", "Mask", "U", "V", "W", "P")


def rel(a, b):
    return float(np.abs(a - b).max()) / max(float(np.abs(b).max()), 1e-300)


def members(non_hydro, tol=1e-10):
    return [
        S.bumps_beach(n=40, m=52, kglob=4, dx=200.0, non_hydro=non_hydro, tol=tol),
        S.beach_runup(mglob=48, nglob=6, kglob=3, non_hydro=non_hydro, tol=tol),
        S.random_scenario(12, mglob=20, nglob=15, kglob=4, non_hydro=non_hydro, tol=tol),
        S.bumps_beach(n=96, kglob=8, dx=20.0, non_hydro=non_hydro, tol=tol),
        S.bumps_beach(n=33, m=70, kglob=5, dx=50.0, non_hydro=non_hydro, tol=tol),
    ]


def alone(s, steps, precision):
    with Model(s, precision=precision) as m:
        m.step(steps)
        return m.download(*FIELDS), m.time(), m.stats()


def test_batch_hydrostatic_bit_identical(gpu):
    sc = members(False)
    with Batch(sc) as b:
        b.step(12)
        got = [(m.download(*FIELDS), m.time()) for m in b.members]
    for s, (f, t) in zip(sc, got):
        ref, tref, _ = alone(s, 12, capi.NHW_FP64)
        assert t == tref, s.name
        bad = [n for n in FIELDS if not np.array_equal(f[n], ref[n])]
        assert bad == [], (s.name, bad)


def test_batch_nonhydrostatic_matches_alone(gpu):
    sc = members(True, tol=1e-11)
    with Batch(sc) as b:
        b.step(6)
        got = [(m.download(*FIELDS), m.stats()) for m in b.members]
    for s, (f, st) in zip(sc, got):
        ref, _, st_ref = alone(s, 6, capi.NHW_FP64)
        assert st.not_converged == 0 and st.last_relres <= s.tol, (s.name, st.last_relres)
        assert np.array_equal(f["Mask"], ref["Mask"]), s.name
        g = np.s_[..., G:-G, G:-G]
        assert rel(f["Eta"][g], ref["Eta"][g]) < 1e-11, s.name
        for n in ("U", "W", "P"):
            assert rel(f[n][g], ref[n][g]) < 1e-7, (s.name, n, rel(f[n][g], ref[n][g]))


def test_batch_mixed_matches_alone(gpu):
    sc = members(True, tol=1e-8)
    with Batch(sc, precision=capi.NHW_MIXED) as b:
        b.step(6)
        got = [(m.download(*FIELDS), m.stats()) for m in b.members]
    for s, (f, st) in zip(sc, got):
        ref, _, _ = alone(s, 6, capi.NHW_MIXED)
        assert st.not_converged == 0 and st.last_relres <= s.tol, (s.name, st.last_relres)
        g = np.s_[..., G:-G, G:-G]
        assert rel(f["Eta"][g], ref["Eta"][g]) < 1e-7, s.name


def test_batch_autotuner_reports_choices(gpu):
    with Batch(members(True)[:3]) as b:
        t = b.tuning()
    assert len(t) == 7
    for r in t:
        assert r["threads"] in (64, 128, 256)
        assert all(v > 0 for v in r["ms"].values()), r


def test_batch_mixed_ensemble_with_shrinking_active_set(gpu):
    """Members converge after different numbers of iterations and
    refinement rounds, so the solve's launch tables shrink between polls
    (regression: the tables are rewritten only after the batch's stream has
    drained).  Every member still reaches TOL and matches its run alone."""
    rng = np.random.default_rng(5)
    sc = [S.bumps_beach(n=int(n), kglob=int(k), dx=10.0, seed=int(sd), amp=float(a), tol=1e-8)
          for n, k, sd, a in zip(rng.integers(96, 200, 12), rng.integers(3, 11, 12),
                                 rng.integers(1, 1 << 20, 12), rng.uniform(0.5, 5.0, 12))]
    with Batch(sc, precision=capi.NHW_MIXED) as b:
        b.step(4)
        got = [(m.download("Eta"), m.stats()) for m in b.members]
    for s, (f, st) in zip(sc, got):
        assert st.not_converged == 0 and st.last_relres <= s.tol, (s.name, st.last_relres)
        ref, _, _ = alone(s, 4, capi.NHW_MIXED)
        g = np.s_[G:-G, G:-G]
        assert rel(f["Eta"][g], ref["Eta"][g]) < 1e-7, s.name


@pytest.mark.parametrize("precision", [capi.NHW_FP64, capi.NHW_MIXED])
def test_batch_graph_replay_matches_eager(gpu, precision):
    """The explicit launches between Poisson solves are captured once into
    HIP graphs and replayed every stage, with the eager solve and a host
    synchronisation between replays; every replayed step must give exactly
    the state of the same batch launched eagerly."""
    sc = members(True, tol=1e-9)
    got = {}
    for graphs in (True, False):
        with Batch(sc, precision=precision, autotune=False, graphs=graphs) as b:
            states = []
            for _ in range(4):
                b.step(1)
                states.append([(m.download(*FIELDS), m.time()) for m in b.members])
            got[graphs] = states
    for step, (sg, se) in enumerate(zip(got[True], got[False])):
        for s, (fg, tg), (fe, te) in zip(sc, sg, se):
            assert tg == te, (s.name, step)
            bad = [n for n in FIELDS if not np.array_equal(fg[n], fe[n])]
            assert bad == [], (s.name, step, bad)


def forced_members(non_hydro, tol=1e-10):
    """Wavemaker and sponge flumes (one-row and wider) beside unforced members."""
    return [
        S.wave_flume(mglob=120, wavemaker="LEF_SOL", amp=0.06, depth=0.21, dx=0.05, bottom="beach",
                     non_hydro=non_hydro, tol=tol),
        S.bumps_beach(n=40, m=52, kglob=4, dx=200.0, non_hydro=non_hydro, tol=tol),
        S.wave_flume(mglob=150, nglob=5, wavemaker="LEF_LIN", amp=0.02, per=2.02, bottom="bar",
                     sponge=(0.0, 0.8, 0.0, 0.0), non_hydro=non_hydro, tol=tol),
        S.wave_flume(mglob=40, nglob=12, dx=0.05, kglob=4, wavemaker="LEF_STK", amp=0.03, per=1.5,
                     bottom="flat", sponge=(0.3, 0.5, 0.2, 0.2), cur=0.05, tramp=0.3, non_hydro=non_hydro,
                     tol=tol),
        S.random_scenario(12, mglob=20, nglob=15, kglob=4, non_hydro=non_hydro, tol=tol),
    ]


def test_batch_forcing_hydrostatic_bit_identical(gpu):
    """Members with a west wavemaker and sponges, batched with unforced ones:
    each gets the state it gets alone, bit for bit."""
    sc = forced_members(False)
    with Batch(sc) as b:
        b.step(25)
        got = [(m.download(*FIELDS), m.time()) for m in b.members]
    for s, (f, t) in zip(sc, got):
        ref, tref, _ = alone(s, 25, capi.NHW_FP64)
        assert t == tref, s.name
        bad = [n for n in FIELDS if not np.array_equal(f[n], ref[n])]
        assert bad == [], (s.name, bad)


@pytest.mark.parametrize("precision", [capi.NHW_FP64, capi.NHW_MIXED])
def test_batch_forcing_nonhydrostatic(gpu, precision):
    """The same members non-hydrostatic: graph replays equal eager launches
    bit for bit, and every member stays within the solve tolerance of its
    run alone."""
    sc = forced_members(True, tol=1e-10)
    got = {}
    for graphs in (True, False):
        with Batch(sc, precision=precision, autotune=False, graphs=graphs) as b:
            b.step(8)
            got[graphs] = [(m.download(*FIELDS), m.stats()) for m in b.members]
    for s, (fg, st), (fe, _) in zip(sc, got[True], got[False]):
        assert st.not_converged == 0, s.name
        bad = [n for n in FIELDS if not np.array_equal(fg[n], fe[n])]
        assert bad == [], (s.name, bad)
        ref, _, _ = alone(s, 8, precision)
        assert np.array_equal(fg["Mask"], ref["Mask"]), s.name
        g = np.s_[..., G:-G, G:-G]
        tol = 1e-10 if precision == capi.NHW_FP64 else 1e-6
        assert rel(fg["Eta"][g], ref["Eta"][g]) < tol, (s.name, rel(fg["Eta"][g], ref["Eta"][g]))


def test_batch_bottom_friction_bit_identical(gpu):
    """Members with the friction-law (BC_Z0 = 5, different Cd0) and no-slip
    (BC_Z0 = 2) bottoms, hydrostatic: each gets the state it gets alone, bit
    for bit."""
    from dataclasses import replace
    sc = [replace(s, bc=(1, 1, 1, 1, bz), cd0=cd) for s, bz, cd in
          zip(members(False)[:4], (5, 5, 2, 5), (0.004, 0.01, 0.0025, 0.002))]
    with Batch(sc) as b:
        b.step(12)
        got = [(m.download(*FIELDS), m.time()) for m in b.members]
    for s, (f, t) in zip(sc, got):
        ref, tref, _ = alone(s, 12, capi.NHW_FP64)
        assert t == tref, s.name
        bad = [n for n in FIELDS if not np.array_equal(f[n], ref[n])]
        assert bad == [], (s.name, bad)


KE_FIELDS = ("Tke", "Eps", "CmuVt", "DTke", "DEps")


def viscous_members(non_hydro, tol=1e-11):
    """Viscous members (constant viscosities with wind, and a wavemaker
    flume with a sponge; the k-epsilon model with the friction-law bottom;
    wall friction side walls, BC 5, in both) beside inviscid ones, and
    internal-wavemaker basins, of uneven sizes."""
    from dataclasses import replace
    kw = dict(non_hydro=non_hydro, tol=tol)
    r = S.random_scenario(21, **kw)
    f = S.wave_flume(mglob=60, nglob=3, wavemaker="LEF_LIN", amp=0.02, per=1.5, bottom="bar",
                     sponge=(0.0, 0.6, 0.0, 0.0), non_hydro=non_hydro)
    k1 = S.random_scenario(23, **kw)
    k2 = S.random_scenario(29, mglob=40, nglob=3, kglob=7, dx=0.5, dy=0.5, **kw)
    return [replace(r, viscous=True, visc=1e-3, cvs=0.03, chs=0.08, wind_u=12.0, wind_v=-7.0, bc=(5, 2, 1, 5, 1),
                    cd0=0.004),
            S.random_scenario(12, mglob=20, nglob=15, kglob=4, **kw),
            replace(k1, viscous=True, turb=10, visc=1e-3, cvs=0.03, chs=0.08, bc=(5, 5, 5, 5, 5), ibot=1,
                    cd0=0.003, wind_u=9.0),
            replace(f, viscous=True, visc=1e-3, cvs=0.001, chs=0.002, tol=tol),
            S.bumps_beach(n=40, m=52, kglob=4, dx=200.0, **kw),
            replace(k2, viscous=True, turb=10, visc=1e-4, schmidt=0.8, rng=False, bc=k2.bc[:4] + (5,), ibot=1,
                    cd0=0.002),
            # the internal wavemaker (its mass source launched per member),
            # inviscid and with the k-epsilon model reset inside the box
            S.internal_wave_basin(mglob=60, nglob=10, dirs=(0.0, 12.0), seed=5, **kw),
            replace(S.internal_wave_basin(mglob=44, nglob=8, box=(1.0, 1.2, 0.1, 0.3), seed=6, **kw),
                    viscous=True, turb=10, visc=1e-6, cvs=0.001, chs=0.001, bc=(1, 1, 1, 1, 5), cd0=0.004)]


def member_state(m, s):
    names = FIELDS + (KE_FIELDS if s.viscous and s.turb == 10 else ())
    return m.download(*names)


@pytest.mark.parametrize("graphs", [True, False])
def test_batch_viscous_members_bit_identical(gpu, graphs):
    """A batch with viscous members (constant viscosities and the k-epsilon
    model) beside inviscid ones, hydrostatic, replayed from HIP graphs or
    launched eagerly: every member gets the state and the time steps it gets
    alone, bit for bit, the k-epsilon arrays included."""
    sc = viscous_members(False)
    with Batch(sc, graphs=graphs) as b:
        b.step(10)
        got = [(member_state(m, s), m.time()) for m, s in zip(b.members, sc)]
    for s, (f, t) in zip(sc, got):
        with Model(s) as m:
            m.step(10)
            ref, tref = member_state(m, s), m.time()
        assert t == tref, s.name
        bad = [n for n in ref if not np.array_equal(f[n], ref[n])]
        assert bad == [], (s.name, bad)


def test_batch_viscous_members_nonhydrostatic(gpu):
    """The same members non-hydrostatic: each converges and stays within the
    Poisson tolerance of its run alone (the batched solve's reductions sum in
    another order), with the same wet-dry mask."""
    sc = viscous_members(True)
    with Batch(sc) as b:
        b.step(6)
        got = [(member_state(m, s), m.stats()) for m, s in zip(b.members, sc)]
    for s, (f, st) in zip(sc, got):
        with Model(s) as m:
            m.step(6)
            ref = member_state(m, s)
        assert st.not_converged == 0, s.name
        assert np.array_equal(f["Mask"], ref["Mask"]), s.name
        g = np.s_[..., G:-G, G:-G]
        for n in ("Eta", "U", "W"):
            assert rel(f[n][g], ref[n][g]) < 1e-8, (s.name, n, rel(f[n][g], ref[n][g]))


def test_batch_viscous_members_mixed_bit_identical(gpu):
    """The viscous members in the mixed form (FP32 storage of the 3-D state,
    FP64 arithmetic), hydrostatic, replayed from HIP graphs: every member
    equals its mixed run alone bit for bit."""
    sc = viscous_members(False)
    with Batch(sc, precision=capi.NHW_MIXED) as b:
        b.step(10)
        got = [(member_state(m, s), m.time()) for m, s in zip(b.members, sc)]
    for s, (f, t) in zip(sc, got):
        with Model(s, precision=capi.NHW_MIXED) as m:
            m.step(10)
            ref, tref = member_state(m, s), m.time()
        assert t == tref, s.name
        bad = [n for n in ref if not np.array_equal(f[n], ref[n])]
        assert bad == [], (s.name, bad)
