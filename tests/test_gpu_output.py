"""Output and restart path: device `statistics` against the values the
reference itself prints (tests/golden/stats_*.json, E12.4), against the
reference's formulas on the downloaded state, and against closed forms;
a checkpoint continues a run bit for bit; an asynchronous snapshot holds the
state of the moment it was taken while stepping goes on."""
import json
from pathlib import Path

import numpy as np
import pytest

from nhwave_amd import capi
from nhwave_amd import scenario as S
from nhwave_amd.model import Model

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
G = S.NGHOST
KEYS = ("mass_volume", "energy", "max_eta", "min_eta", "max_u", "max_v", "max_w")
STATE = ("Eta", "D", "Mask", "U", "V", "W", "DU", "DV", "DW", "P", "Omega", "Hfx0", "Hfy0")


def printed_close(ours: float, printed: float, floor: float = 0.0) -> bool:
    """E12.4 keeps 4 significant digits: half a unit in the last one.  `floor`
    covers quantities at round-off level (a transverse velocity in a 1-D
    channel), where the two Poisson solvers' last bits decide the value."""
    return abs(ours - printed) <= 6e-4 * abs(printed) + floor + 1e-300


@pytest.mark.parametrize("case", ["beach_hydro", "bumps48_nh", "solitary512_nh"])
def test_statistics_vs_reference_log(gpu, case):
    rec = json.loads((GOLDEN / f"stats_{case}.json").read_text())
    s = getattr(S, rec["generator"])(**rec["kwargs"])
    with Model(s) as m:
        for n, ref in enumerate(rec["steps"], start=1):
            m.step(1)
            st = m.statistics()
            t, dt, _ = m.time()
            assert printed_close(t, ref["time"]) and printed_close(dt, ref["dt"]), (case, n)
            vel = 1e-8 * max(ref["max_u"], ref["max_v"], ref["max_w"])
            eta = 1e-8 * max(abs(ref["max_eta"]), abs(ref["min_eta"]))
            floor = dict(max_u=vel, max_v=vel, max_w=vel, max_eta=eta, min_eta=eta)
            bad = [k for k in KEYS if not printed_close(st[k], ref[k], floor.get(k, 0.0))]
            assert bad == [], (case, n, {k: (st[k], ref[k]) for k in bad})


def reference_formula(s, f, dsig, sigc):
    """statistics, src/nhwave.F:3137-3154, on the downloaded interior (the
    reference's serial loop order: j, i, then k)."""
    rho0, grav = 1000.0, float(np.float32(9.81))
    D = f["D"][G:-G, G:-G]
    Hc = f["Hc"][G:-G, G:-G]
    mass = 0.0
    energy = 0.0
    for j in range(s.nglob):
        for i in range(s.mglob):
            mass += D[j, i] * s.dx * s.dy
            for k in range(s.kglob):
                cm = rho0 * dsig[k] * D[j, i] * s.dx * s.dy
                u, v, w = (f[n][G + k, G + j, G + i] for n in ("U", "V", "W"))
                energy += cm * grav * (D[j, i] * sigc[k] - Hc[j, i]) + 0.5 * cm * (u * u + v * v + w * w)
    inner = np.s_[G:-G, G:-G, G:-G]
    return dict(mass_volume=mass, energy=energy, max_eta=f["Eta"][G:-G, G:-G].max(),
                min_eta=f["Eta"][G:-G, G:-G].min(), max_u=np.abs(f["U"][inner]).max(),
                max_v=np.abs(f["V"][inner]).max(), max_w=np.abs(f["W"][inner]).max())


@pytest.mark.parametrize("precision", [capi.NHW_FP64, capi.NHW_MIXED])
def test_statistics_vs_formula(gpu, precision):
    s = S.bumps_beach(n=24, m=31, kglob=5, dx=100.0, steps=10, tol=1e-10)
    dsig, _, sigc = S.vertical_grid(s.kglob)
    with Model(s, precision=precision) as m:
        m.step(6)
        st = m.statistics()
        f = m.download("D", "Hc", "Eta", "U", "V", "W")
    ref = reference_formula(s, f, dsig[G:G + s.kglob], sigc[G:G + s.kglob])
    for k in KEYS:
        tol = 1e-12 if k in ("mass_volume", "energy") else 0.0
        assert abs(st[k] - ref[k]) <= tol * abs(ref[k]), (k, st[k], ref[k])


def test_statistics_lake_at_rest_closed_form(gpu):
    """Flat bottom h, still water: MassVolume = M N h dx dy and
    Energy = Rho0 g dx dy M N h^2 (sum_k dsig_k sigc_k - 1) with U = V = W = 0
    (sum_k dsig_k sigc_k = 1/2 for the uniform sigma grid)."""
    h = 0.5
    s = S.solitary_channel(mglob=20, nglob=14, kglob=4, dx=0.05, depth=h, amp_ratio=0.0, steps=5)
    dsig, _, sigc = S.vertical_grid(s.kglob)
    with Model(s) as m:
        st = m.statistics()
    k0 = slice(G, G + s.kglob)
    mass = s.mglob * s.nglob * h * s.dx * s.dy
    energy = 1000.0 * float(np.float32(9.81)) * s.dx * s.dy * s.mglob * s.nglob * h * h * (
        float(np.sum(dsig[k0] * sigc[k0])) - 1.0)
    assert abs(st["mass_volume"] - mass) <= 1e-13 * mass
    assert abs(st["energy"] - energy) <= 1e-12 * abs(energy)
    assert st["max_u"] == st["max_v"] == st["max_w"] == 0.0
    assert st["max_eta"] == st["min_eta"] == 0.0


@pytest.mark.parametrize("precision,non_hydro", [(capi.NHW_FP64, True), (capi.NHW_MIXED, True),
                                                  (capi.NHW_FP64, False)])
def test_checkpoint_continues_bit_for_bit(gpu, tmp_path, precision, non_hydro):
    s = S.bumps_beach(n=40, m=33, kglob=5, dx=100.0, steps=20, tol=1e-10, non_hydro=non_hydro)
    path = tmp_path / "run.ckpt"
    with Model(s, precision=precision) as a:
        a.step(5)
        a.checkpoint_save(path)
        a.step(5)
        want, t_want = a.download(*STATE), a.time()
    with Model(s, precision=precision) as b:
        b.step(1)  # a different state, overwritten by the load
        b.checkpoint_load(path)
        b.step(5)
        got, t_got = b.download(*STATE), b.time()
    assert t_got == t_want
    bad = [n for n in STATE if not np.array_equal(got[n], want[n])]
    assert bad == [], bad


def test_checkpoint_refuses_other_grid(gpu, tmp_path):
    a_s = S.bumps_beach(n=16, kglob=4, dx=100.0, steps=5)
    b_s = S.bumps_beach(n=18, kglob=4, dx=100.0, steps=5)
    path = tmp_path / "a.ckpt"
    with Model(a_s) as a:
        a.checkpoint_save(path)
    with Model(b_s) as b:
        with pytest.raises(RuntimeError, match="differ"):
            b.checkpoint_load(path)


@pytest.mark.parametrize("change", [dict(dx=90.0), dict(bc=[1, 1, 1, 1, 5], cd0=0.003),
                                    dict(cfl=0.4), dict(tol=1e-9), dict(min_dep=0.02)])
def test_checkpoint_refuses_other_physics(gpu, tmp_path, change):
    """Same grid, other physics: the load names the differing setting and
    leaves the context as it was."""
    a_s = S.bumps_beach(n=16, kglob=4, dx=100.0, steps=5, tol=1e-10)
    b_s = a_s.with_(**change)
    path = tmp_path / "a.ckpt"
    with Model(a_s) as a:
        a.step(2)
        a.checkpoint_save(path)
    with Model(b_s) as b:
        before = b.download("Eta", "U")
        with pytest.raises(RuntimeError, match="differs from the context"):
            b.checkpoint_load(path)
        after = b.download("Eta", "U")
    assert all(np.array_equal(before[n], after[n]) for n in before)


@pytest.mark.parametrize("precision", [capi.NHW_FP64, capi.NHW_MIXED])
def test_snapshot_holds_state_while_stepping(gpu, precision):
    s = S.bumps_beach(n=48, m=40, kglob=6, dx=100.0, steps=20, tol=1e-10)
    names = ("Eta", "Mask", "U", "W", "P", "Omega")
    with Model(s, precision=precision) as a:
        a.step(4)
        snap = a.snapshot_begin(*names)
        a.step(3)  # queued behind the snapshot copies; the drain overlaps them
        a.snapshot_wait()
        later = a.download("Eta")
    with Model(s, precision=precision) as b:
        b.step(4)
        want = b.download(*names)
    bad = [n for n in names if not np.array_equal(snap[n], want[n])]
    assert bad == [], bad
    assert not np.array_equal(later["Eta"], want["Eta"])


@pytest.mark.parametrize("non_hydro", [False, True])
def test_checkpoint_continues_forced_run(gpu, tmp_path, non_hydro):
    """A wavemaker and sponge run restarts bit for bit: the inflow values are
    a function of the device TIME, which the checkpoint carries."""
    s = S.wave_flume(mglob=120, nglob=3, wavemaker="LEF_LIN", amp=0.02, per=1.5, bottom="bar",
                     sponge=(0.0, 0.5, 0.0, 0.0), steps=30, tol=1e-10, non_hydro=non_hydro)
    path = tmp_path / "flume.ckpt"
    with Model(s) as a:
        a.step(7)
        a.checkpoint_save(path)
        a.step(6)
        want, t_want = a.download(*STATE), a.time()
    with Model(s) as b:
        b.checkpoint_load(path)
        b.step(6)
        got, t_got = b.download(*STATE), b.time()
    assert t_got == t_want
    bad = [n for n in STATE if not np.array_equal(got[n], want[n])]
    assert bad == [], bad
