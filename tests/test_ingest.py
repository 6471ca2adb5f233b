"""Scenario ingest (nhwave_amd/ingest.py): a run directory written in the
reference's formats reads back to the same scenario, the input.txt grammar
follows src/mod_util.F:273-378 (comments, first occurrence, 80-column
records, type rules), the list-directed array reader handles records that
split rows, repeat counts and commas, the CELL_GRID depth averaging with
permanent dry points follows src/initialize.F:1770-1804, and options off the
GPU path are refused.  CPU only."""
from pathlib import Path

import numpy as np
import pytest

from nhwave_amd import ingest as I
from nhwave_amd import scenario as S

GENERATORS = [
    ("solitary_channel", dict(mglob=40, nglob=6, steps=7)),
    ("bumps_beach", dict(n=21, m=17, kglob=5, dx=50.0, steps=9, tol=1e-9)),
    ("beach_runup", dict(mglob=30, nglob=5, steps=11, non_hydro=False)),
    ("random_scenario", dict(seed=4, mglob=13, nglob=9, kglob=3, steps=3)),
    ("solitary_channel", dict(mglob=32, nglob=6, steps=5, time_order=3, non_hydro=False)),
    ("wave_flume", dict(mglob=30, nglob=2, wavemaker="LEF_LIN", amp=0.03, per=1.7, theta=5.0,
                        sponge=(0.0, 0.2, 0.0, 0.1), cur=0.1, tramp=0.5, steps=4)),
    ("wave_flume", dict(mglob=20, wavemaker="LEF_SOL", amp=0.05, depth=0.3, bottom="beach", steps=4)),
]
SCALARS = ("mglob", "nglob", "kglob", "dx", "dy", "cfl", "min_dep", "dt_ini", "dt_min", "dt_max",
           "froude_cap", "visc", "slat", "wind_u", "wind_v", "time_order", "non_hydro", "bc", "itmax",
           "tol", "ivgrd", "grd_r", "steps", "wavemaker", "amp_wave", "per_wave", "dep_wave", "theta_wave",
           "cur_wave", "tramp", "sponge")


@pytest.mark.parametrize("gen,kw", GENERATORS)
def test_run_dir_round_trip(tmp_path, gen, kw):
    s = getattr(S, gen)(**kw)
    S.write_reference_inputs(s, tmp_path)
    r = I.scenario_from_run_dir(tmp_path)
    for a in SCALARS:
        assert getattr(r, a) == getattr(s, a), a
    assert np.array_equal(r.hcg, s.hcg) and np.array_equal(r.etag, s.etag)
    for a in ("ug", "vg", "wg"):
        want = getattr(s, a)
        want = np.zeros((s.kglob, s.nglob, s.mglob)) if want is None else want
        assert np.array_equal(getattr(r, a), want), a


def test_grammar(tmp_path):
    p = tmp_path / "input.txt"
    p.write_text(
        "! a comment line\n"
        "  Mglob   =   64   ! trailing comment\n"
        "Mglob = 99\n"                        # the first occurrence wins
        "DX = 0.5\n"
        "NOPE = 1\n"
        "TITLE = run one\n"
        "FLAG = T\n"
        "BADF = 3\n"
        "LONG = 1." + "0" * 70 + "7\n"        # past column 80: cut
        "DEC = 1.5d-3\n"
    )
    inp = I.Input(p)
    assert inp.int("Mglob") == 64
    assert inp.float("DX") == 0.5
    assert inp.str("TITLE") == "run one" and inp.bool("FLAG") is True
    assert inp.float("LONG") == 1.0
    with pytest.raises(I.InputError, match="should be float but is string"):
        inp.float("DEC")  # 'd' is not among Check_Arg_Type's number characters
    with pytest.raises(I.InputError, match="should be float"):
        inp.float("BADF")
    with pytest.raises(I.InputError, match="should be integer"):
        inp.int("DX")
    with pytest.raises(I.InputError, match="does not exist"):
        inp.float("MISSING")
    assert inp.float("MISSING", 2.5) == 2.5
    assert [I.arg_type(v) for v in ("T", "F", "TF", "1.", "12", "1e3", "abc", "1.0E+5")] == [
        "logical", "logical", "string", "float", "integer", "integer", "string", "float"]


def test_list_directed_rows(tmp_path):
    p = tmp_path / "a.txt"
    # row 1 over two records; row 2 with a repeat count and commas; the rest
    # of row 2's record is skipped by the next READ; a d exponent
    p.write_text("1 2\n3\n2*4.5, 6 99 99\n7 8 9.0d0\n")
    a = I.read_rows(p, 3, 3)
    assert np.array_equal(a, [[1, 2, 3], [4.5, 4.5, 6], [7, 8, 9]])
    with pytest.raises(I.InputError, match="end of file"):
        I.read_rows(p, 4, 3)


def test_cell_grid_depth(tmp_path):
    s = S.solitary_channel(mglob=6, nglob=4, kglob=3, steps=2)
    S.write_reference_inputs(s, tmp_path)
    txt = (tmp_path / "input.txt").read_text().replace("DEPTH_TYPE = CELL_CENTER", "DEPTH_TYPE = CELL_GRID")
    (tmp_path / "input.txt").write_text(txt)
    hg = np.arange(35, dtype=float).reshape(5, 7) + 1.0
    hg[2, 3] = -2000.0  # a permanent dry point: left out of its four cells
    (tmp_path / "depth.txt").write_text("\n".join(" ".join(f"{v:.17e}" for v in row) for row in hg) + "\n")
    r = I.scenario_from_run_dir(tmp_path)
    mp = (hg >= -1000.0).astype(float)
    eps = float(np.float32(1e-16))
    for j in range(4):
        for i in range(6):
            num = (hg[j, i] * mp[j, i] + hg[j, i + 1] * mp[j, i + 1] + hg[j + 1, i] * mp[j + 1, i]
                   + hg[j + 1, i + 1] * mp[j + 1, i + 1])
            den = mp[j, i] + mp[j, i + 1] + mp[j + 1, i] + mp[j + 1, i + 1] + eps
            assert r.hcg[j, i] == num / den
    assert r.hcg[1, 2] == (hg[1, 2] + hg[1, 3] + hg[2, 2]) / (3.0 + eps)


@pytest.mark.parametrize("line,what", [("WAVEMAKER = LEF_SPC", "WAVEMAKER"), ("WAVEMAKER = FLUX_LR", "WAVEMAKER"),
                                       ("BC_X0 = 3", "BC_X0"), ("PERIODIC_X = T", "PERIODIC_X")])
def test_refuses_options_off_the_path(tmp_path, line, what):
    s = S.solitary_channel(mglob=8, nglob=4, kglob=3, steps=2)
    S.write_reference_inputs(s, tmp_path)
    key = line.split("=")[0].strip()
    lines = [ln for ln in (tmp_path / "input.txt").read_text().splitlines() if ln.split("=")[0].strip() != key]
    (tmp_path / "input.txt").write_text(line + "\n" + "\n".join(lines) + "\n")
    with pytest.raises(I.InputError, match=what):
        I.scenario_from_run_dir(tmp_path)


def test_viscous_keys(tmp_path):
    """VISCOUS_FLOW = T with IVTURB = IHTURB = 1 is read with its constants,
    IVTURB = 10 with IHTURB >= 10 as the k-epsilon model (Kglob >= 3); the
    other turbulence models are refused."""
    from dataclasses import replace
    s = replace(S.solitary_channel(mglob=8, nglob=4, kglob=3, steps=2), viscous=True, visc=2e-3, cvs=0.03,
                chs=0.05, schmidt=0.7, visc_number=0.2, rng=False)
    S.write_reference_inputs(s, tmp_path)
    r = I.scenario_from_run_dir(tmp_path)
    assert r.viscous and (r.visc, r.cvs, r.chs, r.schmidt, r.visc_number, r.rng) == (2e-3, 0.03, 0.05, 0.7, 0.2, False)
    assert r.turb == 1
    base = (tmp_path / "input.txt").read_text()
    (tmp_path / "input.txt").write_text(base.replace("IVTURB = 1", "IVTURB = 10").replace("IHTURB = 1", "IHTURB = 10"))
    assert I.scenario_from_run_dir(tmp_path).turb == 10
    for iv, ih, what in ((10, 1, "IVTURB = 10, IHTURB = 1"), (20, 20, "IVTURB = 20"), (1, 10, "IHTURB = 10")):
        txt = base.replace("IVTURB = 1", f"IVTURB = {iv}").replace("IHTURB = 1", f"IHTURB = {ih}")
        (tmp_path / "input.txt").write_text(txt)
        with pytest.raises(I.InputError, match=what):
            I.scenario_from_run_dir(tmp_path)
    txt = base.replace("IVTURB = 1", "IVTURB = 10").replace("IHTURB = 1", "IHTURB = 10").replace(
        "Kglob = 3", "Kglob = 2")
    (tmp_path / "input.txt").write_text(txt)
    with pytest.raises(I.InputError, match="k-epsilon"):
        I.scenario_from_run_dir(tmp_path)
    (tmp_path / "input.txt").write_text(base.replace("VISCOUS_FLOW = T", "VISCOUS_FLOW = F").replace(
        "IVTURB = 1", "IVTURB = 20"))
    assert not I.scenario_from_run_dir(tmp_path).viscous


def test_barotropic_off_is_taken(tmp_path):
    """BAROTROPIC = F runs as T (the density is Rho0 without salinity); the
    salinity input that would make it matter is refused."""
    s = S.solitary_channel(mglob=8, nglob=4, kglob=3, steps=2)
    S.write_reference_inputs(s, tmp_path)
    base = (tmp_path / "input.txt").read_text()
    assert "BAROTROPIC = T" in base
    (tmp_path / "input.txt").write_text(base.replace("BAROTROPIC = T", "BAROTROPIC = F"))
    r = I.scenario_from_run_dir(tmp_path)
    assert (r.mglob, r.nglob, r.kglob) == (8, 4, 3)
    (tmp_path / "input.txt").write_text(base.replace("BAROTROPIC = T", "BAROTROPIC = F").replace(
        "INITIAL_SALI = F", "INITIAL_SALI = T"))
    with pytest.raises(I.InputError, match="INITIAL_SALI"):
        I.scenario_from_run_dir(tmp_path)


def test_bottom_friction_keys(tmp_path):
    """BC_Z0 = 5 reads Ibot, Cd0 and Zob; BC_Z0 = 2 needs none of them."""
    from dataclasses import replace
    s = replace(S.solitary_channel(mglob=8, nglob=4, kglob=3, steps=2), bc=(1, 1, 1, 1, 5), ibot=2, cd0=0.004,
                zob=2e-3)
    S.write_reference_inputs(s, tmp_path)
    r = I.scenario_from_run_dir(tmp_path)
    assert r.bc[4] == 5 and (r.ibot, r.cd0, r.zob) == (2, 0.004, 2e-3)
    txt = (tmp_path / "input.txt").read_text().replace("BC_Z0 = 5", "BC_Z0 = 2")
    (tmp_path / "input.txt").write_text(txt)
    assert I.scenario_from_run_dir(tmp_path).bc[4] == 2


def test_total_time_and_process_grid(tmp_path):
    s = S.solitary_channel(mglob=10, nglob=4, kglob=3, steps=50)
    s.total_time = 0.75
    S.write_reference_inputs(s, tmp_path)
    txt = (tmp_path / "input.txt").read_text()
    # PX/PY (the MPI process grid) do not change the run; latin-1 comment text reads
    (tmp_path / "input.txt").write_bytes(b"! \xa3 note\nPX = 4\nPY = 2\n" + txt.encode())
    r = I.scenario_from_run_dir(tmp_path)
    assert r.total_time == 0.75 and r.steps == 50


def simulate(t, dt_old, total, dt_max, growth_seq):
    """TIME after each step of the reference's loop, for a dt sequence
    dt_n = min(g_n dt_{n-1}, dt_max) with g_n <= float32(1.05)."""
    out = []
    d = dt_old
    for g in growth_seq:
        if not t < total:
            break
        d = min(d * g, dt_max)
        t = t + d
        out.append(t)
    return out


@pytest.mark.parametrize("seed", range(6))
def test_steps_before_total_time_never_overshoots(seed):
    from nhwave_amd.serve import steps_before_total_time
    rng = np.random.default_rng(seed)
    g_max = float(np.float32(1.05))
    for _ in range(200):
        t = float(rng.uniform(0, 5))
        dt_old = float(10 ** rng.uniform(-4, -1))
        dt_max = float(rng.choice([dt_old * rng.uniform(0.5, 3), 1e9]))
        total = t + float(rng.uniform(0, 40)) * dt_old
        # growth at the limit, below it, or dt held at DT_MAX: the bound's
        # tight cases
        kind = rng.integers(3)
        gs = [g_max] * 400 if kind == 0 else list(rng.uniform(0.3, g_max, 400)) if kind == 1 else [1.0] * 400
        times = simulate(t, dt_old, total, dt_max, gs)
        n = steps_before_total_time(t, dt_old, total, dt_max, 400)
        assert n <= len(times)                          # the run takes at least n steps
        assert all(x < total for x in times[:n - 1])    # and none of the first n-1 ends it
        if kind == 0 and dt_max > 1e8:
            assert n >= len(times) - 1                  # tight when dt grows at the limit


def test_steps_before_total_time_edges():
    from nhwave_amd.serve import steps_before_total_time
    assert steps_before_total_time(1.0, 0.1, 1.0, 1.0, 10) == 0     # already at the end
    assert steps_before_total_time(0.0, 0.1, 1e-12, 1.0, 10) == 1   # one step ends it
    assert steps_before_total_time(0.0, 0.1, 1e9, 1.0, 7) == 7      # the cap


REF_EXAMPLES = Path("/root/reference/examples")


@pytest.mark.skipif(not REF_EXAMPLES.is_dir(), reason="reference examples not present")
def test_reference_example_inputs_parse():
    """Every input.txt the reference ships parses under its own grammar and
    is either taken or refused naming the options off the GPU path."""
    seen = 0
    for p in sorted(REF_EXAMPLES.rglob("input.txt")):
        inp = I.Input(p)
        assert inp.int("Mglob") > 0 and inp.float("TOTAL_TIME") > 0.0
        try:
            I._check_supported(inp)
        except I.InputError as e:
            assert "off the GPU path" in str(e)
        seen += 1
    assert seen >= 10


def test_flux_lr_flume_ingests(tmp_path):
    """WAVEMAKER = FLUX_LR is read with its open boundaries BC_X0 = BC_Xn = 3,
    Zob and TRAMP; FLUX_LR with a closed east wall is refused."""
    s = S.steady_flume(mglob=24, nglob=2, kglob=3, steps=2)
    S.write_reference_inputs(s, tmp_path)
    r = I.scenario_from_run_dir(tmp_path)
    assert r.wavemaker == "FLUX_LR" and r.bc[:2] == (3, 3) and r.zob == s.zob and r.tramp == s.tramp
    p = tmp_path / "input.txt"
    p.write_text("\n".join("BC_Xn = 1" if ln.split("=")[0].strip() == "BC_Xn" else ln
                           for ln in p.read_text().splitlines()) + "\n")
    with pytest.raises(I.InputError, match="FLUX_LR"):
        I.scenario_from_run_dir(tmp_path)
