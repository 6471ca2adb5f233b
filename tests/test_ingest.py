"""Scenario ingest (nhwave_amd/ingest.py): a run directory written in the
reference's formats reads back to the same scenario, the input.txt grammar
follows src/mod_util.F:273-378 (comments, first occurrence, 80-column
records, type rules), the list-directed array reader handles records that
split rows, repeat counts and commas, the CELL_GRID depth averaging with
permanent dry points follows src/initialize.F:1770-1804, and options off the
GPU path are refused.  CPU only."""
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
]
SCALARS = ("mglob", "nglob", "kglob", "dx", "dy", "cfl", "min_dep", "dt_ini", "dt_min", "dt_max",
           "froude_cap", "visc", "slat", "wind_u", "wind_v", "time_order", "non_hydro", "bc", "itmax",
           "tol", "ivgrd", "grd_r", "steps")


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


@pytest.mark.parametrize("line,what", [("VISCOUS_FLOW = T", "VISCOUS_FLOW"), ("WAVEMAKER = LEF_SOL", "WAVEMAKER"),
                                       ("BC_X0 = 3", "BC_X0"), ("PERIODIC_X = T", "PERIODIC_X")])
def test_refuses_options_off_the_path(tmp_path, line, what):
    s = S.solitary_channel(mglob=8, nglob=4, kglob=3, steps=2)
    S.write_reference_inputs(s, tmp_path)
    key = line.split("=")[0].strip()
    lines = [ln for ln in (tmp_path / "input.txt").read_text().splitlines() if ln.split("=")[0].strip() != key]
    (tmp_path / "input.txt").write_text(line + "\n" + "\n".join(lines) + "\n")
    with pytest.raises(I.InputError, match=what):
        I.scenario_from_run_dir(tmp_path)
