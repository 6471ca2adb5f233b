"""The product's host-side set-up (nhwave_amd.scenario) restates
generate_grid/read_bathymetry; it must equal the oracle's restatement bit
for bit, and the reference input files it writes must be accepted by the
reference parser's grammar."""
import numpy as np
import pytest

import oracle as O
from nhwave_amd import scenario as S


@pytest.mark.parametrize("ivgrd,k", [(1, 3), (1, 10), (2, 6)])
def test_vertical_grid_matches_oracle(ivgrd, k):
    s = S.random_scenario(1, mglob=5, nglob=4, kglob=k).with_(ivgrd=ivgrd, grd_r=1.1)
    o = O.Oracle(s)
    dsig, sig, sigc = S.vertical_grid(k, ivgrd, 1.1)
    assert np.array_equal(dsig, o.field("dsig"))
    assert np.array_equal(sig, o.field("sig"))
    assert np.array_equal(sigc, o.field("sigc"))


@pytest.mark.parametrize("seed", [1, 2])
def test_bathymetry_fields_match_oracle(seed):
    s = S.random_scenario(seed, mglob=11, nglob=7, kglob=3)
    o = O.Oracle(s)
    f = S.bathymetry_fields(s)
    for name in ("Hc", "Hfx", "Hfy", "DelxH", "DelyH", "Mask_Struct"):
        assert np.array_equal(f[name], o.field(name)), name


def test_reference_input_grammar(tmp_path):
    """Every float value carries a '.', lines fit 80 characters, and float
    text round-trips to the same binary64 value (src/mod_util.F:273-415)."""
    s = S.solitary_channel(mglob=16, nglob=4, kglob=3, steps=2)
    S.write_reference_inputs(s, tmp_path)
    for line in (tmp_path / "input.txt").read_text().splitlines():
        assert len(line) <= 80
        key, val = (t.strip() for t in line.split("=", 1))
        if key in ("DX", "DY", "TOL", "CFL", "MinDep"):
            assert "." in val and float(val) == getattr(s, {"DX": "dx", "DY": "dy", "TOL": "tol",
                                                            "CFL": "cfl", "MinDep": "min_dep"}[key])
    depth = np.loadtxt(tmp_path / "depth.txt").ravel()
    assert np.array_equal(depth, s.hcg.ravel())
    uvw = np.loadtxt(tmp_path / "uvw0.txt").ravel()
    assert np.array_equal(uvw[: s.cells], s.ug.ravel())


@pytest.mark.skipif(not O.REF_BIN.exists(), reason="reference build absent")
def test_reference_accepts_inputs(tmp_path):
    s = S.beach_runup(mglob=24, nglob=4, steps=2, non_hydro=False)
    wall, dumps = O.run_reference(s, tmp_path, [2])
    assert 2 in dumps and dumps[2]["step"] == 2
