"""Two of the reference's own example runs on the GPU path, read from run
directories rebuilt from fixtures of their input.txt keys and depth arrays
(tests/golden/example_<name>.json, _depth.npy) through the ingest module:
hydrostatic states against the reference build's (to the example's own
TOTAL_TIME for the solitary wave), the non-hydrostatic example as shipped
against the oracle's direct solve, and the shipped non-hydrostatic example
run to its end."""
import json

import numpy as np
import pytest

import oracle as O
from nhw_testutil import GOLD, example_run_dir, hydro_overrides, interior_state, rel_errors
from nhwave_amd import ingest as I
from nhwave_amd.model import Model

pytestmark = pytest.mark.gpu

INDEX = json.loads((GOLD / "examples_index.json").read_text())
FIELDS = ("Eta", "D", "Mask", "U", "V", "W", "DU", "DV", "DW", "P", "Omega")


def golden(name, st):
    data = np.load(GOLD / f"example_{name}_hydro.npz")
    return {k: data[f"s{st}_{k}"] for k in FIELDS}


@pytest.mark.parametrize("name", sorted(INDEX))
def test_gpu_example_hydrostatic_vs_reference(gpu, tmp_path, name):
    """Masks exact; the state within 1e-12 of the reference (the wavemaker's
    functions of TIME come from the device libm) over the first 300 steps,
    and within 1e-10 after the solitary wave has broken and run up the
    slope at the example's TOTAL_TIME (step 1464)."""
    meta = INDEX[name]
    to_end = meta.get("to_total_time", False)
    example_run_dir(name, tmp_path, hydro_overrides(None if to_end else max(meta["steps"])))
    s = I.scenario_from_run_dir(tmp_path)
    with Model(s) as m:
        for st in meta["steps"]:
            last = st == max(meta["steps"])
            if last and to_end:
                m.run()             # the reference's loop: to TOTAL_TIME
            else:
                m.step(st - m.time()[2])
            t, dt, rs = m.time()
            assert rs == st, (name, rs, st)
            ours = interior_state(s, m.download(*FIELDS).__getitem__)
            ref = golden(name, st)
            assert np.array_equal(ours["Mask"], ref["Mask"]), (name, st)
            err = rel_errors(ours, ref)
            tol = 1e-10 if st > 300 else 1e-12
            bad = {f: e for f, e in err.items() if f not in ("P", "Omega") and e > tol}
            assert bad == {}, (name, st, err)


def test_gpu_example_nonhydrostatic_vs_oracle(gpu, tmp_path):
    """examples/solitary_wave_breaking as shipped (non-hydrostatic, one row:
    beyond the serial reference's NSPCG), solved to TOL 1e-12, against the
    oracle's direct solve over 12 steps."""
    example_run_dir("solitary_wave_breaking", tmp_path, {"TOL": "1.0e-12", "NSTAT": "0"})
    s = I.scenario_from_run_dir(tmp_path)
    o = O.Oracle(s)
    with Model(s) as m:
        for st in (1, 12):
            o.step(st - o.time()[2])
            m.step(st - m.time()[2])
            ours = interior_state(s, m.download(*FIELDS).__getitem__)
            ref = interior_state(s, o.field)
            assert np.array_equal(ours["Mask"], ref["Mask"])
            err = rel_errors(ours, ref)
            assert all(err[f] < 1e-9 for f in ("Eta", "D", "U", "W", "DU", "DW")), (st, err)
            assert m.stats().not_converged == 0


def test_gpu_example_nonhydrostatic_to_total_time(gpu, tmp_path):
    """examples/solitary_wave_breaking as shipped, run to its TOTAL_TIME of
    8 s: every Poisson solve converges to the example's TOL, the state stays
    finite, the wave enters (the west inflow) and runs up the slope (the
    shoreline moves: cells dry out or flood)."""
    example_run_dir("solitary_wave_breaking", tmp_path, {"NSTAT": "0"})
    s = I.scenario_from_run_dir(tmp_path)
    with Model(s) as m:
        mask0 = m.download("Mask")["Mask"].copy()
        taken = m.run()
        t, dt, rs = m.time()
        f = m.download("Eta", "U", "W", "Mask")
        st = m.stats()
    assert t >= 8.0 and rs == taken and st.not_converged == 0 and st.solves == 2 * taken
    for n in ("Eta", "U", "W"):
        assert np.isfinite(f[n]).all(), n
    assert not np.array_equal(f["Mask"], mask0)
    print(f"solitary_wave_breaking (non-hydrostatic) to t = {t:.4f} s in {taken} steps, "
          f"{st.total_iters / st.solves:.1f} Krylov iterations per solve")
