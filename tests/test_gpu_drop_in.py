"""The reference's own Fortran driver, unmodified, linked against
libnhwave_amd through the ISO_C_BINDING shim (nhwave_amd/fortran/): the
drop-in boundary exercised end to end.  The driver reads the same input.txt
and data files as the reference executable that wrote the golden vectors;
the shim's full-precision snapshots after chosen steps are compared with
those vectors."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from nhw_testutil import bit_equal, golden_case, golden_state, interior_state, rel_errors
from nhwave_amd import scenario as S

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parents[1] / "nhwave_amd" / "fortran" / "nhwave_gpu"
CASES = ["solitary512_hydro", "beach_hydro", "solitary_rk3_hydro", "solitary512_tight",
         "beach_nh_tight", "bumps48_nh_tight", "flume_lin_bar_hydro", "flume_sol_beach_nh_tight",
         "visc_bumps48_hydro", "visc_random3_nh_tight", "bot5_visc_bumps48_hydro", "bot2_visc_random7_hydro",
         "ke_random5_hydro", "ke_random7_nh_tight", "ke_random5_baroclinic_hydro", "flume_con_beach_hydro",
         "flume_con_bar_nh_tight", "int_irr_basin_hydro", "int_irr_basin_nh_tight", "flux_lr_steady_hydro",
         "flux_lr_steady_nh_tight", "wall5_visc_random1_hydro", "ke_wall5_random3_nh_tight"]


def run_drop_in(s, steps, run_dir: Path, solver: str = "mg", overrides: dict | None = None):
    S.write_reference_inputs(s, run_dir)
    if overrides:   # input.txt keys the golden's reference run was given
        p = run_dir / "input.txt"
        lines = [f"{k} = {overrides[k]}" if (k := ln.split("=")[0].strip()) in overrides else ln
                 for ln in p.read_text().splitlines()]
        p.write_text("\n".join(lines) + "\n")
    env = dict(os.environ, NHW_SNAPSHOT_STEPS=",".join(str(k) for k in steps))
    if solver == "line":
        env["NHW_SOLVER"] = "line"
    res = subprocess.run([str(BIN)], cwd=run_dir, env=env, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0 and "Normal Termination" in res.stdout, res.stdout[-3000:] + res.stderr[-3000:]
    return {st: S.read_dump(run_dir / f"nhw_snap_{st:05d}.bin") for st in steps}


@pytest.mark.parametrize("name", CASES)
def test_drop_in_driver_vs_reference(gpu, name, tmp_path):
    if not BIN.exists():
        pytest.fail(f"{BIN} missing: run nhwave_amd/fortran/build_drop_in.sh where the reference sources are")
    s, meta, data = golden_case(name)
    snaps = run_drop_in(s, meta["steps"], tmp_path, overrides=meta.get("input_overrides"))
    # wavemaker runs: the device libm's last bit (see test_gpu_parity)
    hydro = name.endswith("_hydro") and not name.startswith(("flume_", "int_irr_", "flux_lr_"))
    tol = 1e-12 if name.endswith("_hydro") else 1e-10
    for st in meta["steps"]:
        d = snaps[st]
        assert d["step"] == st
        ours = interior_state(s, d.__getitem__)
        ref = golden_state(data, st)
        assert np.array_equal(ours["Mask"], ref["Mask"]), (name, st)
        if hydro:
            # the explicit path is bit-exact, and so are the driver's clock
            assert d["TIME"] == meta["time"][str(st)] and d["dt"] == meta["dt"][str(st)]
            assert bit_equal(ours, ref) == [], (name, st, rel_errors(ours, ref))
        else:
            err = rel_errors(ours, ref)
            for f in ("Eta", "D", "U", "V", "W", "DU", "DV", "DW"):
                assert err[f] < tol, (name, st, f, err)
            assert err["P"] < 1e-8 and err["Omega"] < 1e-8, (name, st, err)


def test_drop_in_runs_shipped_example_nonhydrostatic(gpu, tmp_path):
    """examples/solitary_wave_breaking as shipped (non-hydrostatic, one row,
    a solitary wave from the west wavemaker), run by the reference's own
    driver linked to the GPU path: the serial reference's NSPCG solve cannot
    run it (NaN on one-row grids); with the Poisson solve on the GPU it runs,
    and its snapshots match the oracle's direct solve (TOL 1e-12)."""
    import oracle as O
    from nhw_testutil import example_run_dir
    from nhwave_amd import ingest as I
    if not BIN.exists():
        pytest.fail(f"{BIN} missing: run nhwave_amd/fortran/build_drop_in.sh where the reference sources are")
    example_run_dir("solitary_wave_breaking", tmp_path,
                    {"TOL": "1.0e-12", "NSTAT": "0", "SIM_STEPS": "12", "PLOT_INTV": "1.0e9",
                     "SCREEN_INTV": "1.0e9"})
    env = dict(os.environ, NHW_SNAPSHOT_STEPS="1,12")
    res = subprocess.run([str(BIN)], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0 and "Normal Termination" in res.stdout, res.stdout[-3000:] + res.stderr[-3000:]
    s = I.scenario_from_run_dir(tmp_path)
    o = O.Oracle(s)
    for st in (1, 12):
        d = S.read_dump(tmp_path / f"nhw_snap_{st:05d}.bin")
        o.step(st - o.time()[2])
        ours = interior_state(s, d.__getitem__)
        ref = interior_state(s, o.field)
        assert np.array_equal(ours["Mask"], ref["Mask"]), st
        err = rel_errors(ours, ref)
        assert all(err[f] < 1e-9 for f in ("Eta", "D", "U", "W", "DU", "DW")), (st, err)
