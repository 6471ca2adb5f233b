"""Writes fixtures from four of the reference's own example runs
(examples/solitary_wave_breaking: a solitary wave from the west wavemaker
breaking on a slope; examples/submerged_bar: linear waves over a submerged
bar with an east sponge; the analytical simple-beach benchmark's
time_solution run: a solitary wave on a 1:19.85 beach with the k-epsilon
closure and the friction-law bottom; examples/tingkirby: cnoidal waves
(LEF_CON) breaking on a beach, k-epsilon, 20 sigma levels, one row): their
input.txt key/value pairs and
depth arrays (tests/golden/example_<name>.json / _depth.npy, data only), and
golden states of the reference build (oracle/_ref/nhwave_ref_dump) on the
same inputs for a bounded number of steps: run hydrostatically
(example_<name>_hydro.npz), and, where the grid has more than one row,
non-hydrostatically with TOL 1e-13 (example_<name>_nh_tight.npz).

The first two examples and tingkirby are non-hydrostatic one-row flumes; the serial
reference's NSPCG solve returns NaN on one-row grids, so their golden states
are taken with NON_HYDRO = F; their non-hydrostatic runs are checked against
the oracle.

Run in the build container (needs /root/reference and oracle/_ref):
    python tests/golden/make_example_fixtures.py
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

sys.path.insert(0, str(ROOT / "tests"))

from nhwave_amd import ingest as I  # noqa: E402
from nhwave_amd import scenario as S  # noqa: E402
import oracle as O  # noqa: E402
from nhw_testutil import example_run_dir, hydro_overrides, nh_tight_overrides  # noqa: E402

OUT = Path(__file__).resolve().parent
EXAMPLES = Path("/root/reference/examples")
# name -> (directory under examples/, hydrostatic golden steps ("end" = the
# example's own TOTAL_TIME), non-hydrostatic TOL 1e-13 golden steps)
CASES = {
    "solitary_wave_breaking": ("solitary_wave_breaking", [1, 50, 300, "end"], []),
    "submerged_bar": ("submerged_bar", [1, 50, 300], []),
    "simple_beach_time_solution": (
        "inundation_benchmarks/Solitary_wave_on_a_simple_beach_analytical/time_solution/input", [1, 50, 200],
        [1, 10]),
    "tingkirby": ("tingkirby", [1, 50, 200], []),
    # the internal wavemaker's spectrum of 399 frequencies is set up again at
    # every call of the reference's source_terms (a dispersion-relation
    # solve per component and source cell): about 12 s a step here.  Its
    # non-hydrostatic states (50 MB for two steps of 492 x 243 x 4 cells)
    # are not kept; int_irr_basin_nh_tight covers that path
    "monai_valley": ("inundation_benchmarks/Runup_onto_a_complex_3-D_beach_Monai_Valley/input", [1, 10, 25],
                     []),
    # the steady current over a submerged obstacle: FLUX_LR with wall
    # friction side walls (BC 5), the k-epsilon closure and the log-law
    # bottom on 656 x 100 x 15 cells; the flow starts from rest, so the
    # early hydrostatic states are mostly zeros and stay small as fixtures.
    # Its non-hydrostatic states (118 MB for two steps: P fills every cell)
    # are not kept; ke_flux_lr_walls_nh_tight covers that path
    "steady_flow": ("current_benchmarks/Steady_flow_over_submerged_obstacle/input", [1, 5], []),
}
TURB = ("Tke", "Eps", "CmuVt")


def states(s, run_dir: Path, steps) -> tuple:
    """Interior states of the dumps after `steps` (plus the k-epsilon state of
    the interior columns, ghost levels included, when the run has it)."""
    g = S.NGHOST
    arrays, time_, dt_ = {}, {}, {}
    for st in steps:
        d = S.read_dump(run_dir / f"dump_{st:05d}.bin")
        for k in ("Eta", "D"):
            arrays[f"s{st}_{k}"] = d[k][g:-g, g:-g]
        arrays[f"s{st}_Mask"] = d["Mask"][g:-g, g:-g].astype(np.int8)
        for k in ("U", "V", "W", "DU", "DV", "DW"):
            arrays[f"s{st}_{k}"] = d[k][g:-g, g:-g, g:-g]
        arrays[f"s{st}_P"] = d["P"][g:g + s.kglob, g:-g, g:-g]
        arrays[f"s{st}_Omega"] = d["Omega"][g:g + s.kglob + 1, g:-g, g:-g]
        if s.viscous and s.turb == 10:
            for k in TURB:
                arrays[f"s{st}_{k}"] = d[k][:, g:-g, g:-g]
        time_[str(st)] = d["TIME"]
        dt_[str(st)] = d["dt"]
    return arrays, time_, dt_


def _big_stack():
    # the reference keeps grid-sized automatic arrays on the stack
    import resource
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run_ref(run_dir: Path, steps) -> float:
    env = dict(os.environ, NHW_DUMP_STEPS=",".join(map(str, steps)))
    t0 = time.perf_counter()
    res = subprocess.run([str(O.REF_DUMP_BIN)], cwd=run_dir, env=env, capture_output=True, text=True,
                         timeout=3600, preexec_fn=_big_stack)
    if res.returncode != 0 or "Normal Termination" not in res.stdout:
        raise RuntimeError(res.stdout[-2000:] + res.stderr[-2000:])
    return time.perf_counter() - t0


def last_step(run_dir: Path) -> int:
    """RUN_STEP of the last line estimate_dt logged (src/nhwave.F:3856)."""
    last = 0
    for ln in (run_dir / "log.txt").read_text(errors="replace").splitlines():
        t = ln.split()
        if len(t) == 3 and t[0].isdigit():
            last = int(t[0])
    return last


def main(names=None) -> None:
    index_path = OUT / "examples_index.json"
    index = json.loads(index_path.read_text()) if index_path.exists() else {}
    for name, (rel, steps, nh_steps) in CASES.items():
        if names and name not in names:
            continue
        src = EXAMPLES / rel
        keys = I.read_input(src / "input.txt")
        inp = I.Input(src / "input.txt")
        m, n = inp.int("Mglob"), inp.int("Nglob")
        depth = I.read_rows(src / "depth.txt", n, m)
        (OUT / f"example_{name}.json").write_text(json.dumps(
            {"source": f"examples/{rel}/input.txt, depth.txt", "keys": keys}, indent=1))
        np.save(OUT / f"example_{name}_depth.npy", depth)
        if (src / "irr2d.txt").exists():  # the internal wavemaker's spectrum, as data
            f, di, sp, ph = I.read_irr2d(src / "irr2d.txt")
            np.savez(OUT / f"example_{name}_irr2d.npz", freq=f, dire=di, spc=sp, phs=ph)
        tmp = Path(tempfile.mkdtemp(prefix=f"example_{name}_"))
        try:
            to_end = "end" in steps
            steps = [st for st in steps if st != "end"]
            example_run_dir(name, tmp, hydro_overrides(None if to_end else max(steps)))
            s = I.scenario_from_run_dir(tmp)
            if to_end:  # find the step the example ends on, then dump there too
                run_ref(tmp, [0])
                steps = steps + [last_step(tmp)]
            wall = run_ref(tmp, steps)
            arrays, time_, dt_ = states(s, tmp, steps)
            meta = {"steps": steps, "time": time_, "dt": dt_, "reference_wall_s": wall, "to_total_time": to_end}
            np.savez_compressed(OUT / f"example_{name}_hydro.npz", **arrays)
            print(f"{name}: {wall:.1f} s, steps {steps}", flush=True)
            if nh_steps:
                shutil.rmtree(tmp, ignore_errors=True)
                example_run_dir(name, tmp, nh_tight_overrides(max(nh_steps)))
                s = I.scenario_from_run_dir(tmp)
                wall = run_ref(tmp, nh_steps)
                arrays, time_, dt_ = states(s, tmp, nh_steps)
                meta["nh_tight"] = {"steps": nh_steps, "time": time_, "dt": dt_, "reference_wall_s": wall}
                np.savez_compressed(OUT / f"example_{name}_nh_tight.npz", **arrays)
                print(f"{name} (non-hydrostatic, TOL 1e-13): {wall:.1f} s, steps {nh_steps}", flush=True)
            index[name] = meta
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
    index_path.write_text(json.dumps(index, indent=1, sort_keys=True))


if __name__ == "__main__":
    main(sys.argv[1:] or None)
