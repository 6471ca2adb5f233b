"""Writes the golden fixtures in tests/golden/ by running the serial reference
(built from its own sources by oracle/build_ref.sh, `preview` replaced by the
raw-FP64 dump hook) on the synthetic scenarios below.

Fixtures hold the reference's state after the listed steps, interior cells
only (plus the mask), as FP64.  Scenario inputs are not stored: they are
regenerated from the named generator and parameters, and a SHA-256 of the
inputs is stored to catch drift.

Run in the build container (needs oracle/_ref/nhwave_ref_dump):
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import shutil
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

from nhwave_amd import scenario as S  # noqa: E402
import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent

# name -> (generator, kwargs, dump steps)
CASES = {
    # the CPU-runnable solitary-wave channel at its own TOL
    "solitary512_tol1e-10": ("solitary_channel", dict(steps=400, tol=1e-10), [1, 10, 100, 400]),
    # same, both sides solved tight (NSPCG clamps TOL at 3.55e-12).  Beyond
    # ~150 steps a transverse mode of this narrow channel, seeded by round-off,
    # grows ~10%/step in the reference itself, so pins stop at step 100.
    "solitary512_tight": ("solitary_channel", dict(steps=100, tol=1e-13), [1, 10, 100]),
    "solitary512_hydro": ("solitary_channel", dict(steps=100, non_hydro=False), [1, 100]),
    "beach_hydro": ("beach_runup", dict(steps=300, non_hydro=False), [1, 100, 300]),
    "beach_nh_tight": ("beach_runup", dict(steps=100, tol=1e-13), [1, 50, 100]),
    "random1_hydro": ("random_scenario", dict(seed=1, steps=5, non_hydro=False), [1, 5]),
    "random2_hydro": ("random_scenario", dict(seed=2, steps=5, non_hydro=False), [1, 5]),
    "random3_nh_tight": ("random_scenario", dict(seed=3, steps=3, tol=1e-13), [1, 3]),
    "bumps48_hydro": ("bumps_beach", dict(n=48, kglob=4, dx=200.0, steps=20, non_hydro=False),
                      [1, 20]),
    "bumps48_nh_tight": ("bumps_beach", dict(n=48, kglob=4, dx=200.0, steps=10, tol=1e-13),
                         [1, 10]),
    "solitary_rk3_hydro": ("solitary_channel",
                           dict(mglob=64, nglob=6, x0=1.5, steps=30, non_hydro=False, time_order=3),
                           [1, 30]),
    # boundary forcing: west wavemakers (BC_X0 = 3) and sponges, one-row
    # flumes like the reference's own examples, a wet-dry beach.  The
    # non-hydrostatic ones have three rows: with Nglob = 1 the serial
    # reference's NSPCG solve returns NaN (its +M and +MN stencil offsets
    # coincide), so one-row non-hydrostatic flumes are not pinned by it
    "flume_sol_beach_hydro": ("wave_flume", dict(mglob=200, wavemaker="LEF_SOL", amp=0.06, depth=0.21,
                                                 dx=0.04, bottom="beach", non_hydro=False, steps=150),
                              [1, 20, 150]),
    "flume_sol_beach_nh_tight": ("wave_flume", dict(mglob=200, nglob=3, wavemaker="LEF_SOL", amp=0.06, depth=0.21,
                                                    dx=0.04, bottom="beach", tol=1e-13, steps=60),
                                 [1, 60]),
    "flume_lin_bar_hydro": ("wave_flume", dict(mglob=240, wavemaker="LEF_LIN", amp=0.02, per=2.02,
                                               bottom="bar", sponge=(0.0, 1.0, 0.0, 0.0), non_hydro=False,
                                               steps=120), [1, 30, 120]),
    "flume_lin_bar_nh_tight": ("wave_flume", dict(mglob=240, nglob=3, wavemaker="LEF_LIN", amp=0.02, per=2.02,
                                                  bottom="bar", sponge=(0.0, 1.0, 0.0, 0.0), tol=1e-13,
                                                  steps=60), [1, 60]),
    "flume_stk_hydro": ("wave_flume", dict(mglob=100, nglob=3, wavemaker="LEF_STK", amp=0.03, per=1.5,
                                           bottom="flat", non_hydro=False, steps=80), [1, 80]),
    "flume_lin_oblique_hydro": ("wave_flume", dict(mglob=60, nglob=10, dx=0.05, wavemaker="LEF_LIN",
                                                   amp=0.02, per=1.2, theta=12.0, bottom="flat",
                                                   sponge=(0.0, 0.6, 0.2, 0.2), non_hydro=False,
                                                   steps=60, cur=0.05, tramp=0.5), [1, 60]),
    # the cnoidal wavemaker (LEF_CON): a moderate wave onto a beach, and the
    # steep wave of the reference's tingkirby example (AMP 0.125, PER 2.0,
    # DEP 0.4) over a bar, non-hydrostatic and with the k-epsilon model
    "flume_con_beach_hydro": ("wave_flume", dict(mglob=200, wavemaker="LEF_CON", amp=0.05, per=2.0, depth=0.4,
                                                 dx=0.04, bottom="beach", non_hydro=False, steps=150),
                              [1, 40, 150]),
    "flume_con_bar_nh_tight": ("wave_flume", dict(mglob=160, nglob=3, wavemaker="LEF_CON", amp=0.125, per=2.0,
                                                  depth=0.4, dx=0.025, bottom="bar", tol=1e-13, steps=60),
                               [1, 60]),
    "ke_con_beach_hydro": ("variant", dict(base="wave_flume", base_kw=dict(mglob=120, nglob=3, kglob=6,
                                                                          wavemaker="LEF_CON", amp=0.125, per=2.0,
                                                                          depth=0.4, dx=0.025, bottom="beach",
                                                                          non_hydro=False, steps=40),
                                           bc=[3, 1, 1, 1, 5], viscous=True, visc=1.0e-6, cvs=0.001, chs=0.001,
                                           cd0=0.006, turb=10), [1, 40]),
    # the internal irregular wavemaker (INT_IRR, a spectrum of irr2d.txt
    # with random phases, as in the reference's Monai Valley run): oblique
    # components whose directions the reference moves to the periodic-y
    # admissible ones, a wet-dry beach, non-hydrostatic, and the k-epsilon
    # model whose state is reset inside the source box
    "int_irr_basin_hydro": ("internal_wave_basin", dict(dirs=[0.0, 15.0, -20.0], non_hydro=False, steps=60),
                            [1, 20, 60]),
    "int_irr_basin_nh_tight": ("internal_wave_basin", dict(dirs=[0.0, 10.0], tol=1e-13, steps=20), [1, 20]),
    "ke_int_irr_hydro": ("variant", dict(base="internal_wave_basin",
                                         base_kw=dict(non_hydro=False, steps=20, box=[1.0, 1.2, 0.1, 0.5]),
                                         bc=[1, 1, 1, 1, 5], viscous=True, visc=1.0e-6, cvs=0.001, chs=0.001,
                                         cd0=0.004, turb=10), [1, 20]),
    # FLUX_LR: a ramped logarithmic inflow at the west and the open east
    # boundary with Orlanski radiation, over a submerged obstacle with the
    # log-law bottom, as in the reference's steady-flow current benchmark
    "flux_lr_steady_hydro": ("steady_flume", dict(non_hydro=False, steps=150), [1, 50, 150]),
    "flux_lr_steady_nh_tight": ("steady_flume", dict(tol=1e-13, steps=40), [1, 40]),
    "ke_flux_lr_hydro": ("variant", dict(base="steady_flume", base_kw=dict(non_hydro=False, steps=60),
                                         viscous=True, visc=1.0e-6, cvs=0.001, chs=0.001, turb=10), [1, 60]),
    # wall friction side boundaries (BC 5 at the lateral walls, viscous runs):
    # the ghost factor 1 - Phi of the tangential velocities with Cd0 or the
    # log law, and kepsilon_3D's wall function at the first interior cells;
    # with FLUX_LR as in the reference's steady-flow benchmark.  The
    # non-hydrostatic flume stops at step 10: V across this symmetric channel
    # is of round-off size, and the open end's Orlanski speed for V, a ratio
    # of two such differences clipped to [0, 1], turns the solvers' 1e-12
    # differences into 1.7e-9 of the flow speed by step 20 (measured against
    # the oracle; the hydrostatic twin stays bit-exact for 60 steps)
    "wall5_visc_random1_hydro": ("variant", dict(base="random_scenario", base_kw=dict(seed=1, steps=5,
                                                 non_hydro=False), bc=[5, 5, 5, 5, 5], viscous=True,
                                                 visc=1.0e-3, cvs=0.02, chs=0.05, cd0=0.004), [1, 5]),
    "ke_wall5_random6_log_hydro": ("variant", dict(base="random_scenario", base_kw=dict(seed=6, steps=5,
                                                   non_hydro=False), bc=[5, 5, 5, 5, 5], viscous=True,
                                                   visc=1.0e-3, cvs=0.02, chs=0.05, ibot=2, zob=1.0e-3, turb=10),
                                   [1, 2, 5]),
    "ke_wall5_random3_nh_tight": ("variant", dict(base="random_scenario", base_kw=dict(seed=3, steps=3, tol=1e-13),
                                                  bc=[5, 2, 5, 1, 5], viscous=True, visc=1.0e-3, cvs=0.02,
                                                  chs=0.05, cd0=0.004, turb=10), [1, 3]),
    "ke_flux_lr_walls_hydro": ("variant", dict(base="steady_flume", base_kw=dict(nglob=6, bc_y=5, non_hydro=False,
                                                                                steps=60),
                                               viscous=True, visc=1.0e-6, cvs=0.001, chs=0.001, turb=10), [1, 60]),
    "ke_flux_lr_walls_nh_tight": ("variant", dict(base="steady_flume", base_kw=dict(nglob=6, bc_y=5, tol=1e-13,
                                                                                   steps=10),
                                                  viscous=True, visc=1.0e-6, cvs=0.001, chs=0.001, turb=10),
                                  [1, 10]),
    # viscous branch with constant viscosities (IVTURB = IHTURB = 1): random
    # islands with initial velocities, a wet-dry bumpy basin, a case where the
    # explicit viscous dt limit binds, and non-hydrostatic runs
    "visc_random1_hydro": ("viscous_case", dict(base="random_scenario", base_kw=dict(seed=1, steps=5,
                                                non_hydro=False), visc=1.0e-3, cvs=0.02, chs=0.05), [1, 5]),
    "visc_random4_dtlimit_hydro": ("viscous_case", dict(base="random_scenario", base_kw=dict(
                                       seed=4, steps=8, non_hydro=False), visc=1.0e-3, cvs=0.05, chs=2.5),
                                   [1, 2, 8]),
    "visc_bumps48_hydro": ("viscous_case", dict(base="bumps_beach", base_kw=dict(
                               n=48, kglob=4, dx=200.0, steps=20, non_hydro=False), visc=1.0e-6, cvs=0.05,
                               chs=20.0), [1, 20]),
    "visc_random3_nh_tight": ("viscous_case", dict(base="random_scenario", base_kw=dict(seed=3, steps=3,
                                                   tol=1e-13), visc=1.0e-3, cvs=0.02, chs=0.05), [1, 3]),
    "visc_solitary_rk3_nh_tight": ("viscous_case", dict(base="solitary_channel", base_kw=dict(
                                       mglob=96, nglob=4, kglob=5, x0=2.0, steps=40, tol=1e-13, time_order=3),
                                       visc=1.0e-4, cvs=0.002, chs=0.004, schmidt=0.7), [1, 40]),
    # bottom stress: friction law (BC_Z0 = 5, constant Cd0 or the log law) and
    # no-slip (BC_Z0 = 2), inviscid and viscous
    "bot5_random5_hydro": ("variant", dict(base="random_scenario", base_kw=dict(seed=5, steps=5, non_hydro=False),
                                           bc=[1, 1, 1, 1, 5], cd0=0.005), [1, 5]),
    "bot5_visc_random6_hydro": ("variant", dict(base="random_scenario", base_kw=dict(seed=6, steps=5,
                                                non_hydro=False), bc=[1, 1, 1, 1, 5], viscous=True,
                                                visc=1.0e-3, cvs=0.02, chs=0.05, ibot=2, zob=1.0e-3), [1, 5]),
    "bot5_visc_bumps48_hydro": ("variant", dict(base="bumps_beach", base_kw=dict(n=48, kglob=4, dx=200.0, steps=20,
                                                non_hydro=False), bc=[1, 1, 1, 1, 5], viscous=True, visc=1.0e-6,
                                                cvs=0.05, chs=20.0, cd0=0.003), [1, 20]),
    "bot2_visc_random7_hydro": ("variant", dict(base="random_scenario", base_kw=dict(seed=7, steps=5,
                                                non_hydro=False), bc=[1, 1, 1, 1, 2], viscous=True,
                                                visc=1.0e-3, cvs=0.02, chs=0.05), [1, 5]),
    "bot5_visc_random8_nh_tight": ("variant", dict(base="random_scenario", base_kw=dict(seed=8, steps=3,
                                                   tol=1e-13), bc=[1, 1, 1, 1, 5], viscous=True, visc=1.0e-3,
                                                   cvs=0.02, chs=0.05, cd0=0.004), [1, 3]),
    # no-slip bottom without viscosity (the bottom-face states zeroed, the
    # bottom-face slopes still entering the ghosts), and a well-conditioned
    # non-hydrostatic friction-law case held to the default tolerances
    "bot2_random9_hydro": ("variant", dict(base="random_scenario", base_kw=dict(seed=9, steps=5, non_hydro=False),
                                           bc=[1, 1, 1, 1, 2]), [1, 5]),
    "bot5_bumps48_nh_tight": ("variant", dict(base="bumps_beach", base_kw=dict(n=48, kglob=4, dx=200.0, steps=10,
                                                                               tol=1e-13),
                                              bc=[1, 1, 1, 1, 5], cd0=0.003), [1, 10]),
    # the 512x8x3 channel solved tight through step 400: against the TOL 1e-10
    # fixture it shows how far the reference drifts from itself once the
    # channel's transverse mode has grown (the band of the step-400 checks)
    "solitary512_tol1e-13_step400": ("solitary_channel", dict(steps=400, tol=1e-13), [400]),
    # the 3-D k-epsilon closure (IVTURB = IHTURB = 10) with the friction-law
    # bottom as every shipped viscous example runs it: random islands with
    # strong near-bed flow (constant Cd0 and the log law), a run-up beach
    # with shallow and drying columns, the bump basin, non-hydrostatic
    "ke_random5_hydro": ("variant", dict(base="random_scenario", base_kw=dict(seed=5, steps=5, non_hydro=False),
                                         bc=[1, 1, 1, 1, 5], viscous=True, visc=1.0e-3, cvs=0.02, chs=0.05,
                                         cd0=0.004, turb=10), [1, 2, 5]),
    "ke_random6_log_hydro": ("variant", dict(base="random_scenario", base_kw=dict(seed=6, steps=5, non_hydro=False),
                                             bc=[1, 1, 1, 1, 5], viscous=True, visc=1.0e-3, cvs=0.02, chs=0.05,
                                             ibot=2, zob=1.0e-3, turb=10), [1, 5]),
    "ke_beach_hydro": ("variant", dict(base="beach_runup", base_kw=dict(mglob=64, nglob=4, kglob=4, steps=60,
                                                                        non_hydro=False),
                                       bc=[1, 1, 1, 1, 5], viscous=True, visc=1.0e-6, cvs=0.001, chs=0.001,
                                       cd0=0.006, turb=10), [1, 30, 60]),
    "ke_bumps48_hydro": ("variant", dict(base="bumps_beach", base_kw=dict(n=48, kglob=5, dx=50.0, steps=20,
                                                                          non_hydro=False),
                                         bc=[1, 1, 1, 1, 5], viscous=True, visc=1.0e-6, cvs=0.001, chs=0.001,
                                         cd0=0.003, turb=10), [1, 20]),
    "ke_random7_nh_tight": ("variant", dict(base="random_scenario", base_kw=dict(seed=7, steps=3, tol=1e-13),
                                            bc=[1, 1, 1, 1, 5], viscous=True, visc=1.0e-3, cvs=0.02, chs=0.05,
                                            cd0=0.004, turb=10), [1, 3]),
    # twins of two k-epsilon cases run with BAROTROPIC = F, as most of the
    # shipped viscous examples set it: without the salinity, temperature or
    # sediment builds the density stays Rho0, and baropg_z's gradient is
    # -0.0; the reference's own output must equal the BAROTROPIC = T runs
    "ke_random5_baroclinic_hydro": ("variant", dict(base="random_scenario",
                                                    base_kw=dict(seed=5, steps=5, non_hydro=False),
                                                    bc=[1, 1, 1, 1, 5], viscous=True, visc=1.0e-3, cvs=0.02,
                                                    chs=0.05, cd0=0.004, turb=10), [1, 2, 5],
                                    {"BAROTROPIC": "F"}),
    "ke_random7_baroclinic_nh_tight": ("variant", dict(base="random_scenario",
                                                       base_kw=dict(seed=7, steps=3, tol=1e-13),
                                                       bc=[1, 1, 1, 1, 5], viscous=True, visc=1.0e-3, cvs=0.02,
                                                       chs=0.05, cd0=0.004, turb=10), [1, 3],
                                       {"BAROTROPIC": "F"}),
}

FIELDS_2D = ("Eta", "D")
FIELDS_3D = ("U", "V", "W", "DU", "DV", "DW")


def make_scenario(gen: str, kw: dict):
    return getattr(S, gen)(**kw)


def input_digest(s) -> str:
    h = hashlib.sha256()
    for a in (s.hcg, s.etag, s.ug, s.vg, s.wg):
        if a is not None:
            h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def interior(s, d: dict) -> dict:
    g = S.NGHOST
    out = {}
    for n in FIELDS_2D:
        out[n] = d[n][g:-g, g:-g]
    out["Mask"] = d["Mask"][g:-g, g:-g].astype(np.int8)
    for n in FIELDS_3D:
        out[n] = d[n][g:-g, g:-g, g:-g]
    out["P"] = d["P"][g:g + s.kglob, g:-g, g:-g]
    out["Omega"] = d["Omega"][g:g + s.kglob + 1, g:-g, g:-g]
    if "Tke" in d and getattr(s, "turb", 1) == 10:
        # the k-epsilon state of the interior columns, ghost levels included
        # (the eddy viscosity of the ghost levels enters the vertical solves)
        for n in ("Tke", "Eps", "CmuVt"):
            out[n] = d[n][:, g:-g, g:-g]
    return out


def main(names=None, check=False) -> int:
    """Writes the fixtures; with check=True writes nothing and compares a fresh
    run of the reference build with each committed fixture, bit for bit
    (arrays, TIME and dt).  Returns the number of fixtures that differ."""
    if not O.REF_DUMP_BIN.exists():
        raise SystemExit("oracle/_ref/nhwave_ref_dump missing: run oracle/build_ref.sh first")
    meta_all = {}
    meta_path = OUT / "golden_index.json"
    if meta_path.exists():
        meta_all = json.loads(meta_path.read_text())
    differ = 0
    for name, case in CASES.items():
        if names and name not in names:
            continue
        gen, kw, steps = case[:3]
        overrides = case[3] if len(case) > 3 else {}
        s = make_scenario(gen, kw)
        tmp = Path(tempfile.mkdtemp(prefix="golden_"))
        wall, dumps = O.run_reference(s, tmp, steps, timeout=3600, input_overrides=overrides)
        arrays = {}
        meta = {"generator": gen, "kwargs": kw, "steps": steps, "input_sha256": input_digest(s),
                "reference_wall_s": wall, "time": {}, "dt": {}}
        if overrides:
            meta["input_overrides"] = overrides
        for st in steps:
            d = dumps[st]
            for k, v in interior(s, d).items():
                arrays[f"s{st}_{k}"] = v
            meta["time"][str(st)] = d["TIME"]
            meta["dt"][str(st)] = d["dt"]
        shutil.rmtree(tmp, ignore_errors=True)
        if check:
            old = np.load(OUT / f"{name}.npz")
            om = meta_all[name]
            bad = sorted(set(old.files) ^ set(arrays))
            bad += [k for k in arrays if k in old.files and not (
                old[k].dtype == arrays[k].dtype and old[k].shape == arrays[k].shape
                and old[k].tobytes() == np.ascontiguousarray(arrays[k]).tobytes())]
            bad += [f"{q}@{st}" for q in ("time", "dt") for st in meta[q] if om[q].get(st) != meta[q][st]]
            if om["input_sha256"] != meta["input_sha256"]:
                bad.append("input_sha256")
            differ += bool(bad)
            print(f"{name}: {'identical' if not bad else 'DIFFERS in ' + ', '.join(bad[:8])}", flush=True)
            continue
        np.savez_compressed(OUT / f"{name}.npz", **arrays)
        meta_all[name] = meta
        print(f"{name}: {wall:.1f} s, steps {steps}", flush=True)
    if not check:
        meta_path.write_text(json.dumps(meta_all, indent=1, sort_keys=True))
    return differ


if __name__ == "__main__":
    args = sys.argv[1:]
    chk = "--check" in args
    sys.exit(1 if main([a for a in args if a != "--check"] or None, check=chk) else 0)
