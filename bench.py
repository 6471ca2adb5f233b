"""Benchmark of the NHWAVE time-stepping hot path on one MI355X.

Workload (default): the plain-FP64 single run on 1024 x 1024 x 10 cells,
dx = 10 m, seeded Gaussian-bump bathymetry of 2-200 m with a 1:20 drying
beach on the east side, a 2 m Gaussian hump of radius 500 m, SSP-RK2,
CFL 0.5, non-hydrostatic with Poisson relative residual 1e-10.  A step is
one full time step (estimate_dt, update_mask, update_vars and two RK stages
of source_terms, fluxes and eval_duvw with its Poisson solve).

Metric: cell-stage updates/s = Mglob*Nglob*Kglob*It_Order*steps / wall.

Prints ONE JSON line with the contract's fields plus `roofline` (dominant
kernel, from HIP events recorded on the library's stream during the timed
steps) and `cpu_baseline` (the reference's serial build, or the oracle port
when that build is absent, on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from nhwave_amd import capi  # noqa: E402
from nhwave_amd import scenario as S  # noqa: E402
from nhwave_amd.model import Model  # noqa: E402

HBM_PEAK_GBS = 8000.0
PROFILES = ROOT / "profiles"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(size: int, kglob: int, tol: float) -> S.Scenario:
    return S.bumps_beach(n=size, kglob=kglob, dx=10.0, steps=2000, tol=tol)


def kernel_table(m: Model) -> list[dict]:
    lib = m.lib
    out = []
    kid = 0
    while True:
        name = C.c_char_p()
        launches = C.c_int64()
        total = C.c_double()
        bpl = C.c_double()
        rc = lib.nhw_kernel_stats(m.ctx, kid, C.byref(name), C.byref(launches), C.byref(total), C.byref(bpl))
        if rc != 0:
            break
        if launches.value:
            out.append(dict(kernel=name.value.decode(), launches=launches.value, total_ms=total.value,
                            avg_ms=total.value / launches.value, bytes_per_launch=bpl.value))
        kid += 1
    return out


def traffic_for(kernel: str, workload_name: str):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if one
    exists for this kernel and workload."""
    p = PROFILES / f"traffic_{kernel}_{workload_name}.json"
    if not p.exists():
        return None
    try:
        return float(json.loads(p.read_text())["bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline(s_full: S.Scenario, window: int, steps: int) -> dict:
    """Times the reference's serial build (oracle/_ref) on a centred
    window x window x K cut of the same bathymetry and initial state; the
    oracle's C port stands in when that build is absent."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O  # noqa: E402  (checker only)
    g0 = (s_full.nglob - window) // 2
    g1 = (s_full.mglob - window) // 2
    sl = np.s_[g0:g0 + window, g1:g1 + window]
    z = np.zeros((s_full.kglob, window, window))
    s = s_full.with_(name=f"{s_full.name}_window{window}", mglob=window, nglob=window,
                     hcg=s_full.hcg[sl].copy(), etag=s_full.etag[sl].copy(), ug=z, vg=z.copy(),
                     wg=z.copy(), steps=steps)
    cells_stages = window * window * s.kglob * s.time_order * steps
    sample = (f"{window}x{window}x{s.kglob} centre window of the same bathymetry and hump, "
              f"{steps} steps incl. start-up, TOL {s.tol:g}")
    if O.REF_BIN.exists():
        with tempfile.TemporaryDirectory(prefix="nhw_cpu_") as d:
            wall, _ = O.run_reference(s, Path(d), [], timeout=600, binary=O.REF_BIN)
        kind = "reference"
        sample = "serial reference build (NSPCG IC3+GMRES), " + sample
    else:
        t0 = time.perf_counter()
        o = O.Oracle(s)
        o.step(steps)
        wall = time.perf_counter() - t0
        kind = "port"
        sample = "oracle C restatement (banded direct Poisson solve), " + sample
    return dict(value=cells_stages / wall, unit="cell-stage updates/s", cores=1, kind=kind,
                sample=sample, wall_s=wall)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--kglob", type=int, default=10)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--cpu-window", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=3,
                    help="extra steps run after the timed ones with per-launch HIP events, for the "
                         "kernel table and roofline (0: none, e.g. under rocprofv3)")
    ap.add_argument("--no-kernel-timing", action="store_true", help="same as --profile-steps 0")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        lib = capi.load()
        nbytes = args.copy_calibration << 20
        ms = C.c_double()
        capi.check(lib.nhw_bench_copy(nbytes, 20, C.byref(ms)), "nhw_bench_copy")
        print(json.dumps({"kernel": "k_copy", "bytes_read": nbytes, "bytes_written": nbytes,
                          "avg_ms": ms.value, "gbs": 2 * nbytes / (ms.value * 1e-3) / 1e9}), flush=True)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is a "
                                   "design note, not a measured path", "n_gpus": args.gpus}))
        return 2
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "launch bench.py directly, one process, one GPU"}))
        return 2

    s = workload(args.size, args.kglob, args.tol)
    wl = f"bumps{args.size}x{args.size}x{args.kglob}_fp64"
    log(f"workload {wl}: building scenario and device state")
    m = Model(s)
    log(f"warmup {args.warmup} steps")
    m.step(args.warmup)
    m.sync()
    st0 = m.stats()
    log(f"timing {args.steps} steps")
    m.sync()
    t0 = time.perf_counter()
    m.step(args.steps)
    m.sync()
    wall = time.perf_counter() - t0
    st1 = m.stats()
    # kernel table: further steps with an event pair around every launch
    # (kept out of the timed region above)
    nprof = 0 if args.no_kernel_timing else args.profile_steps
    kt = []
    prof_wall = None
    if nprof > 0:
        log(f"{nprof} instrumented steps for the kernel table")
        m.lib.nhw_kernel_timing(m.ctx, 1)
        tp = time.perf_counter()
        m.step(nprof)
        m.sync()
        prof_wall = time.perf_counter() - tp
        kt = kernel_table(m)
        m.lib.nhw_kernel_timing(m.ctx, 0)
    t_sim, dt, run_step = m.time()
    m.close()

    cells = s.mglob * s.nglob * s.kglob
    value = cells * s.time_order * args.steps / wall
    solves = st1.solves - st0.solves
    iters = st1.total_iters - st0.total_iters
    roof = None
    if kt:
        dom = max(kt, key=lambda r: r["total_ms"])
        achieved = dom["bytes_per_launch"] / (dom["avg_ms"] * 1e-3) / 1e9
        roof = dict(bound="hbm", kernel=dom["kernel"], achieved=achieved, peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                    traffic=traffic_for(dom["kernel"], wl),
                    avg_launch_ms=dom["avg_ms"], launches=dom["launches"],
                    algorithmic_bytes_per_launch=dom["bytes_per_launch"],
                    share_of_kernel_time=dom["total_ms"] / sum(r["total_ms"] for r in kt),
                    instrumented_steps=nprof, instrumented_ms_per_step=prof_wall / nprof * 1e3)
        model_bytes = sum(r["bytes_per_launch"] * r["launches"] for r in kt)
        roof["all_kernels_model_gbs"] = model_bytes / prof_wall / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        try:
            cpu = cpu_baseline(s, args.cpu_window, args.cpu_steps)
        except Exception as e:  # report, do not hide
            cpu = dict(error=str(e))
    out = {
        "metric": "cell-stage updates/s (plain FP64)",
        "value": value,
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl, "mglob": s.mglob, "nglob": s.nglob, "kglob": s.kglob,
                   "dx_m": s.dx, "time_order": s.time_order, "non_hydro": True, "tol": s.tol,
                   "cfl": s.cfl, "seed": s.meta["seed"], "parallelism": "single GPU"},
        "krylov_iters_per_solve": iters / max(solves, 1),
        "sim_time_s": t_sim,
        "run_step": run_step,
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": kt,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
