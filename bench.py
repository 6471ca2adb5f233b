"""Benchmark of the NHWAVE time-stepping hot path on one MI355X.

Headline workload (default): the plain-FP64 single run on 1024 x 1024 x 10
cells, dx = 10 m, seeded Gaussian-bump bathymetry of 2-200 m with a 1:20
drying beach on the east side, a 2 m Gaussian hump of radius 500 m, SSP-RK2,
CFL 0.5, non-hydrostatic with Poisson relative residual 1e-10.  A step is one
full time step (estimate_dt, update_mask, update_wind, update_vars and two RK
stages of source_terms, fluxes and eval_duvw with its Poisson solve).

Metric: cell-stage updates/s = Mglob*Nglob*Kglob*It_Order*steps / wall.

The same JSON line carries the other forms measured in the same run:
  variants.mixed_2048   2048 x 2048 x 8, FP32 storage / inner solves with FP64
                        refinement to 1e-8, and its errors against the plain
                        FP64 form run on the same input for the same steps;
  variants.batched_64   64 scenarios (grids log-uniform 128^2..768^2, Kglob
                        3..10, hump amplitude 0.5-5 m, radius 200-2000 m) in
                        one batch, mixed precision, autotuned sweeps;
  variants.viscous_1024 the headline input with the viscous branch on
                        (constant viscosities, plain FP64).
`--workload serve` runs the serving loop instead (Poisson arrivals, grids
64^2x3 .. 1024^2x10, 1000 steps each) and prints p50/p99 time-to-result.
`--workload basin` runs a forced wave basin (2048 x 1024 x 8, dx = 5 cm, a
linear wave train entering at 10 degrees from the west wavemaker over a
submerged bar, sponges on the east, south and north sides) with the same
kernel table, for the boundary-forcing kernels.
`--workload viscous` runs the 1024 x 1024 x 10 bumps case with the viscous
branch on (constant viscosities), for the diffusion and vertical-solve kernels.
`--workload steady` runs a steady current over a submerged obstacle at the
size of the reference's steady-flow example (656 x 100 x 15, dx = 1.5 cm:
FLUX_LR between wall friction side walls, the k-epsilon closure, the log-law
bottom, non-hydrostatic), for the FLUX_LR and wall-function kernels.

Prints ONE JSON line with the contract's fields plus `roofline` (dominant
kernel, from HIP events recorded on the library's stream during extra
instrumented steps) and `cpu_baseline` (the reference's serial build, or the
oracle port when that build is absent, on a bounded sample of the workload).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
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
from nhwave_amd.batch import Batch  # noqa: E402
from nhwave_amd.model import Model  # noqa: E402

HBM_PEAK_GBS = 8000.0
PROFILES = ROOT / "profiles"
G = S.NGHOST


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(size: int, kglob: int, tol: float) -> S.Scenario:
    return S.bumps_beach(n=size, kglob=kglob, dx=10.0, steps=2000, tol=tol)


def kernel_table(lib, ctx) -> list[dict]:
    out = []
    kid = 0
    while True:
        name = C.c_char_p()
        launches = C.c_int64()
        total = C.c_double()
        bpl = C.c_double()
        rc = lib.nhw_kernel_stats(ctx, kid, C.byref(name), C.byref(launches), C.byref(total), C.byref(bpl))
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


def cpu_baseline(s_full: S.Scenario, window: int, steps: int, timeout: float = 600.0) -> dict:
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
            wall, _ = O.run_reference(s, Path(d), [], timeout=timeout, binary=O.REF_BIN)
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


def roofline(kt: list[dict], wl: str, prof_wall: float, nprof: int):
    if not kt:
        return None
    dom = max(kt, key=lambda r: r["total_ms"])
    achieved = dom["bytes_per_launch"] / (dom["avg_ms"] * 1e-3) / 1e9
    roof = dict(bound="hbm", kernel=dom["kernel"], achieved=achieved, peak=HBM_PEAK_GBS,
                unit="GB/s", frac=achieved / HBM_PEAK_GBS, traffic=traffic_for(dom["kernel"], wl),
                traffic_source=(f"profiles/traffic_{dom['kernel']}_{wl}.json: rocprofv3 --pmc FETCH_SIZE and "
                                "WRITE_SIZE runs of this workload, committed, scaled by a copy-kernel "
                                "calibration (not measured in this run)"),
                avg_launch_ms=dom["avg_ms"], launches=dom["launches"],
                algorithmic_bytes_per_launch=dom["bytes_per_launch"],
                share_of_kernel_time=dom["total_ms"] / sum(r["total_ms"] for r in kt),
                instrumented_steps=nprof, instrumented_ms_per_step=prof_wall / nprof * 1e3)
    model_bytes = sum(r["bytes_per_launch"] * r["launches"] for r in kt)
    roof["all_kernels_model_gbs"] = model_bytes / prof_wall / 1e9
    return roof


def run_single(s: S.Scenario, precision: int, steps: int, warmup: int, nprof: int, keep=(),
               solver: int = capi.NHW_SOLVER_MG, late=None):
    m = Model(s, precision=precision, solver=solver)
    m.step(warmup)
    m.sync()
    st0 = m.stats()
    t0 = time.perf_counter()
    m.step(steps)
    m.sync()
    wall = time.perf_counter() - t0
    st1 = m.stats()
    kt, prof_wall = [], None
    if nprof > 0:
        m.lib.nhw_kernel_timing(m.ctx, 1)
        tp = time.perf_counter()
        m.step(nprof)
        m.sync()
        prof_wall = time.perf_counter() - tp
        kt = kernel_table(m.lib, m.ctx)
        m.lib.nhw_kernel_timing(m.ctx, 0)
    fields = m.download(*keep) if keep else {}
    t_sim = m.time()
    cells = s.mglob * s.nglob * s.kglob
    late_out = None
    if late:
        # the same run carried on (untimed) to step `late[0]`, then `late[1]`
        # steps timed there: the solver's work later in the run
        to, window = late
        log(f"late window: advancing to step {to}")
        m.step(max(to - t_sim[2], 0))
        m.sync()
        la = m.stats()
        tl = time.perf_counter()
        m.step(window)
        m.sync()
        wl = time.perf_counter() - tl
        lb = m.stats()
        late_out = dict(from_step=m.time()[2] - window, steps=window, ms_per_step=wl / window * 1e3,
                        value=cells * s.time_order * window / wl,
                        iters_per_solve=(lb.total_iters - la.total_iters) / max(lb.solves - la.solves, 1),
                        max_iters=lb.max_iters, not_converged=lb.not_converged, sim_time_s=m.time()[0])
    m.close()
    solves = st1.solves - st0.solves
    return dict(value=cells * s.time_order * steps / wall, wall=wall, ms_per_step=wall / steps * 1e3,
                iters_per_solve=(st1.total_iters - st0.total_iters) / max(solves, 1),
                last_relres=st1.last_relres, not_converged=st1.not_converged, kt=kt, prof_wall=prof_wall,
                fields=fields, time=t_sim, late=late_out)


def rel(a, b):
    return float(np.abs(a - b).max()) / max(float(np.abs(b).max()), 1e-300)


def variant_mixed(size: int, kglob: int, steps: int, warmup: int, compare: bool = True) -> dict:
    """2048x2048x8 mixed form; errors against the plain FP64 form on the
    same input after the same steps (the FP64 reference of this size).
    compare=False skips the FP64 run (a trace of the mixed form alone)."""
    s = S.bumps_beach(n=size, kglob=kglob, dx=10.0, steps=2000, tol=1e-8)
    keep = ("Eta", "U", "V", "W", "D", "Mask")
    log(f"mixed {size}x{size}x{kglob}")
    a = run_single(s, capi.NHW_MIXED, steps, warmup, 0, keep)
    if not compare:
        return dict(workload=f"bumps{size}x{size}x{kglob}_mixed", value=a["value"], ms_per_step=a["ms_per_step"],
                    steps=steps, warmup=warmup, tol=1e-8, krylov_inner_iters_per_solve=a["iters_per_solve"],
                    not_converged=a["not_converged"])
    log(f"fp64 reference of the mixed run {size}x{size}x{kglob}")
    b = run_single(s, capi.NHW_FP64, steps, warmup, 0, keep)
    fa, fb = a["fields"], b["fields"]
    err = {"eta": rel(fa["Eta"][G:-G, G:-G], fb["Eta"][G:-G, G:-G])}
    ok = (fa["Mask"][G:-G, G:-G] == 1) & (fb["Mask"][G:-G, G:-G] == 1) & (fb["D"][G:-G, G:-G] > 0.5)
    for n in ("U", "V", "W"):
        x, y = fa[n][G:-G, G:-G, G:-G][:, ok], fb[n][G:-G, G:-G, G:-G][:, ok]
        err[n.lower()] = float(np.abs(x - y).max()) / max(float(np.abs(y).max()), 1e-300)
    err["mask_cells_differing"] = int((fa["Mask"] != fb["Mask"]).sum())
    return dict(workload=f"bumps{size}x{size}x{kglob}_mixed", value=a["value"], ms_per_step=a["ms_per_step"],
                steps=steps, warmup=warmup, tol=1e-8, krylov_inner_iters_per_solve=a["iters_per_solve"],
                not_converged=a["not_converged"], fp64_same_input_value=b["value"],
                errors_vs_fp64_after_steps=err,
                error_note="relative max errors over interior cells (u, v, w over wet cells deeper than 0.5 m)")


def variant_batched(count: int, steps: int, warmup: int, precision: int, seed: int = 11) -> dict:
    sc = S.ensemble(count, seed)
    cells = sum(s.mglob * s.nglob * s.kglob for s in sc)
    log(f"batched ensemble of {count} ({cells / 1e6:.1f}M cells): building")
    with Batch(sc, precision=precision, autotune=True) as b:
        b.step(warmup)
        t0 = time.perf_counter()
        b.step(steps)
        wall = time.perf_counter() - t0
        st = [m.stats() for m in b.members]
        tune = b.tuning()
    return dict(workload=f"ensemble{count}_{'mixed' if precision else 'fp64'}", scenarios=count, cells=cells,
                value=cells * 2 * steps / wall, ms_per_step=wall / steps * 1e3, steps=steps, warmup=warmup,
                tol=1e-8, not_converged=int(sum(x.not_converged for x in st)),
                iters_per_solve_max=max(x.max_iters for x in st),
                autotune={r["kernel"]: r["threads"] for r in tune})


def variant_serving(lam: float, count: int, steps: int, seed: int, tol: float, classes) -> dict:
    """One arrival rate of the serving loop (mixed precision, size classes):
    p50/p99 time-to-result and throughput."""
    from nhwave_amd.serve import Server, poisson_stream, summary
    reqs = poisson_stream(lam, count, seed=seed, steps=steps, tol=tol)
    log(f"serving variant: lambda={lam}/s, {count} requests of {steps} steps")
    with Server(precision=capi.NHW_MIXED, hbm_budget=200e9, chunk=20, classes=classes) as srv:
        t0 = time.perf_counter()
        done = srv.run(reqs)
        wall = time.perf_counter() - t0
        sm = summary(done, wall)
        sm["peak_concurrent"] = srv.peak_active
    sm.update(workload="serving", lambda_per_s=lam, steps_per_request=steps, tol=tol,
              grids="log-uniform 64^2x3 .. 1024^2x10", size_classes_cells=classes, seed=seed)
    return sm


def basin_mode(args) -> int:
    """Forced wave basin: the wavemaker and sponge kernels at full size."""
    s = S.wave_flume(mglob=2048, nglob=1024, kglob=8, dx=0.05, wavemaker="LEF_LIN", amp=0.02, per=1.0,
                     theta=10.0, bottom="bar", sponge=(0.0, 2.0, 1.0, 1.0), tol=args.tol, steps=2000)
    nprof = 0 if args.no_kernel_timing else args.profile_steps
    log(f"basin {s.name}: {args.warmup} warmup + {args.steps} timed steps")
    r = run_single(s, capi.NHW_FP64, args.steps, args.warmup, nprof)
    forcing = {}
    for k in r["kt"]:
        if k["kernel"] in ("k_wave_bc", "k_sponge"):
            k["achieved_gbs"] = k["bytes_per_launch"] / (k["avg_ms"] * 1e-3) / 1e9
            k["share_of_kernel_time"] = k["total_ms"] / sum(x["total_ms"] for x in r["kt"])
            forcing[k["kernel"]] = k
    out = {"metric": "cell-stage updates/s (forced wave basin, plain FP64)", "value": r["value"],
           "unit": "cell-stage updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": "basin2048x1024x8_fp64", "wavemaker": "LEF_LIN", "theta_deg": 10.0,
                      "sponge_m": s.sponge, "tol": s.tol, "parallelism": "single GPU"},
           "krylov_iters_per_solve": r["iters_per_solve"], "not_converged": r["not_converged"],
           "forcing_kernels": forcing, "kernels": r["kt"]}
    print(json.dumps(out), flush=True)
    return 0


def steady_mode(args) -> int:
    """Steady current at the steady-flow example's size: FLUX_LR, friction
    side walls, k-epsilon, non-hydrostatic."""
    from dataclasses import replace
    s = replace(S.steady_flume(mglob=656, nglob=100, kglob=15, dx=0.015, depth=0.3, obstacle=0.1, width=0.2,
                               tramp=5.0, zob=1.5e-5, steps=2000, tol=args.tol, bc_y=5),
                viscous=True, turb=10, visc=1.0e-6, cvs=0.001, chs=0.001)
    nprof = 0 if args.no_kernel_timing else args.profile_steps
    log(f"steady {s.name}: {args.warmup} warmup + {args.steps} timed steps")
    r = run_single(s, capi.NHW_FP64, args.steps, args.warmup, nprof)
    names = ("k_wave_bc", "k_kepsilon", "k_uvw_a", "k_diffusion", "k_visc_solve")
    total = sum(x["total_ms"] for x in r["kt"]) or 1.0
    shares = {k["kernel"]: k["total_ms"] / total for k in r["kt"]}
    out = {"metric": "cell-stage updates/s (steady current, plain FP64)", "value": r["value"],
           "unit": "cell-stage updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": "steady656x100x15_ke_fp64", "wavemaker": "FLUX_LR", "bc": s.bc,
                      "turbulence": "k-epsilon", "tol": s.tol, "parallelism": "single GPU"},
           "krylov_iters_per_solve": r["iters_per_solve"], "not_converged": r["not_converged"],
           "kernel_shares": {n: shares[n] for n in names if n in shares}, "kernels": r["kt"]}
    print(json.dumps(out), flush=True)
    return 0


def viscous_mode(args) -> int:
    """The 1024 x 1024 x 10 bumps case with the viscous branch on (constant
    viscosities): the diffusion and implicit vertical-solve kernels at full size."""
    from dataclasses import replace
    s = replace(S.bumps_beach(tol=args.tol), viscous=True, visc=1.0e-6, cvs=0.01, chs=1.0)
    nprof = 0 if args.no_kernel_timing else args.profile_steps
    log(f"viscous {s.name}: {args.warmup} warmup + {args.steps} timed steps")
    r = run_single(s, capi.NHW_FP64, args.steps, args.warmup, nprof)
    visc = {}
    for k in r["kt"]:
        if k["kernel"] in ("k_diffusion", "k_visc_sigma", "k_visc_solve"):
            k["achieved_gbs"] = k["bytes_per_launch"] / (k["avg_ms"] * 1e-3) / 1e9
            k["share_of_kernel_time"] = k["total_ms"] / sum(x["total_ms"] for x in r["kt"])
            visc[k["kernel"]] = k
    out = {"metric": "cell-stage updates/s (viscous, plain FP64)", "value": r["value"],
           "unit": "cell-stage updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": "bumps1024x1024x10_viscous_fp64", "visc": s.visc, "cvs": s.cvs, "chs": s.chs,
                      "tol": s.tol, "parallelism": "single GPU"},
           "krylov_iters_per_solve": r["iters_per_solve"], "not_converged": r["not_converged"],
           "viscous_kernels": visc, "kernels": r["kt"]}
    print(json.dumps(out), flush=True)
    return 0


def serve_mode(args) -> int:
    from nhwave_amd.serve import Server, poisson_stream, summary
    lams = [float(x) for x in args.lams.split(",")]
    res = []
    for lam in lams:
        reqs = poisson_stream(lam, args.requests, seed=args.seed, steps=args.serve_steps, tol=args.tol)
        log(f"serving lambda={lam}/s, {len(reqs)} requests")
        classes = None if args.classes == "none" else [float(x) for x in args.classes.split(",")]
        with Server(precision=capi.NHW_MIXED if args.precision == "mixed" else capi.NHW_FP64,
                    hbm_budget=args.hbm_budget * 1e9, chunk=args.chunk, classes=classes) as srv:
            t0 = time.perf_counter()
            done = srv.run(reqs)
            wall = time.perf_counter() - t0
            sm = summary(done, wall)
            sm["peak_concurrent"] = srv.peak_active
            sm["peak_footprint_gb"] = srv.peak_used / 1e9
            per = {}
            for q, ln in enumerate(srv.lanes):
                mine = [r for r in done if srv.lane_of(r) == q]
                if mine:
                    lat = np.array([r.latency for r in mine])
                    per[f"<= {ln.bound:g} cells"] = dict(requests=len(mine), p50_s=float(np.percentile(lat, 50)),
                                                         p99_s=float(np.percentile(lat, 99)))
            sm["per_class"] = per
        sm["lambda_per_s"] = lam
        res.append(sm)
        log(f"lambda {lam}: p50 {sm['p50_s']:.2f}s p99 {sm['p99_s']:.2f}s value {sm['value']:.3e}")
    best = max(res, key=lambda r: r["value"])
    out = {"metric": "cell-stage updates/s and p50/p99 time-to-result (serving)", "value": best["value"],
           "unit": "cell-stage updates/s", "n_gpus": 1, "steps": args.serve_steps, "warmup": 0,
           "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64" if args.precision == "mixed" else "f64", "data": "synthetic",
           "config": {"workload": "serving", "grids": "log-uniform 64^2x3 .. 1024^2x10",
                      "steps_per_request": args.serve_steps, "requests_per_lambda": args.requests,
                      "size_classes_cells": args.classes,
                      "precision": args.precision, "parallelism": "single GPU, continuous batching"},
           "per_lambda": res}
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--kglob", type=int, default=10)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--workload", choices=("fp64", "mixed", "batched", "serve", "basin", "viscous", "steady"), default="fp64",
                    help="fp64: the headline line (with the mixed and batched variants unless "
                         "--no-variants); mixed / batched: only that variant, as its own line")
    ap.add_argument("--solver", choices=("mg", "line"), default="mg",
                    help="Poisson solver of the fp64 workload: multigrid-preconditioned or "
                         "line-preconditioned BiCGStab")
    ap.add_argument("--no-variants", action="store_true", help="skip the mixed and batched measurements")
    ap.add_argument("--variant-steps", type=int, default=10)
    ap.add_argument("--no-compare", action="store_true",
                    help="--workload mixed: skip the FP64 run on the same input (a trace of the mixed form alone)")
    ap.add_argument("--ensemble", type=int, default=64)
    ap.add_argument("--lams", default="0.5,2,8")
    ap.add_argument("--requests", type=int, default=12)
    ap.add_argument("--serve-requests", type=int, default=12,
                    help="requests of the serving variant of the default line (lambda = 8/s); 0: skip it")
    ap.add_argument("--late-step", type=int, default=1000,
                    help="the headline run carried on to this step, then --late-window steps timed there "
                         "(variants.late_window; 0: skip)")
    ap.add_argument("--late-window", type=int, default=10)
    ap.add_argument("--serve-steps", type=int, default=1000)
    ap.add_argument("--chunk", type=int, default=20)
    ap.add_argument("--classes", default="131072,1048576",
                    help="serving size classes: ascending upper bounds on a request's cells, one batch "
                         "per class on its own stream and host thread; 'none' for one batch")
    ap.add_argument("--hbm-budget", type=float, default=200.0, help="GB")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="mixed")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--cpu-window", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-only", action="store_true",
                    help="only time the CPU baseline (--cpu-window, --cpu-steps) and print it; touches no GPU")
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
    if args.workload == "serve":
        return serve_mode(args)
    if args.workload == "basin":
        return basin_mode(args)
    if args.workload == "viscous":
        return viscous_mode(args)
    if args.workload == "steady":
        return steady_mode(args)
    if args.workload in ("mixed", "batched"):
        v = (variant_mixed(2048, 8, args.variant_steps, 2, compare=not args.no_compare) if args.workload == "mixed" else
             variant_batched(args.ensemble, args.variant_steps, 2, capi.NHW_MIXED))
        v["metric"] = f"cell-stage updates/s ({args.workload} variant)"
        v["unit"] = "cell-stage updates/s"
        v["n_gpus"] = 1
        print(json.dumps(v), flush=True)
        return 0

    s = workload(args.size, args.kglob, args.tol)
    if args.cpu_only:
        print(json.dumps(dict(cpu_baseline(s, args.cpu_window, args.cpu_steps, timeout=6 * 3600.0),
                              window=args.cpu_window, steps=args.cpu_steps)), flush=True)
        return 0
    line = args.solver == "line"
    wl = f"bumps{args.size}x{args.size}x{args.kglob}_fp64" + ("_line" if line else "")
    log(f"workload {wl}: {args.warmup} warmup + {args.steps} timed steps")
    nprof = 0 if args.no_kernel_timing else args.profile_steps
    r = run_single(s, capi.NHW_FP64, args.steps, args.warmup, nprof,
                   solver=capi.NHW_SOLVER_LINE if line else capi.NHW_SOLVER_MG,
                   late=None if args.no_variants or args.late_step <= 0 else (args.late_step, args.late_window))
    roof = roofline(r["kt"], wl, r["prof_wall"], nprof) if nprof else None
    variants = {}
    if r["late"]:
        variants["late_window"] = dict(workload=wl, **r["late"])
    if not args.no_variants:
        try:
            variants["mixed_2048"] = variant_mixed(2048, 8, args.variant_steps, 2)
        except Exception as e:  # report, do not hide
            variants["mixed_2048"] = dict(error=str(e))
        try:
            variants["batched_64"] = variant_batched(args.ensemble, args.variant_steps, 2, capi.NHW_MIXED)
        except Exception as e:
            variants["batched_64"] = dict(error=str(e))
        try:  # the headline input with the viscous branch on (constant viscosities)
            from dataclasses import replace
            sv = replace(s, viscous=True, visc=1.0e-6, cvs=0.01, chs=1.0)
            log("viscous variant")
            rv = run_single(sv, capi.NHW_FP64, args.variant_steps, 2, 0)
            variants["viscous_1024"] = dict(workload="bumps1024x1024x10_viscous_fp64", value=rv["value"],
                                            ms_per_step=rv["ms_per_step"], steps=args.variant_steps,
                                            visc=sv.visc, cvs=sv.cvs, chs=sv.chs)
        except Exception as e:
            variants["viscous_1024"] = dict(error=str(e))
        if args.serve_requests > 0:
            try:
                classes = None if args.classes == "none" else [float(x) for x in args.classes.split(",")]
                variants["serving"] = variant_serving(8.0, args.serve_requests, args.serve_steps, args.seed,
                                                      args.tol, classes)
            except Exception as e:
                variants["serving"] = dict(error=str(e))
        for v in variants.values():
            if "value" in v:
                v["vs_plain_fp64"] = v["value"] / r["value"]
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        try:
            cpu = cpu_baseline(s, args.cpu_window, args.cpu_steps)
        except Exception as e:  # report, do not hide
            cpu = dict(error=str(e))
    t_sim, dt, run_step = r["time"]
    out = {
        "metric": "cell-stage updates/s (plain FP64)",
        "value": r["value"],
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl, "mglob": s.mglob, "nglob": s.nglob, "kglob": s.kglob,
                   "dx_m": s.dx, "time_order": s.time_order, "non_hydro": True, "tol": s.tol,
                   "cfl": s.cfl, "seed": s.meta["seed"],
                   "poisson_solver": "line-BiCGStab" if line else "multigrid-BiCGStab",
                   "parallelism": "single GPU"},
        "krylov_iters_per_solve": r["iters_per_solve"],
        "sim_time_s": t_sim,
        "run_step": run_step,
        "roofline": roof,
        "cpu_baseline": cpu,
        "variants": variants,
        "kernels": r["kt"],
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
