"""Writes what the reference's Poisson solves achieve as true residuals
(tests/golden/solves_<name>.json): the serial reference built with NSPCG's
driver wrapped by oracle/ref_solve_log.f90 (oracle/_ref/nhwave_ref_solves)
logs ||Rhs - A Phi|| before and after every solve, with A the matrix
generate_coef_rhs assembled.  NSPCG stops on its own test (IPARM(1) = 2 by
default, a preconditioned residual against RPARM(1) = TOL,
src/pressure.F:88-93), so the true residual it leaves can differ from TOL;
these records say by how much, per solve.

Scenarios are the golden fixtures' (make_golden.CASES), run to `steps`.

Run in the build container (needs oracle/_ref/nhwave_ref_solves):
    python tests/golden/make_solve_golden.py
"""
from __future__ import annotations

import json
import shutil
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(Path(__file__).resolve().parent))

import oracle as O  # noqa: E402
import make_golden as G  # noqa: E402

OUT = Path(__file__).resolve().parent
BIN = O.HERE / "_ref" / "nhwave_ref_solves"
# golden case -> steps to run
CASES = {"solitary512_tol1e-10": 100, "solitary512_tight": 100, "bumps48_nh_tight": 10,
         "beach_nh_tight": 100, "random3_nh_tight": 3}


def main() -> None:
    if not BIN.exists():
        raise SystemExit(f"{BIN} missing: run oracle/build_ref.sh first")
    for name, steps in CASES.items():
        gen, kw = G.CASES[name][:2]
        kw = dict(kw)
        kw["steps"] = steps
        s = G.make_scenario(gen, kw)
        tmp = Path(tempfile.mkdtemp(prefix="solves_"))
        try:
            O.run_reference(s, tmp, [], binary=BIN, timeout=3600)
            rows = [ln.split() for ln in (tmp / "nhw_solves.txt").read_text().splitlines()]
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        a = np.array([[float(v) for v in r] for r in rows])
        rec = {"generator": gen, "kwargs": kw, "tol": s.tol, "steps": steps,
               "run_step": a[:, 0].astype(int).tolist(), "neqns": int(a[0, 1]),
               "rhs_norm": a[:, 2].tolist(), "initial_resid": a[:, 3].tolist(), "final_resid": a[:, 4].tolist(),
               "ier": a[:, 5].astype(int).tolist()}
        (OUT / f"solves_{name}.json").write_text(json.dumps(rec, indent=0))
        rel = a[:, 4] / a[:, 2]
        print(f"{name}: {len(a)} solves, true relres min {rel.min():.3e} median {np.median(rel):.3e} "
              f"max {rel.max():.3e} (TOL {s.tol:g})", flush=True)


if __name__ == "__main__":
    main()
