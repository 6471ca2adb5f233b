"""Shared helpers of the test suite (not a test module)."""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np

from nhwave_amd import scenario as S

GOLD = Path(__file__).resolve().parent / "golden"
INDEX = json.loads((GOLD / "golden_index.json").read_text())

FIELDS_2D = ("Eta", "D")
FIELDS_3D = ("U", "V", "W", "DU", "DV", "DW")
VEL_GROUPS = {"U": "vel", "V": "vel", "W": "vel", "DU": "mom", "DV": "mom", "DW": "mom"}


def input_digest(s) -> str:
    h = hashlib.sha256()
    for a in (s.hcg, s.etag, s.ug, s.vg, s.wg):
        if a is not None:
            h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def golden_case(name: str):
    meta = INDEX[name]
    s = getattr(S, meta["generator"])(**meta["kwargs"])
    assert input_digest(s) == meta["input_sha256"], f"{name}: scenario generator drifted"
    data = np.load(GOLD / f"{name}.npz")
    return s, meta, data


def interior_state(s, get) -> dict:
    """Interior cells of the compared fields; `get(name)` returns a full array."""
    g = S.NGHOST
    out = {}
    for n in FIELDS_2D:
        out[n] = get(n)[g:-g, g:-g]
    out["Mask"] = get("Mask")[g:-g, g:-g].astype(np.int8)
    for n in FIELDS_3D:
        out[n] = get(n)[g:-g, g:-g, g:-g]
    out["P"] = get("P")[g:g + s.kglob, g:-g, g:-g]
    out["Omega"] = get("Omega")[g:g + s.kglob + 1, g:-g, g:-g]
    return out


TURB = ("Tke", "Eps", "CmuVt")


def turb_golden(data, step: int) -> dict:
    """The k-epsilon state of a fixture at `step` ({} if it has none)."""
    return {n: data[f"s{step}_{n}"] for n in TURB if f"s{step}_{n}" in data.files}


def turb_interior(get) -> dict:
    """Interior columns, every level, of the k-epsilon state."""
    g = S.NGHOST
    return {n: get(n)[:, g:-g, g:-g] for n in TURB}


def golden_state(data, step: int) -> dict:
    keys = ["Eta", "D", "Mask", *FIELDS_3D, "P", "Omega"]
    return {k: data[f"s{step}_{k}"] for k in keys}


def rel_errors(a: dict, b: dict) -> dict:
    """max|a-b| / scale per field; velocities and momenta share one scale
    per group (max over the group's components), so a component that is
    identically zero in one solution is measured against the flow speed."""
    scale = {}
    for grp in ("vel", "mom"):
        comps = [n for n, g in VEL_GROUPS.items() if g == grp]
        scale[grp] = max(float(np.abs(b[n]).max()) for n in comps)
    out = {}
    for n in ("Eta", "D", *FIELDS_3D, "P", "Omega"):
        sc = scale.get(VEL_GROUPS.get(n), None) or float(np.abs(b[n]).max())
        sc = max(sc, 1e-300)
        out[n] = float(np.abs(a[n] - b[n]).max()) / sc
    return out


def bit_equal(a: dict, b: dict, names=("Eta", "D", "Mask", *FIELDS_3D, "P", "Omega")) -> list:
    """Names of fields that differ in value (== comparison; -0.0 == 0.0)."""
    return [n for n in names if not np.array_equal(a[n], b[n])]


def example_run_dir(name: str, run_dir, overrides: dict) -> None:
    """A run directory rebuilt from the fixture of one of the reference's
    example runs (tests/golden/example_<name>.json: its input.txt key/value
    pairs; example_<name>_depth.npy: its depth rows), with `overrides`."""
    import json
    from pathlib import Path
    run_dir = Path(run_dir)
    rec = json.loads((GOLD / f"example_{name}.json").read_text())
    depth = np.load(GOLD / f"example_{name}_depth.npy")
    keys = dict(rec["keys"])
    keys.update(overrides)
    run_dir.mkdir(parents=True, exist_ok=True)
    (run_dir / "results").mkdir(exist_ok=True)
    (run_dir / "input.txt").write_text("".join(f"{k} = {v}\n" for k, v in keys.items()))
    S._write_rows(run_dir / "depth.txt", depth)
    irr = GOLD / f"example_{name}_irr2d.npz"
    if irr.exists():  # the internal wavemaker's spectrum file
        z = np.load(irr)
        sp = S.Scenario(name="", mglob=1, nglob=1, kglob=1, dx=1.0, dy=1.0, hcg=None, etag=None,
                        irr_freq=z["freq"], irr_dire=z["dire"], irr_spc=z["spc"], irr_phs=z["phs"])
        S.write_irr2d(sp, run_dir / "irr2d.txt")


def hydro_overrides(steps) -> dict:
    """Hydrostatic; bounded to `steps` (None: the example's own SIM_STEPS and
    TOTAL_TIME); an output call after every step (the reference build's dump
    hook replaces `preview`); no probes (their stat.txt is not a fixture)."""
    o = {"NON_HYDRO": "F", "PLOT_INTV": "0.0", "PLOT_START": "0.0", "SCREEN_INTV": "1.0e9", "NSTAT": "0"}
    if steps is not None:
        o["SIM_STEPS"] = str(steps)
    return o


def nh_tight_overrides(steps) -> dict:
    """Non-hydrostatic as shipped, solved to TOL 1e-13, bounded to `steps`."""
    o = hydro_overrides(steps)
    o.update(NON_HYDRO="T", TOL="1.0e-13")
    return o


# JCoef offsets (di, dj, dk) of the 15 diagonals of the Poisson matrix,
# src/pressure.F:817-831
JCOEF = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, -1, 1), (-1, 0, 1), (0, 0, 1), (1, 0, 1), (0, 1, 1),
         (-1, 0, 0), (0, -1, 0), (0, 1, -1), (1, 0, -1), (0, 0, -1), (-1, 0, -1), (0, -1, -1)]


def spmv(coef, x, s, absolute=False):
    """y_r = sum_d coef[d, r] x[r + JCoef(d)] over the interior (rows i
    fastest, then j, then k); a neighbour outside the interior contributes
    nothing.  absolute: |coef| |x|, the scale of each row's sum."""
    M, N, K = s.mglob, s.nglob, s.kglob
    X = (np.abs(x) if absolute else x).reshape(K, N, M)
    Xp = np.pad(X, 1)
    y = np.zeros((K, N, M))
    for d, (di, dj, dk) in enumerate(JCOEF):
        c = coef[d].reshape(K, N, M)
        y += (np.abs(c) if absolute else c) * Xp[1 + dk:1 + dk + K, 1 + dj:1 + dj + N, 1 + di:1 + di + M]
    return y.reshape(-1)
