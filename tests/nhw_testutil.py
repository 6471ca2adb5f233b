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
