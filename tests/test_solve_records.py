"""What the reference's Poisson solves achieve as true residuals
(tests/golden/solves_*.json, written by make_solve_golden.py from the
reference build with NSPCG's driver wrapped): every solve ends without an
NSPCG error (TOL below 500 eps is reset to it, with a warning) and below its
starting residual, the logged right side is the
oracle's for the same state, and NSPCG's own stopping test leaves true
relative residuals well above TOL at TOL 1e-10 (the records' content is the
point: the GPU tests hold the GPU's true residual against them)."""
import json

import numpy as np
import pytest

import oracle as O
from nhw_testutil import GOLD
from nhwave_amd import scenario as S

RECORDS = sorted(p.stem[len("solves_"):] for p in GOLD.glob("solves_*.json"))


def record(name):
    return json.loads((GOLD / f"solves_{name}.json").read_text())


@pytest.mark.parametrize("name", RECORDS)
def test_solve_record_is_consistent(name):
    r = record(name)
    n = len(r["ier"])
    assert n == 2 * r["steps"]  # two stages per step (TIME_ORDER = SECOND)
    assert r["run_step"] == [1 + q // 2 for q in range(n)]
    # ier = 2: ZETA below 500 times the machine relative precision, reset to
    # it (src/nspcg.f:17438-17445); the solve goes on
    assert all(e == (0 if r["tol"] >= 500 * np.finfo(float).eps else 2) for e in r["ier"])
    b, r0, r1 = (np.array(r[k]) for k in ("rhs_norm", "initial_resid", "final_resid"))
    assert np.all(r1 < r0) and np.all(b > 0)


@pytest.mark.parametrize("name", RECORDS)
def test_first_solve_right_side_is_the_oracles(name):
    """The first solve's ||Rhs|| equals that of the oracle's system for the
    same state (the state is bit-identical up to the first solve)."""
    r = record(name)
    s = getattr(S, r["generator"])(**r["kwargs"])
    o = O.Oracle(s)
    o.call("orc_estimate_dt")
    o.call("orc_update_mask")
    o.call("orc_update_vars")
    o.call("orc_source_terms")
    o.call("orc_fluxes")
    assert o.call("orc_eval_duvw", 1) == 0
    bn = float(np.sqrt(np.sum(o.field("Rhs") ** 2)))
    assert abs(bn - r["rhs_norm"][0]) <= 1e-13 * bn, (bn, r["rhs_norm"][0])


def test_nspcg_leaves_true_residuals_above_tol():
    """At TOL 1e-10 NSPCG's test stops with true relative residuals of a few
    1e-9 (the 512 x 8 x 3 channel); at TOL 1e-13 with about 1e-10."""
    a = record("solitary512_tol1e-10")
    rel = np.array(a["final_resid"]) / np.array(a["rhs_norm"])
    assert 1e-9 < rel.min() and rel.max() < 1e-8
    t = record("solitary512_tight")
    rel = np.array(t["final_resid"]) / np.array(t["rhs_norm"])
    assert 1e-11 < rel.min() and rel.max() < 1e-9
