"""The C ABI library loads and exports every entry point include/nhwave_amd.h
declares; argument validation runs on the host (no device call)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from nhwave_amd import capi
from nhwave_amd.model import make_config
from nhwave_amd import scenario as S

HEADER = Path(__file__).resolve().parents[1] / "include" / "nhwave_amd.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(nhw_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("nhw_fluxes", "nhw_eval_duvw", "nhw_update_mask", "nhw_poisson_solver",
                 "nhw_estimate_dt", "nhw_update_vars", "nhw_source_terms", "nhw_step"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []
    assert set(declared_functions()) <= set(capi.EXPORTS)


def _create(cfg):
    lib = capi.load()
    ctx = C.c_void_p()
    rc = lib.nhw_create(C.byref(cfg), C.byref(ctx))
    return rc, ctx, capi.last_error()


@pytest.mark.parametrize("patch,msg", [
    (dict(mglob=1), "grid"),
    (dict(kglob=1), "grid"),
    (dict(hllc=1), "HLLC"),
    (dict(bc_x0=3), "lateral BC"),
    (dict(bc_z0=3), "BC_Z0"),
    (dict(bc_z0=5, ibot=2, zob=0.0), "Zob"),
    (dict(time_order=4), "TIME_ORDER"),
    (dict(dx=0.0), "DX"),
    (dict(tol=0.0), "TOL"),
    (dict(precision=7), "precision"),
    (dict(solver=5), "solver"),
    (dict(kglob=17), "Kglob <= 16"),
    (dict(precision=1, solver=1), "mixed form runs the multigrid"),
    (dict(bc_y0=5), "lateral BC type 5"),
    (dict(viscous=1, schmidt=1.0, bc_x0=5, visc=0.0), "VISCOSITY > 0"),
    (dict(viscous=1, schmidt=1.0, kglob=33, solver=1), "Kglob <= 32"),
    (dict(viscous=1, schmidt=0.0), "Schmidt"),
    (dict(viscous=1, schmidt=1.0, turb=5), "turbulence model unsupported"),
    (dict(viscous=1, schmidt=1.0, turb=10, kglob=2), "k-epsilon model takes Kglob >= 3"),
])
def test_create_rejects_unsupported_config(patch, msg):
    cfg = make_config(S.solitary_channel(mglob=16, nglob=4, kglob=3))
    for k, v in patch.items():
        setattr(cfg, k, v)
    rc, ctx, err = _create(cfg)
    assert rc == capi.NHW_EINVAL and not ctx.value
    assert msg in err


def test_null_arguments():
    lib = capi.load()
    assert lib.nhw_create(None, None) == capi.NHW_EINVAL
    assert lib.nhw_step(None, 1) == capi.NHW_EINVAL
    assert lib.nhw_destroy(None) == 0


@pytest.mark.parametrize("patch,msg", [
    (dict(precision=1), "share precision"),
    (dict(time_order=3), "share precision, TIME_ORDER"),
    (dict(solver=1), "multigrid"),
])
def test_batch_create_rejects_mixed_members(patch, msg):
    lib = capi.load()
    a = make_config(S.solitary_channel(mglob=16, nglob=4, kglob=3))
    b = make_config(S.solitary_channel(mglob=24, nglob=6, kglob=4))
    for k, v in patch.items():
        setattr(b, k, v)
    cfgs = (capi.NhwConfig * 2)(a, b)
    h = C.c_void_p()
    rc = lib.nhw_batch_create(cfgs, 2, C.byref(h))
    assert rc == capi.NHW_EINVAL and not h.value
    assert msg in capi.last_error()
