"""nhw_snapshot_static, the snapshot entry point that allocates nothing,
against nhw_snapshot_begin + nhw_snapshot_wait through the C ABI: the same
bytes, on two calls in a row (the second with a smaller request on a smaller
context), in FP64 and in the mixed form (whose FP32 fields it widens on the
device), and a request beyond its fixed staging size refused."""
import ctypes as C

import numpy as np
import pytest

from nhwave_amd import capi
from nhwave_amd import scenario as S
from nhwave_amd.model import INT_FIELDS, Model, field_shape

pytestmark = pytest.mark.gpu

BIG = ("Eta", "D", "Mask", "Mask9", "Hfx0", "EtayR", "U", "V", "W", "DU", "DW0", "P", "Omega", "Ex", "Gy", "Hz")
SMALL = ("Eta", "Mask", "W", "P")
LIMIT = 128 << 20


def _arrays(m, names):
    return {n: np.full(field_shape(m.s, n), -7, dtype=np.int32 if n in INT_FIELDS else np.float64)
            for n in names}


def static_snapshot(m, names):
    out = _arrays(m, names)
    f = m._fields_struct(out)
    capi.check(m.lib.nhw_snapshot_static(m.ctx, C.byref(f)), "nhw_snapshot_static")
    return out


def reference_snapshot(m, names):
    out = m.snapshot_begin(*names)
    m.snapshot_wait()
    return out


def same(a, b):
    return [n for n in a if a[n].tobytes() != b[n].tobytes()]


@pytest.mark.parametrize("precision", [capi.NHW_FP64, capi.NHW_MIXED])
def test_static_snapshot_equals_snapshot_twice(gpu, precision):
    big = S.bumps_beach(n=96, m=80, kglob=6, dx=100.0, steps=10, tol=1e-10)
    small = S.bumps_beach(n=24, m=20, kglob=3, dx=100.0, steps=10, tol=1e-10)
    with Model(big, precision=precision) as a, Model(small, precision=precision) as b:
        a.step(3)
        b.step(2)
        got1 = static_snapshot(a, BIG)
        want1 = reference_snapshot(a, BIG)
        assert same(got1, want1) == []
        got2 = static_snapshot(b, SMALL)   # second call: smaller request, smaller context
        want2 = reference_snapshot(b, SMALL)
        assert same(got2, want2) == []
        # stepping on between calls: the next static snapshot sees the new state
        a.step(1)
        got3 = static_snapshot(a, ("Eta", "U"))
        want3 = reference_snapshot(a, ("Eta", "U"))
        assert same(got3, want3) == []
        assert not np.array_equal(got3["Eta"], got1["Eta"])
    if precision == capi.NHW_MIXED:  # FP32 storage widened exactly
        u = got1["U"]
        assert np.array_equal(u.astype(np.float32).astype(np.float64), u)


def test_static_snapshot_refuses_oversize_request(gpu):
    s = S.bumps_beach(n=384, kglob=16, dx=100.0, steps=5, non_hydro=False)
    names = ("U", "V", "W", "DU", "DV", "DW", "DU0", "DV0", "DW0")
    with Model(s) as m:
        need = sum(-(-int(np.prod(field_shape(s, n))) * 8 // 256) * 256 for n in names)
        assert need > LIMIT
        out = _arrays(m, names)
        f = m._fields_struct(out)
        rc = m.lib.nhw_snapshot_static(m.ctx, C.byref(f))
        assert rc == capi.NHW_EINVAL
        assert "more than the" in capi.last_error()
        assert all(np.all(out[n] == -7) for n in names)   # nothing written
        # the context is still usable, and a request that fits works after the refusal
        m.step(1)
        got = static_snapshot(m, ("Eta", "U"))
        want = reference_snapshot(m, ("Eta", "U"))
        assert same(got, want) == []
