"""The Poisson system on the GPU against generate_coef_rhs
(src/pressure.F:475-840): the matrix and right side the GPU assembles
(nhw_generate_coef_rhs, the line solver's operator) entry by entry against
the oracle's assembly of the same state, and the multigrid solver's
matrix-free operator (nhw_poisson_apply) against the product with that
assembled matrix."""
from dataclasses import replace

import numpy as np
import pytest

import oracle as O
from nhwave_amd import capi
from nhwave_amd import scenario as S
from nhwave_amd.model import Model
from nhw_testutil import spmv

pytestmark = pytest.mark.gpu

# the state generate_coef_rhs reads (ghost cells included)
STATE = ("Eta", "D", "DelxH", "DelyH", "DelxEta", "DelyEta", "U", "V", "W", "Mask")

def cases():
    r = S.random_scenario(12, mglob=20, nglob=15, kglob=4, steps=3, tol=1e-13)
    b = S.bumps_beach(n=40, kglob=5, dx=250.0, steps=3, tol=1e-12)
    f = S.wave_flume(mglob=60, nglob=1, wavemaker="LEF_SOL", amp=0.05, depth=0.25, dx=0.05, bottom="beach",
                     steps=3, tol=1e-13)
    c = replace(S.solitary_channel(mglob=48, nglob=4, kglob=6, x0=1.2, steps=3, tol=1e-13), ivgrd=2, grd_r=1.2)
    return {"random_islands": r, "bumps_wetdry": b, "flume_1row": f, "channel_stretched_sigma": c}


CASES = cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_poisson_system_vs_oracle(gpu, name):
    s = CASES[name]
    rng = np.random.default_rng(5)
    n = s.mglob * s.nglob * s.kglob
    with Model(s) as m:
        m.step(2)
        st = m.download(*STATE)
        t, dt, rs = m.time()
        coef, rhs = m.poisson_system()
        x = rng.standard_normal(n)
        y = m.poisson_apply(x)
    o = O.Oracle(s)
    for k, a in st.items():
        o.field(k)[...] = a
    o.call("orc_set_time", t, dt, rs)
    o.call("orc_interpolate_velocity_to_faces")
    o.call("orc_generate_coef_rhs")
    oc, orhs = o.field("Coef"), o.field("Rhs")
    # the assembled matrix and right side: entry by entry
    bad = [d for d in range(15) if not np.array_equal(coef[d], oc[d])]
    assert bad == [], (name, bad, [float(np.abs(coef[d] - oc[d]).max()) for d in bad])
    assert np.array_equal(rhs, orhs), (name, float(np.abs(rhs - orhs).max()))
    # the matrix-free operator: products by precomputed reciprocals instead
    # of quotients, so agreement to rounding of each row's sum
    ya = spmv(coef, x, s)
    scale = spmv(coef, x, s, absolute=True)
    err = float(np.max(np.abs(y - ya) / scale))
    assert err < 1e-14, (name, err)


def test_gpu_poisson_apply_mixed_context(gpu):
    """The mixed form keeps the FP64 operator for its outer residual: the same
    matrix-free product as the plain form on the same state."""
    s = CASES["random_islands"]
    x = np.random.default_rng(6).standard_normal(s.mglob * s.nglob * s.kglob)
    with Model(s) as a, Model(s, precision=capi.NHW_MIXED) as b:
        # put the plain context's state into the mixed one (its 3-D fields are
        # FP32, so copy only what the operator reads: the 2-D fields)
        a.step(1)
        st = a.download("D", "DelxH", "DelyH", "DelxEta", "DelyEta", "Eta")
        b.upload(st)
        ya, yb = a.poisson_apply(x), b.poisson_apply(x)
    assert np.array_equal(ya, yb)
