"""GPU checks at the full bench sizes through size-independent properties,
plus the host-forced failure paths."""
import numpy as np
import pytest

from nhwave_amd import capi
from nhwave_amd import scenario as S
from nhwave_amd.model import Model

pytestmark = pytest.mark.gpu


def test_determinism_and_sanity_at_bumps_1024(gpu):
    """1024x1024x10 Gaussian-bump bathymetry with a drying beach: two runs of
    3 steps are bit-identical, finite, mass-consistent and converged."""
    s = S.bumps_beach(n=1024, kglob=10, steps=3)
    outs = []
    for _ in range(2):
        with Model(s) as m:
            m.step(3)
            outs.append(m.download("Eta", "D", "U", "V", "W", "P", "Mask"))
            st = m.stats()
            assert st.not_converged == 0 and st.solves == 6
            assert st.last_relres < s.tol
    a, b = outs
    for n in a:
        assert np.array_equal(a[n], b[n]), n
    g = S.NGHOST
    for n in ("Eta", "U", "V", "W", "P"):
        assert np.isfinite(a[n]).all(), n
    dry = int((a["Mask"][g:-g, g:-g] == 0).sum())
    assert dry > 0  # the beach is dry land


@pytest.mark.parametrize("non_hydro", [False, True])
def test_lake_at_rest_large(gpu, non_hydro):
    s = S.bumps_beach(n=256, kglob=8, dx=40.0, amp=0.0, non_hydro=non_hydro)
    s = s.with_(hcg=np.maximum(s.hcg, 1.0))  # fully wet: no fronts
    with Model(s) as m:
        m.step(10)
        f = m.download("Eta", "U", "V", "W")
    g = S.NGHOST
    assert np.abs(f["Eta"][g:-g, g:-g]).max() < 1e-12
    for n in ("U", "V", "W"):
        assert np.abs(f[n][g:-g, g:-g, g:-g]).max() < 1e-11


def test_mass_conservation_large(gpu):
    s = S.bumps_beach(n=256, kglob=6, dx=40.0)
    s = s.with_(hcg=np.maximum(s.hcg, 3.0))
    g = S.NGHOST
    with Model(s) as m:
        v0 = m.field("D")[g:-g, g:-g].sum()
        m.step(8)
        v1 = m.field("D")[g:-g, g:-g].sum()
    assert abs(v1 - v0) <= 1e-12 * v0


def test_dt_below_dt_min_is_an_error(gpu):
    """estimate_dt's `stop` (src/nhwave.F:3840-3850) becomes NHW_EDT."""
    s = S.solitary_channel(mglob=32, nglob=4, kglob=3, x0=0.8, non_hydro=False).with_(dt_min=1.0)
    with Model(s) as m:
        with pytest.raises(capi.NhwError) as e:
            m.step(1)
        assert e.value.code == capi.NHW_EDT


def test_itmax_reached_is_counted(gpu):
    """An unconverged Poisson solve is not fatal (the reference ignores
    NSPCG's ier); it is counted in the solver statistics."""
    s = S.solitary_channel(mglob=32, nglob=4, kglob=3, x0=0.8).with_(itmax=1, tol=1e-14)
    with Model(s) as m:
        m.step(2)
        st = m.stats()
    assert st.not_converged == st.solves == 4
    assert st.max_iters == 1


@pytest.mark.parametrize("shape", [(256, 256, 10), (203, 77, 7), (96, 5, 16)])
def test_multigrid_and_line_solvers_agree(gpu, shape):
    """The two Poisson solvers (multigrid-preconditioned matrix-free BiCGStab
    and line-preconditioned BiCGStab on the assembled matrix) solve the same
    system: after two steps the states agree to the solve tolerance, and the
    multigrid solve needs far fewer iterations.  Odd and thin grids exercise
    the uneven coarsening paths."""
    m_, n_, k_ = shape
    s = S.bumps_beach(n=n_, m=m_, kglob=k_, dx=10.0, tol=1e-11)
    res = {}
    for solver in (capi.NHW_SOLVER_MG, capi.NHW_SOLVER_LINE):
        with Model(s, solver=solver) as m:
            m.step(2)
            res[solver] = (m.download("Eta", "U", "W", "P"), m.stats())
    (a, sa), (b, sb) = res[capi.NHW_SOLVER_MG], res[capi.NHW_SOLVER_LINE]
    assert sa.not_converged == sb.not_converged == 0
    assert sa.last_relres < s.tol and sb.last_relres < s.tol
    g = S.NGHOST
    for name, tol in (("Eta", 1e-12), ("U", 1e-8), ("W", 1e-7), ("P", 1e-8)):
        x, y = a[name][..., g:-g, g:-g], b[name][..., g:-g, g:-g]
        err = float(np.abs(x - y).max()) / max(float(np.abs(y).max()), 1e-300)
        assert err < tol, (shape, name, err)
    assert sa.max_iters <= 30, sa.max_iters
    if m_ * n_ >= 256 * 256:
        assert sa.max_iters * 5 < sb.max_iters, (sa.max_iters, sb.max_iters)


@pytest.mark.parametrize("case", ["mixed", "batch", "east_open_without_flux_lr", "no_roughness"])
def test_flux_lr_refusals(gpu, case):
    """FLUX_LR runs in FP64 single contexts with BC_X0 = BC_Xn = 3 and a
    positive Zob; an open east boundary is FLUX_LR's alone."""
    from nhwave_amd.batch import Batch
    s = S.steady_flume(mglob=24, nglob=2, kglob=3, steps=1, non_hydro=False)
    with pytest.raises((capi.NhwError, RuntimeError)):
        if case == "mixed":
            Model(s, precision=capi.NHW_MIXED).close()
        elif case == "batch":
            Batch([s]).close()
        elif case == "east_open_without_flux_lr":
            Model(s.with_(wavemaker="NONE")).close()
        else:
            Model(s.with_(zob=0.0)).close()
