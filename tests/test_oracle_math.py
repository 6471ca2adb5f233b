"""Checks of the oracle against what the mathematics fixes, independent of
the reference run: closed forms, invariants and brute force on tiny cases."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from nhwave_amd import scenario as S

_DP = C.POINTER(C.c_double)
# 15-point stencil (di, dj, dk) in the reference's JCoef order, written out
# from the diagonal comments of src/pressure.F:817-831
STENCIL = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, -1, 1), (-1, 0, 1), (0, 0, 1), (1, 0, 1), (0, 1, 1),
           (-1, 0, 0), (0, -1, 0), (0, 1, -1), (1, 0, -1), (0, 0, -1), (-1, 0, -1), (0, -1, -1)]


def dense_matrix(s, coef):
    M, N, K = s.mglob, s.nglob, s.kglob
    n = M * N * K
    A = np.zeros((n, n))
    for k in range(K):
        for j in range(N):
            for i in range(M):
                r = i + M * (j + N * k)
                for d, (di, dj, dk) in enumerate(STENCIL):
                    a = coef[d, r]
                    if a == 0.0:
                        continue
                    i1, j1, k1 = i + di, j + dj, k + dk
                    assert 0 <= i1 < M and 0 <= j1 < N and 0 <= k1 < K, "coefficient leaves the domain"
                    A[r, i1 + M * (j1 + N * k1)] += a
    return A


def assembled(s, steps=1):
    o = O.Oracle(s)
    if steps:
        o.step(steps)
    o.call("orc_generate_coef_rhs")
    return o, o.field("Coef").copy(), o.field("Rhs").copy()


def test_flat_poisson_closed_form():
    """Flat bottom, flat surface: DelxS = DelyS = 0, so the 8 cross
    coefficients vanish and the operator is a 7-point Laplacian whose
    vertical coefficients are A1/(0.5*(dsig_k+dsig_k-1)*dsig) with A1 = 1/D^2."""
    s = S.solitary_channel(mglob=7, nglob=5, kglob=4, dx=0.3, depth=2.0, amp_ratio=0.0, steps=1)
    s = s.with_(ug=np.zeros_like(s.ug), dy=0.45)
    o, coef, rhs = assembled(s, steps=0)
    M, N, K = s.mglob, s.nglob, s.kglob
    dsig, _, _ = S.vertical_grid(K)
    D = 2.0
    A1 = (1.0 / D) ** 2
    dx2, dy2 = 1.0 / s.dx ** 2, 1.0 / s.dy ** 2
    for d in (3, 4, 6, 7, 10, 11, 13, 14):
        assert np.all(coef[d] == 0.0)
    for k in range(K):
        ds = dsig[k + 2]
        up = A1 / (0.5 * (ds + ds) * ds)
        for j in range(N):
            for i in range(M):
                r = i + M * (j + N * k)
                c1 = 2 * dx2 + 2 * dy2 + 2 * up
                c2 = c9 = -dx2
                c3 = c10 = -dy2
                c6 = c13 = -up
                if i == 0:
                    c1 += c9; c9 = 0.0
                if i == M - 1:
                    c1 += c2; c2 = 0.0
                if j == 0:
                    c1 += c10; c10 = 0.0
                if j == N - 1:
                    c1 += c3; c3 = 0.0
                if k == 0:
                    c6 += c13; c13 = 0.0
                if k == K - 1:
                    c6 = 0.0
                want = {0: c1, 1: c2, 2: c3, 5: c6, 8: c9, 9: c10, 12: c13}
                for d, v in want.items():
                    assert coef[d, r] == pytest.approx(v, rel=1e-14, abs=1e-12), (i, j, k, d)
    assert np.all(rhs == 0.0)  # fluid at rest


@pytest.mark.parametrize("seed", [3, 4])
def test_poisson_row_sums(seed):
    """Constants lie in the kernel of the Neumann-folded operator: each row
    sums to zero except at the Dirichlet surface row, where the sum is minus
    the dropped upward coupling (c4..c8 before zeroing)."""
    s = S.random_scenario(seed, mglob=9, nglob=7, kglob=4, dry=False)
    o, coef, _ = assembled(s)
    M, N, K = s.mglob, s.nglob, s.kglob
    sums = coef.sum(axis=0).reshape(K, N, M)
    scale = np.abs(coef).max()
    assert np.abs(sums[:-1]).max() < 1e-12 * scale
    assert np.abs(sums[-1]).min() > 1e-8 * scale  # the surface rows carry the Dirichlet term


def test_banded_solver_matches_dense():
    """The oracle's banded LU solve equals a dense solve of the assembled
    system (tiny grid, random bathymetry with dry cells and a non-trivial
    right-hand side)."""
    s = S.random_scenario(5, mglob=6, nglob=5, kglob=3)
    o, coef, rhs = assembled(s)
    A = dense_matrix(s, coef)
    x_dense = np.linalg.solve(A, rhs)
    x = np.zeros_like(rhs)
    cc = np.ascontiguousarray(coef)
    rc = O.load().orc_solve_banded(o.h, cc.ctypes.data_as(_DP), rhs.ctypes.data_as(_DP), x.ctypes.data_as(_DP))
    assert rc == 0
    assert np.abs(x - x_dense).max() <= 1e-11 * np.abs(x_dense).max()
    assert np.abs(A @ x - rhs).max() <= 1e-12 * np.abs(rhs).max()
    # the diagonal-storage product of the GPU tests (nhw_testutil.spmv, the
    # JCoef offsets) is this dense matrix's product
    from nhw_testutil import spmv
    v = np.random.default_rng(3).standard_normal(rhs.size)
    assert np.abs(spmv(coef, v, s) - A @ v).max() <= 1e-13 * np.abs(A).sum(axis=1).max()


@pytest.mark.parametrize("non_hydro", [False, True])
def test_lake_at_rest(non_hydro):
    """Well-balanced bed-slope source (source_terms vs the hydrostatic flux
    term): still water over rough bathymetry stays still."""
    s = S.random_scenario(7, mglob=20, nglob=14, kglob=3, dry=False, non_hydro=non_hydro)
    z = np.zeros_like(s.ug)
    s = s.with_(etag=np.zeros_like(s.etag), ug=z, vg=z.copy(), wg=z.copy(), hcg=np.abs(s.hcg) + 0.3)
    o = O.Oracle(s)
    o.step(30)
    g = S.NGHOST
    assert np.abs(o.field("Eta")[g:-g, g:-g]).max() < 1e-13
    for f in ("U", "V", "W"):
        assert np.abs(o.field(f)[g:-g, g:-g, g:-g]).max() < 1e-12


@pytest.mark.parametrize("non_hydro", [False, True])
def test_mass_conservation(non_hydro):
    """Walls all round and no drying: total volume sum(D) dx dy is conserved."""
    s = S.random_scenario(8, mglob=16, nglob=12, kglob=3, dry=False, non_hydro=non_hydro)
    s = s.with_(hcg=np.abs(s.hcg) + 0.5)
    o = O.Oracle(s)
    g = S.NGHOST
    v0 = o.field("D")[g:-g, g:-g].sum()
    o.step(20)
    v1 = o.field("D")[g:-g, g:-g].sum()
    assert abs(v1 - v0) <= 1e-12 * v0


def test_y_invariance():
    """A y-invariant channel keeps V = 0 and Eta y-invariant to round-off
    (hydrostatic; the wall faces carry the exact pressure term while interior
    faces carry its HLL average, which rounds)."""
    s = S.solitary_channel(mglob=40, nglob=5, kglob=3, x0=1.0, steps=1, non_hydro=False)
    o = O.Oracle(s)
    o.step(25)
    g = S.NGHOST
    eta = o.field("Eta")[g:-g, g:-g]
    umax = np.abs(o.field("U")).max()
    assert np.abs(o.field("V")[g:-g, g:-g, g:-g]).max() < 1e-14 * umax
    assert np.abs(eta - eta[0]).max() < 1e-15


def test_x_mirror_symmetry():
    """Reflecting the scenario in x reflects the solution (U changes sign)."""
    s = S.random_scenario(9, mglob=14, nglob=6, kglob=3, non_hydro=False)
    sm = s.with_(hcg=s.hcg[:, ::-1].copy(), etag=s.etag[:, ::-1].copy(),
                 ug=-s.ug[:, :, ::-1].copy(), vg=s.vg[:, :, ::-1].copy(), wg=s.wg[:, :, ::-1].copy())
    a, b = O.Oracle(s), O.Oracle(sm)
    a.step(4)
    b.step(4)
    g = S.NGHOST
    ea, eb = a.field("Eta")[g:-g, g:-g], b.field("Eta")[g:-g, g:-g][:, ::-1]
    assert np.abs(ea - eb).max() < 1e-13
    ua, ub = a.field("U")[g:-g, g:-g, g:-g], -b.field("U")[g:-g, g:-g, g:-g][:, :, ::-1]
    assert np.abs(ua - ub).max() < 1e-12 * max(np.abs(ua).max(), 1.0)
    assert np.array_equal(a.field("Mask"), b.field("Mask")[:, ::-1])


def test_update_mask_truth_table():
    """Hand-built wet-dry patches for update_mask (src/nhwave.F:5403-5456)."""
    s = S.random_scenario(10, mglob=6, nglob=6, kglob=2, dry=False)
    o = O.Oracle(s)
    md = s.min_dep
    g = S.NGHOST
    mask, eta, d, hc = o.field("Mask"), o.field("Eta"), o.field("D"), o.field("Hc")
    ms = o.field("Mask_Struct")
    mask[:] = 1
    hc[:] = 1.0
    eta[:] = 0.0
    d[:] = 1.0
    ghost_before = mask.copy()
    c = lambda i, j: (g + j, g + i)  # noqa: E731  interior (i,j) -> array index
    # (a) dry cell, wet higher neighbour with D-MinDep > 1e-6: floods
    mask[c(1, 1)] = 0; eta[c(1, 1)] = -0.5; eta[c(0, 1)] = 0.1
    # (b) dry cell whose only higher wet neighbour has D-MinDep = 5e-7: stays dry
    for q in [(3, 2), (5, 2), (4, 3)]:
        eta[c(*q)] = -1.0  # lower neighbours: no flooding from them
    mask[c(4, 2)] = 0; eta[c(4, 2)] = -0.5; d[c(4, 1)] = md + 5e-7; eta[c(4, 1)] = 0.2
    # (c) wet cell with |D-MinDep| <= 1e-6 dries and resets Eta
    d[c(2, 4)] = md + 4e-7
    # (d) structurally dry cell never floods
    ms[c(0, 4)] = 0; mask[c(0, 4)] = 0; eta[c(0, 4)] = -0.9
    o.call("orc_update_mask")
    assert mask[c(1, 1)] == 1
    assert mask[c(4, 2)] == 0
    assert mask[c(2, 4)] == 0 and eta[c(2, 4)] == md - hc[c(2, 4)] and d[c(2, 4)] == eta[c(2, 4)] + hc[c(2, 4)]
    assert mask[c(0, 4)] == 0
    # ghost masks are never touched
    ring = np.ones_like(mask, dtype=bool)
    ring[g:-g, g:-g] = False
    assert np.array_equal(mask[ring], ghost_before[ring])


def _cnoidal_lib():
    lib = O.load()
    for f in ("orc_cnoidal_cn", "orc_cnoidal_ck", "orc_cnoidal_ce"):
        getattr(lib, f).restype = C.c_double
    lib.orc_cnoidal_cn.argtypes = [C.c_double, C.c_double]
    lib.orc_cnoidal_ck.argtypes = [C.c_double]
    lib.orc_cnoidal_ce.argtypes = [C.c_double]
    lib.orc_cnoidal_setup.restype = C.c_int
    lib.orc_cnoidal_setup.argtypes = [C.c_double, C.c_double, C.c_double, _DP]
    return lib


@pytest.mark.parametrize("m", [0.1, 0.5, 0.9, 0.999, 0.99999999])
def test_cnoidal_elliptic_functions_vs_scipy(m):
    """The wavemaker's AGM/Landen helpers (src/waves.F:225-361) are K(m),
    E(m) and cn(u|m) of the parameter convention m = k**2: against scipy's
    independent implementations (the pi literal is binary32, 2.8e-8)."""
    from scipy import special
    lib = _cnoidal_lib()
    assert lib.orc_cnoidal_ck(m) == pytest.approx(special.ellipk(m), rel=1e-7)
    assert lib.orc_cnoidal_ce(m) == pytest.approx(special.ellipe(m), rel=1e-7)
    K = special.ellipk(m)
    for u in np.linspace(-3.0 * K, 3.0 * K, 41):
        assert lib.orc_cnoidal_cn(u, m) == pytest.approx(special.ellipj(u, m)[1], abs=1e-7)


@pytest.mark.parametrize("amp,per,depth", [(0.05, 2.0, 0.4), (0.125, 2.0, 0.4), (0.02, 3.0, 0.3)])
def test_cnoidal_wavemaker_signal(amp, per, depth):
    """The LEF_CON boundary surface (src/waves.F:92-170) over one period:
    periodic in PER, crest-to-trough height AMP, zero mean (the trough level
    of first-order cnoidal theory), and it starts at a zero crossing."""
    lib = _cnoidal_lib()
    out = np.zeros(5)
    assert lib.orc_cnoidal_setup(amp, depth, per, out.ctypes.data_as(_DP)) == 0
    L, cel, yt, m, xs = out
    assert 0.0 < m < 1.0 and L > 0 and yt < 0
    s = S.wave_flume(mglob=8, nglob=3, kglob=3, wavemaker="LEF_CON", amp=amp, per=per, depth=depth,
                     bottom="flat", non_hydro=False)
    o = O.Oracle(s)
    ts = np.linspace(0.0, per, 4001)
    eta = []
    for t in ts:
        o.call("orc_set_time", C.c_double(t), C.c_double(0.01), 1)
        o.call("orc_update_wave_bc")
        e = o.field("Ein_X0")
        assert np.all(e == e[0])  # uniform along the boundary
        eta.append(e[0])
    eta = np.array(eta)
    assert abs(eta[0]) <= 1e-6
    assert abs(eta[-1] - eta[0]) <= 1e-7 * amp
    assert eta.max() - eta.min() == pytest.approx(amp, rel=1e-4)
    assert abs(np.trapezoid(eta, ts) / per) <= 1e-6 * amp


def test_internal_wavemaker_radiates_its_amplitude():
    """The internal wavemaker's mass source (src/source_terms.F:194-305) has
    the volume flux Celerity*Amp*cos(...) per unit width (Source_Area
    normalises over the box), which radiates a wave of amplitude Amp/2 =
    Wave_Spc2d each way: a single long wave (kh = 0.1) in a flat channel,
    hydrostatic, measured 5 m from the source after its front has passed
    (the scheme's damping takes 2.4 % there at dx = 0.25 m, 1.2 % at 0.125 m;
    a factor-2 slip in the source strength or a wrong area would show)."""
    a = 0.002
    s = S.internal_wave_basin(mglob=800, nglob=3, kglob=3, dx=0.25, dy=0.5, depth=0.3, nfreq=1, fmin=0.1,
                              fmax=0.1, dirs=(0.0,), box=(99.5, 100.5, 0.0, 1.5), beach=False, non_hydro=False)
    s = s.with_(irr_spc=np.array([[a]]), irr_phs=np.array([[0.3]]), dt_max=0.2, cfl=0.5)
    o = O.Oracle(s)
    g = S.NGHOST
    i_probe = int(105.0 / 0.25)  # x = 105 m, 5 m east of the source
    eta = []
    while o.time()[0] < 50.0:
        o.step(1)
        if o.time()[0] > 30.0:  # the front (c = 1.7 m/s) has passed by t = 15 s
            eta.append(o.field("Eta")[g + 1, g + i_probe])
    eta = np.array(eta)
    amp = 0.5 * (eta.max() - eta.min())
    assert amp == pytest.approx(a, rel=0.04), amp


def test_binary32_literals():
    """DOUBLE_PRECISION build semantics of default-REAL literals."""
    dsig, sig, sigc = S.vertical_grid(3)
    assert dsig[0] == 0.3333333432674408 and dsig[0] != 1.0 / 3.0
    assert sig[2 + 3] != 1.0  # sig(Kend1) accumulates the binary32 spacing
    dsig10, _, _ = S.vertical_grid(10)
    assert dsig10[0] == 0.10000000149011612
    assert S.GRAV_F32 == 9.8100004196167
