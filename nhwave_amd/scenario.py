"""Scenario definitions, host-side set-up and the reference's input formats.

A scenario is what one NHWAVE run reads: the `input.txt` keys of the hot
path, the cell-centre depth (`depth.txt`, DEPTH_TYPE = CELL_CENTER) and the
initial surface and velocity (`eta0.txt`, `uvw0.txt`, INITIAL_EUVW = T)
(src/initialize.F:1821-1836, 2991-3017).

Array conventions (numpy, C order, so memory order equals the reference's
Fortran order with i fastest):
    global 2-D (Nglob, Mglob), global 3-D (Kglob, Nglob, Mglob)
    local  2-D (Nloc, Mloc),   local 3-D (Kloc, Nloc, Mloc)
    x-face 2-D (Nloc, Mloc1),  y-face 2-D (Nloc1, Mloc)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from pathlib import Path

import numpy as np

NGHOST = 2
GRAV_F32 = float(np.float32(9.81))


@dataclass
class Scenario:
    name: str
    mglob: int
    nglob: int
    kglob: int
    dx: float
    dy: float
    hcg: np.ndarray                  # (Nglob, Mglob) still-water depth, positive down
    etag: np.ndarray                 # (Nglob, Mglob)
    ug: np.ndarray | None = None     # (Kglob, Nglob, Mglob)
    vg: np.ndarray | None = None
    wg: np.ndarray | None = None
    cfl: float = 0.5
    min_dep: float = 0.01
    dt_ini: float = 0.01
    dt_min: float = 1.0e-5
    dt_max: float = 60.0
    froude_cap: float = 100.0
    froude_cap_on: bool = True
    visc: float = 1.0e-6
    slat: float = 0.0
    wind_u: float = 0.0
    wind_v: float = 0.0
    time_order: int = 2
    non_hydro: bool = True
    bc: tuple = (1, 1, 1, 1, 1)      # BC_X0, BC_Xn, BC_Y0, BC_Yn, BC_Z0
    itmax: int = 1000
    tol: float = 1.0e-10
    ivgrd: int = 1
    grd_r: float = 1.1
    steps: int = 1                   # SIM_STEPS
    total_time: float = 1.0e9        # TOTAL_TIME
    # boundary forcing: WAVEMAKER (NONE, LEF_SOL, LEF_LIN, LEF_STK, LEF_CON
    # with BC_X0 = 3; INT_IRR; FLUX_LR with BC_X0 = BC_Xn = 3), AMP, PER, DEP,
    # THETA, CUR, TRAMP; sponge widths
    # (West, East, South, North) or None for SPONGE_ON = F
    wavemaker: str = "NONE"
    amp_wave: float = 0.0
    per_wave: float = 1.0
    dep_wave: float = 1.0
    theta_wave: float = 0.0
    cur_wave: float = 0.0
    tramp: float = 0.0
    sponge: tuple | None = None
    # internal irregular wavemaker (WAVEMAKER = INT_IRR): the source box
    # (Xsource_West, Xsource_East, Ysource_Suth, Ysource_Nrth) and the
    # spectrum of irr2d.txt: frequencies (NumFreq), directions in degrees
    # (NumDir), Wave_Spc2d and Random_Phs as (NumFreq, NumDir) arrays
    source_box: tuple = (0.0, 0.0, 0.0, 0.0)
    irr_freq: np.ndarray | None = None
    irr_dire: np.ndarray | None = None
    irr_spc: np.ndarray | None = None
    irr_phs: np.ndarray | None = None
    # viscous branch with constant viscosities (VISCOUS_FLOW = T, IVTURB =
    # IHTURB = 1): vertical Cvs, horizontal Chs, Schmidt, VISCOUS_NUMBER;
    # the laminar viscosity is `visc` (VISCOSITY)
    viscous: bool = False
    cvs: float = 0.001
    chs: float = 0.001
    schmidt: float = 1.0
    visc_number: float = 0.1666667
    rng: bool = True                 # RNG: coefficient of the initial turbulent viscosity
    # IVTURB = IHTURB: 1 (constant Cvs, Chs) or 10 (the 3-D k-epsilon model)
    turb: int = 1
    # bottom friction of BC_Z0 = 5: Ibot (1: constant Cd0, else log law with Zob)
    ibot: int = 1
    cd0: float = 0.0025
    zob: float = 1.5e-5
    meta: dict = field(default_factory=dict)

    @property
    def mloc(self) -> int:
        return self.mglob + 2 * NGHOST

    @property
    def nloc(self) -> int:
        return self.nglob + 2 * NGHOST

    @property
    def kloc(self) -> int:
        return self.kglob + 2 * NGHOST

    @property
    def cells(self) -> int:
        return self.mglob * self.nglob * self.kglob

    @property
    def fcor(self) -> float:
        """fcor = 2.0*7.29e-5*sin(slat); the leading constant is a binary32
        product (src/initialize.F:494)."""
        return float(np.float32(2.0) * np.float32(7.29e-5)) * math.sin(self.slat)

    def with_(self, **kw) -> "Scenario":
        return replace(self, **kw)


# ---------------------------------------------------------------------------
# host-side set-up (restated from initialize.F; exact in FP64)
# ---------------------------------------------------------------------------

def vertical_grid(kglob: int, ivgrd: int = 1, grd_r: float = 1.1):
    """dsig (Kloc), sig (Kloc1), sigc (Kloc): generate_grid,
    src/initialize.F:1700-1727.  For IVGRD = 1, dsig = 1.0/float(Kglob) is a
    binary32 quotient."""
    kloc = kglob + 2 * NGHOST
    kb, ke = NGHOST + 1, kloc - NGHOST          # 1-based Kbeg, Kend
    dsig = np.zeros(kloc + 1)                   # 1-based scratch
    sig = np.zeros(kloc + 2)
    sigc = np.zeros(kloc + 1)
    if ivgrd == 1:
        dsig[1:] = float(np.float32(1.0) / np.float32(kglob))
    else:
        dsig[kb] = (grd_r - 1.0) / (math.pow(grd_r, float(np.float32(kglob))) - 1.0)
        for k in range(kb + 1, ke + 1):
            dsig[k] = dsig[k - 1] * grd_r
        for k in range(1, NGHOST + 1):
            dsig[kb - k] = dsig[kb + k - 1]
            dsig[ke + k] = dsig[ke - k + 1]
    sig[kb] = 0.0
    for k in range(kb + 1, kloc + 2):
        sig[k] = sig[k - 1] + dsig[k - 1]
        sigc[k - 1] = sig[k - 1] + 0.5 * dsig[k - 1]
    for k in range(kb - 1, 0, -1):
        sig[k] = sig[k + 1] - dsig[k]
        sigc[k] = sig[k + 1] - 0.5 * dsig[k]
    return dsig[1:].copy(), sig[1:].copy(), sigc[1:].copy()


def mirror_ghosts_2d(phi: np.ndarray) -> None:
    """phi_2D_coll (src/nhwave.F:5264-5330): x ghosts from the interior rows,
    then y ghosts over the full width (corners included)."""
    g = NGHOST
    phi[g:-g, 0] = phi[g:-g, g + 1]
    phi[g:-g, 1] = phi[g:-g, g]
    phi[g:-g, -1] = phi[g:-g, -g - 2]
    phi[g:-g, -2] = phi[g:-g, -g - 1]
    phi[0, :] = phi[g + 1, :]
    phi[1, :] = phi[g, :]
    phi[-1, :] = phi[-g - 2, :]
    phi[-2, :] = phi[-g - 1, :]


def bathymetry_fields(s: Scenario) -> dict:
    """read_bathymetry for DEPTH_TYPE = CELL_CENTER (src/initialize.F:1821-2175):
    Hc with mirrored ghosts, Mask_Struct, face depths Hfx/Hfy and DelxH/DelyH."""
    g = NGHOST
    hc = np.zeros((s.nloc, s.mloc))
    hc[g:-g, g:-g] = s.hcg
    mirror_ghosts_2d(hc)
    mask_struct = np.ones((s.nloc, s.mloc), dtype=np.int32)
    interior = mask_struct[g:-g, g:-g]
    interior[hc[g:-g, g:-g] < -1000.0] = 0
    hfx = np.zeros((s.nloc, s.mloc + 1))
    hfx[:, 1:s.mloc] = 0.5 * (hc[:, :-1] + hc[:, 1:])
    hfx[:, 0] = hc[:, 0]
    hfx[:, s.mloc] = hc[:, -1]
    hfy = np.zeros((s.nloc + 1, s.mloc))
    hfy[1:s.nloc, :] = 0.5 * (hc[:-1, :] + hc[1:, :])
    hfy[0, :] = hc[0, :]
    hfy[s.nloc, :] = hc[-1, :]
    delxh = (hfx[:, 1:] - hfx[:, :-1]) / s.dx
    delyh = (hfy[1:, :] - hfy[:-1, :]) / s.dy
    return dict(Hc=hc, Mask_Struct=mask_struct, Hfx=hfx, Hfx0=hfx.copy(), Hfy=hfy,
                Hfy0=hfy.copy(), DelxH=delxh, DelyH=delyh)


def initial_fields(s: Scenario) -> dict:
    """The loaded part of `initial` with INITIAL_EUVW = T (src/initialize.F:
    2991-3045): Eta and U,V,W on the interior, zero elsewhere."""
    g = NGHOST
    eta = np.zeros((s.nloc, s.mloc))
    eta[g:-g, g:-g] = s.etag
    out = {"Eta": eta}
    for name, arr in (("U", s.ug), ("V", s.vg), ("W", s.wg)):
        f = np.zeros((s.kloc, s.nloc, s.mloc))
        if arr is not None:
            f[g:-g, g:-g, g:-g] = arr
        out[name] = f
    return out


# ---------------------------------------------------------------------------
# synthetic scenario families
# ---------------------------------------------------------------------------

def solitary_channel(mglob=512, nglob=8, kglob=3, dx=0.05, depth=0.5, amp_ratio=0.2,
                     x0=5.0, steps=400, tol=1e-10, non_hydro=True, time_order=2) -> Scenario:
    """Flat-bottom channel with a solitary wave (first-order KdV profile):
    eta = A sech^2(k(x-x0)), k = sqrt(3A/(4h^3)), u = c eta/(h+eta),
    c = sqrt(g(h+A)), v = w = 0; walls all round."""
    a = amp_ratio * depth
    xc = (np.arange(mglob) + 0.5) * dx
    kk = math.sqrt(3.0 * a / (4.0 * depth ** 3))
    eta1 = a / np.cosh(kk * (xc - x0)) ** 2
    c = math.sqrt(9.81 * (depth + a))
    u1 = c * eta1 / (depth + eta1)
    etag = np.tile(eta1, (nglob, 1))
    ug = np.tile(u1, (kglob, nglob, 1))
    z = np.zeros((kglob, nglob, mglob))
    return Scenario(
        name=f"solitary_{mglob}x{nglob}x{kglob}", mglob=mglob, nglob=nglob, kglob=kglob,
        dx=dx, dy=dx, hcg=np.full((nglob, mglob), depth), etag=etag, ug=ug, vg=z.copy(),
        wg=z.copy(), dt_ini=0.01, steps=steps, tol=tol, non_hydro=non_hydro,
        time_order=time_order, meta=dict(amplitude=a, x0=x0, depth=depth))


def wave_flume(mglob=300, nglob=1, kglob=3, dx=0.02, dy=None, depth=0.4, wavemaker="LEF_LIN",
               amp=0.02, per=2.02, theta=0.0, bottom="bar", sponge=None, steps=200, tol=1e-10,
               non_hydro=True, time_order=2, cur=0.0, tramp=0.0) -> Scenario:
    """A wave flume driven by a west wavemaker (BC_X0 = 3) and closed by
    walls elsewhere, at rest at t = 0, after the reference's
    submerged_bar / solitary_wave_breaking examples.  bottom: "flat";
    "bar" (a submerged trapezoid: 1:20 up from x = 0.3 L to depth/4, a crest
    of 0.1 L, 1:10 down); "beach" (a 1:15 plane beach from x = 0.45 L
    that rises above the datum, so the wave runs up and the shoreline
    moves)."""
    dy = dx if dy is None else dy
    lx = mglob * dx
    xc = (np.arange(mglob) + 0.5) * dx
    h = np.full(mglob, depth)
    if bottom == "bar":
        x0 = 0.3 * lx
        top = depth / 4.0
        up = np.clip((xc - x0) / 20.0, 0.0, depth - top)
        x1 = x0 + (depth - top) * 20.0 + 0.1 * lx
        down = np.clip((xc - x1) / 10.0, 0.0, depth - top)
        h = depth - up + down
    elif bottom == "beach":
        h = depth - np.clip((xc - 0.45 * lx) / 15.0, 0.0, None)
    elif bottom != "flat":
        raise ValueError(bottom)
    z = np.zeros((kglob, nglob, mglob))
    return Scenario(
        name=f"flume_{wavemaker.lower()}_{bottom}_{mglob}x{nglob}x{kglob}", mglob=mglob, nglob=nglob,
        kglob=kglob, dx=dx, dy=dy, hcg=np.tile(h, (nglob, 1)), etag=np.zeros((nglob, mglob)),
        ug=z.copy(), vg=z.copy(), wg=z.copy(), min_dep=0.005, dt_ini=0.01, dt_max=0.1,
        froude_cap=10.0, steps=steps, tol=tol, non_hydro=non_hydro, time_order=time_order,
        bc=(3, 1, 1, 1, 1), wavemaker=wavemaker, amp_wave=amp, per_wave=per, dep_wave=depth,
        theta_wave=theta, cur_wave=cur, tramp=tramp, sponge=sponge,
        meta=dict(bottom=bottom))


def internal_wave_basin(mglob=80, nglob=12, kglob=3, dx=0.05, dy=0.05, depth=0.3, nfreq=16,
                        dirs=(0.0,), fmin=0.6, fmax=1.6, hs=0.01, box=(1.0, 1.2, 0.0, 0.6), seed=3,
                        beach=True, steps=100, tol=1e-10, non_hydro=True) -> Scenario:
    """A walled basin at rest driven by the internal irregular wavemaker
    (WAVEMAKER = INT_IRR) over the source box `box`, like the reference's
    Monai Valley run: a seeded spectrum of `nfreq` frequencies in [fmin,
    fmax] Hz (a peaked shape of significant height about `hs`) over the
    directions `dirs` (degrees), seeded random phases; a 1:10 beach in the
    east quarter that rises above the datum (wet-dry) when `beach`."""
    rng = np.random.default_rng(seed)
    xc = (np.arange(mglob) + 0.5) * dx
    h = np.full(mglob, depth)
    if beach:
        x0 = 0.75 * mglob * dx
        h = depth - np.clip((xc - x0) / 10.0, 0.0, None) * 2.0
    freq = np.linspace(fmin, fmax, nfreq)
    fp = 0.5 * (fmin + fmax)
    shape = np.exp(-((freq - fp) / (0.25 * (fmax - fmin))) ** 2) if fmax > fmin else np.ones(nfreq)
    nd = len(dirs)
    # component half-amplitudes (the wavemaker uses Amp = 2*Wave_Spc2d)
    spc = np.outer(shape / np.sqrt((shape ** 2).sum()), np.full(nd, 1.0 / np.sqrt(nd))) * hs / 4.0
    phs = rng.uniform(-np.pi, np.pi, (nfreq, nd))
    z = np.zeros((kglob, nglob, mglob))
    return Scenario(
        name=f"internal_{mglob}x{nglob}x{kglob}", mglob=mglob, nglob=nglob, kglob=kglob, dx=dx, dy=dy,
        hcg=np.tile(h, (nglob, 1)), etag=np.zeros((nglob, mglob)), ug=z.copy(), vg=z.copy(), wg=z.copy(),
        min_dep=0.005, dt_ini=0.01, dt_max=0.1, froude_cap=10.0, steps=steps, tol=tol, non_hydro=non_hydro,
        wavemaker="INT_IRR", dep_wave=depth, source_box=tuple(float(v) for v in box),
        irr_freq=freq, irr_dire=np.asarray(dirs, dtype=float), irr_spc=spc, irr_phs=phs,
        meta=dict(seed=seed))


def steady_flume(mglob=120, nglob=4, kglob=4, dx=0.05, dy=None, depth=0.3, obstacle=0.1, width=0.4,
                 tramp=1.0, zob=1.0e-4, steps=100, tol=1e-10, non_hydro=True, bc_y=1) -> Scenario:
    """Steady current over a submerged obstacle, after the reference's
    Steady_flow_over_submerged_obstacle benchmark: WAVEMAKER = FLUX_LR (a
    logarithmic inflow profile of mean 0.115 m/s ramped over `tramp` s at
    the west, an open east boundary with Orlanski radiation: BC_X0 = BC_Xn
    = 3), the log-law bottom (BC_Z0 = 5, Ibot = 2, roughness `zob`), a
    Gaussian obstacle of height `obstacle` and half width `width` at mid
    flume, the water at rest at t = 0; side walls of type `bc_y`."""
    dy = dx if dy is None else dy
    xc = (np.arange(mglob) + 0.5) * dx
    h = depth - obstacle * np.exp(-((xc - 0.5 * mglob * dx) / width) ** 2)
    z = np.zeros((kglob, nglob, mglob))
    return Scenario(
        name=f"steady_{mglob}x{nglob}x{kglob}", mglob=mglob, nglob=nglob, kglob=kglob, dx=dx, dy=dy,
        hcg=np.tile(h, (nglob, 1)), etag=np.zeros((nglob, mglob)), ug=z.copy(), vg=z.copy(), wg=z.copy(),
        min_dep=0.002, dt_ini=0.01, dt_max=1.0, froude_cap=10.0, steps=steps, tol=tol, non_hydro=non_hydro,
        bc=(3, 3, bc_y, bc_y, 5), ibot=2, zob=zob, wavemaker="FLUX_LR", amp_wave=0.0144, per_wave=10.5,
        dep_wave=depth, tramp=tramp, meta=dict(obstacle=obstacle))


def bumps_beach(n=1024, kglob=10, dx=10.0, seed=2025, nbumps=32, h0=60.0,
                hmin=2.0, hmax=200.0, beach_slope=1.0 / 20.0, beach_width=1200.0,
                amp=2.0, radius=500.0, center=None, steps=2000, tol=1e-10,
                m=None, non_hydro=True) -> Scenario:
    """Seeded Gaussian-bump bathymetry with an east-side 1:20 beach that runs
    above the datum (wet-dry), and a Gaussian surface hump at rest.
    depth = clip(h0 + sum_n a_n exp(-r^2/R_n^2), hmin, hmax), a_n ~ U[-80, 140],
    R_n ~ U[200, 2000] m, centres uniform over the domain; east of the beach
    toe the depth is min(depth, (x_toe_to_shore - x)*slope) reaching
    -beach_width*slope/4 at the east wall."""
    mglob = n if m is None else m
    nglob = n
    rng = np.random.default_rng(seed)
    lx, ly = mglob * dx, nglob * dx
    xc = (np.arange(mglob) + 0.5) * dx
    yc = (np.arange(nglob) + 0.5) * dx
    X, Y = np.meshgrid(xc, yc)
    h = np.full((nglob, mglob), h0)
    for _ in range(nbumps):
        cx, cy = rng.uniform(0, lx), rng.uniform(0, ly)
        r = rng.uniform(200.0, 2000.0)
        a = rng.uniform(-80.0, 140.0)
        h += a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / r ** 2)
    h = np.clip(h, hmin, hmax)
    x_shore = lx - 0.25 * beach_width
    beach = (x_shore - X) * beach_slope
    h = np.where(X > lx - beach_width, np.minimum(h, beach), h)
    if center is None:
        center = (0.45 * lx, 0.5 * ly)
    eta = amp * np.exp(-((X - center[0]) ** 2 + (Y - center[1]) ** 2) / radius ** 2)
    z = np.zeros((kglob, nglob, mglob))
    return Scenario(
        name=f"bumps_{mglob}x{nglob}x{kglob}", mglob=mglob, nglob=nglob, kglob=kglob,
        dx=dx, dy=dx, hcg=h, etag=eta, ug=z, vg=z.copy(), wg=z.copy(), dt_ini=0.1,
        dt_max=60.0, steps=steps, tol=tol, non_hydro=non_hydro,
        meta=dict(seed=seed, nbumps=nbumps, h0=h0, amp=amp, radius=radius))


def ensemble(count: int, seed: int, n_range=(128, 768), k_range=(3, 10), tol=1e-8,
             steps=2000) -> list[Scenario]:
    """`count` scenarios of the bump-and-beach family: grids N x N with N
    log-uniform in n_range, Kglob uniform in k_range, hump amplitude
    U[0.5, 5] m and radius U[200, 2000] m, each bathymetry from its own seed
    drawn from `seed` (dx = 10 m)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        k = int(rng.integers(k_range[0], k_range[1] + 1))
        out.append(bumps_beach(n=n, kglob=k, dx=10.0, seed=int(rng.integers(1 << 30)),
                               amp=float(rng.uniform(0.5, 5.0)), radius=float(rng.uniform(200.0, 2000.0)),
                               steps=steps, tol=tol))
    return out


def beach_runup(mglob=96, nglob=6, kglob=3, dx=0.1, depth=0.4, slope=1.0 / 15.0,
                amp=0.05, x0=2.0, steps=200, tol=1e-10, non_hydro=True) -> Scenario:
    """Solitary wave running up a plane beach that emerges above the datum:
    exercises flooding, drying, the front overrides and Hfx0 mutation."""
    xc = (np.arange(mglob) + 0.5) * dx
    toe = 0.45 * mglob * dx
    h1 = np.where(xc < toe, depth, depth - (xc - toe) * slope)
    kk = math.sqrt(3.0 * amp / (4.0 * depth ** 3))
    eta1 = amp / np.cosh(kk * (xc - x0)) ** 2
    eta1 = np.where(h1 > 0.05, eta1, 0.0)
    c = math.sqrt(9.81 * (depth + amp))
    u1 = np.where(h1 > 0.05, c * eta1 / (depth + eta1), 0.0)
    yv = 1.0 + 0.05 * np.cos(np.pi * (np.arange(nglob) + 0.5) / nglob)
    hcg = np.outer(np.ones(nglob), h1)
    etag = np.outer(yv, eta1)
    ug = np.tile(np.outer(yv, u1), (kglob, 1, 1))
    z = np.zeros((kglob, nglob, mglob))
    return Scenario(
        name=f"beach_{mglob}x{nglob}x{kglob}", mglob=mglob, nglob=nglob, kglob=kglob,
        dx=dx, dy=dx, hcg=hcg, etag=etag, ug=ug, vg=z.copy(), wg=z.copy(), dt_ini=0.005,
        steps=steps, tol=tol, non_hydro=non_hydro, meta=dict(slope=slope, amp=amp))


def viscous_case(base: str, base_kw: dict, visc=1.0e-3, cvs=0.01, chs=0.01, schmidt=1.0,
                 visc_number=0.1666667) -> Scenario:
    """Scenario `base(**base_kw)` run with VISCOUS_FLOW = T and constant
    viscosities (IVTURB = IHTURB = 1)."""
    s = globals()[base](**base_kw)
    return replace(s, name=f"{s.name}_visc", viscous=True, visc=visc, cvs=cvs, chs=chs,
                   schmidt=schmidt, visc_number=visc_number)


def variant(base: str, base_kw: dict, **fields) -> Scenario:
    """Scenario `base(**base_kw)` with the given fields replaced (e.g. the
    bottom boundary `bc`, `viscous` and the viscosities, `ibot`, `cd0`, `zob`)."""
    s = globals()[base](**base_kw)
    if "bc" in fields:
        fields["bc"] = tuple(fields["bc"])
    return replace(s, **fields)


def random_scenario(seed: int, mglob=24, nglob=18, kglob=4, dx=1.0, dy=1.3,
                    steps=5, non_hydro=True, dry=True, tol=1e-12) -> Scenario:
    """Small random bathymetry with islands (dry cells), random surface and
    random velocities: a stress case for every branch of the path."""
    rng = np.random.default_rng(seed)
    h = 1.0 + 0.6 * rng.standard_normal((nglob, mglob))
    if dry:
        h[rng.random((nglob, mglob)) < 0.12] = -0.2
    eta = 0.08 * rng.standard_normal((nglob, mglob))
    ug = 0.2 * rng.standard_normal((kglob, nglob, mglob))
    vg = 0.2 * rng.standard_normal((kglob, nglob, mglob))
    wg = 0.02 * rng.standard_normal((kglob, nglob, mglob))
    return Scenario(
        name=f"random{seed}_{mglob}x{nglob}x{kglob}", mglob=mglob, nglob=nglob, kglob=kglob,
        dx=dx, dy=dy, hcg=h, etag=eta, ug=ug, vg=vg, wg=wg, dt_ini=0.01, steps=steps,
        non_hydro=non_hydro, tol=tol, meta=dict(seed=seed))


# ---------------------------------------------------------------------------
# reference input files (src/mod_util.F:273-415 grammar)
# ---------------------------------------------------------------------------

def _f(x: float) -> str:
    """A float the reference's parser accepts (it needs a '.') and that
    round-trips to the same binary64 value."""
    return f"{float(x):.17e}"


def input_keys(s: Scenario) -> list[tuple[str, str]]:
    """Every key the serial reference reads (src/initialize.F:141-1321), with
    the scenario's values for the hot-path keys; the rest keep every optional
    subsystem off."""
    T, F = "T", "F"
    order = {1: "FIRST", 2: "SECOND", 3: "THIRD"}[s.time_order]
    sw = (0.0, 0.0, 0.0, 0.0) if s.sponge is None else tuple(s.sponge)
    keys = [
        ("TITLE", s.name), ("Mglob", str(s.mglob)), ("Nglob", str(s.nglob)),
        ("Kglob", str(s.kglob)), ("PX", "1"), ("PY", "1"), ("DX", _f(s.dx)), ("DY", _f(s.dy)),
        ("IVGRD", str(s.ivgrd)), ("GRD_R", _f(s.grd_r)), ("DT_INI", _f(s.dt_ini)),
        ("DT_MIN", _f(s.dt_min)), ("DT_MAX", _f(s.dt_max)), ("RESULT_FOLDER", "./results/"),
        ("SIM_STEPS", str(s.steps)), ("TOTAL_TIME", _f(s.total_time)), ("PLOT_START", _f(0.0)),
        ("PLOT_INTV", _f(0.0)), ("SCREEN_INTV", _f(1.0e9)), ("HOTSTART", F),
        ("CFL", _f(s.cfl)), ("VISCOUS_NUMBER", _f(s.visc_number)), ("FROUDE_CAP", _f(s.froude_cap)),
        ("MinDep", _f(s.min_dep)), ("VISCOUS_FLOW", T if s.viscous else F),
        ("IVTURB", str(s.turb) if s.viscous else "10"), ("IHTURB", str(s.turb) if s.viscous else "10"),
        ("PRODTYPE", "3"), ("VISCOSITY", _f(s.visc)), ("Schmidt", _f(s.schmidt)), ("Cvs", _f(s.cvs)),
        ("Chs", _f(s.chs)), ("RNG", T if s.rng else F), ("DEPTH_TYPE", "CELL_CENTER"), ("ANA_BATHY", F),
        ("DepConst", _f(0.3)), ("INITIAL_EUVW", T), ("INITIAL_SALI", F), ("Ibot", str(s.ibot)),
        ("Cd0", _f(s.cd0)), ("Zob", _f(s.zob)), ("Dfric_Min", _f(0.0)), ("Creef", _f(0.0)),
        ("Iws", "1"), ("WindU", _f(s.wind_u)), ("WindV", _f(s.wind_v)), ("slat", _f(s.slat)),
        ("BAROTROPIC", T), ("HIGH_ORDER", "SECOND"), ("TIME_ORDER", order),
        ("CONVECTION", "HLPA"), ("HLLC", F), ("TRAMP", _f(s.tramp)),
        ("NON_HYDRO", T if s.non_hydro else F), ("ISOLVER", "2"), ("ITMAX", str(s.itmax)),
        ("TOL", _f(s.tol)), ("PERIODIC_X", F), ("PERIODIC_Y", F), ("EXTERNAL_FORCING", F),
        ("Pgrad0", _f(0.0)), ("BC_X0", str(s.bc[0])), ("BC_Xn", str(s.bc[1])),
        ("BC_Y0", str(s.bc[2])), ("BC_Yn", str(s.bc[3])), ("BC_Z0", str(s.bc[4])),
        ("BC_Zn", "1"), ("WAVEMAKER", s.wavemaker), ("AMP", _f(s.amp_wave)), ("PER", _f(s.per_wave)),
        ("DEP", _f(s.dep_wave)), ("THETA", _f(s.theta_wave)), ("CUR", _f(s.cur_wave)),
        ("sd_return", _f(0.75)),
        ("Xsource_West", _f(s.source_box[0])), ("Xsource_East", _f(s.source_box[1])),
        ("Ysource_Suth", _f(s.source_box[2])), ("Ysource_Nrth", _f(s.source_box[3])), ("nwave", "1"), ("Component_Amp_Type", "1"),
        ("k_center", _f(1.0)), ("f_center", _f(1.0)), ("x_breaking", _f(0.0)),
        ("t_breaking", _f(0.0)), ("Slope_group", _f(0.0)), ("normalized_delta_f", _f(0.5)),
        ("depth_comp", _f(1.0)), ("Hm0", _f(0.0)), ("Tp", _f(1.0)), ("Freq_Min", _f(0.1)),
        ("Freq_Max", _f(1.0)), ("NumFreq", "1"), ("SPONGE_ON", F if s.sponge is None else T),
        ("Sponge_West_Width", _f(sw[0])), ("Sponge_East_Width", _f(sw[1])),
        ("Sponge_South_Width", _f(sw[2])), ("Sponge_North_Width", _f(sw[3])),
        ("WAVE_AVERAGE_ON", F), ("WAVE_AVERAGE_START", _f(0.0)), ("WAVE_AVERAGE_END", _f(0.0)),
        ("WaveheightID", "1"), ("NSTAT", "0"), ("PLOT_INTV_STAT", _f(1.0e9)),
    ]
    for c in "HEUVWPKDSCBAFTGIZM":
        keys.append((f"OUT_{c}", F))
    # keys of optional subsystems compiled out of this build; present so that
    # any unconditional read of them finds a value
    for k in ["Sed_Type", "Veg_Type", "SlideType", "RHEO_OPT", "SLIDE_FILE", "MultiSlidesFile",
              "COUPLING_FILE"]:
        keys.append((k, "NONE"))
    for k in ["BED_LOAD", "COUPLE_FS", "BED_CHANGE", "SLIDE_SHAPE_BOX", "SMFCut", "SlideStop",
              "Slide", "RHEOLOGY_ON", "NON_HYDRO_SLD", "DISP_CORR_SLD", "REDU_GRAV_SLD",
              "NON_HYDRO_UP"]:
        keys.append((k, F))
    keys.append(("ntyws", "1"))
    for k in ["Sedi_Ws", "D50", "Shields_c", "Af", "Tau_ce", "Tau_cd", "Erate", "Mud_Visc",
              "Tim_Sedi", "MorDt", "Veg_X0", "Veg_Xn", "Veg_Y0", "Veg_Yn", "VegH", "VegDens",
              "StemD", "VegDrag", "Cfk", "Cfe", "VegVM", "EI", "alpha_por", "beta_por", "D50_por",
              "Por_n", "Por_X0", "Por_Xn", "Por_Y0", "Por_Yn", "Por_Z0", "Por_Zn", "SlideT",
              "SlideL", "SlideW", "SlideAngle", "SlopeAngle", "SlideX0", "SlideY0", "SlideUt",
              "SlideA0", "SlideDens", "H_slide", "L_slide", "subaer_center2water", "W_slide",
              "e_slide", "Angle_slide", "Coulomb_phi", "X0_slide", "Y0_slide", "Slope_slide",
              "TermV_slide", "sf_slide", "tf_slide", "Smax_slide", "ac_slide", "ACC_lab",
              "SlideVisc", "SlideLambda", "SlideIniU", "Cf_ul", "PhiInt", "PhiBed",
              "Yield_Stress", "Plastic_Visc"]:
        keys.append((k, _f(0.0)))
    return keys


def _write_rows(path: Path, rows) -> None:
    with open(path, "w") as fh:
        for row in rows:
            vals = [f"{float(v):.17e}" for v in row]
            for a in range(0, len(vals), 16):
                fh.write(" ".join(vals[a:a + 16]) + "\n")


def write_irr2d(s: Scenario, path: Path) -> None:
    """The internal wavemaker's spectrum file as src/initialize.F:835-878
    reads it: "NumFreq NumDir", then one value a line: the frequencies, the
    directions, Wave_Spc2d and Random_Phs (direction fastest)."""
    nf, nd = len(s.irr_freq), len(s.irr_dire)
    with open(path, "w") as fh:
        fh.write(f"{nf} {nd}\n")
        for a in (s.irr_freq, s.irr_dire, np.asarray(s.irr_spc).reshape(-1), np.asarray(s.irr_phs).reshape(-1)):
            for v in a:
                fh.write(f"{float(v):.17e}\n")


def write_reference_inputs(s: Scenario, run_dir: Path) -> None:
    """input.txt, depth.txt, eta0.txt and uvw0.txt in the reference's formats
    (and irr2d.txt for the internal wavemaker)."""
    run_dir = Path(run_dir)
    run_dir.mkdir(parents=True, exist_ok=True)
    (run_dir / "results").mkdir(exist_ok=True)
    with open(run_dir / "input.txt", "w") as fh:
        for k, v in input_keys(s):
            line = f"{k} = {v}"
            assert len(line) <= 80, line
            fh.write(line + "\n")
    _write_rows(run_dir / "depth.txt", s.hcg)
    _write_rows(run_dir / "eta0.txt", s.etag)
    z = np.zeros((s.kglob, s.nglob, s.mglob))
    blocks = []
    for arr in (s.ug, s.vg, s.wg):
        a = z if arr is None else arr
        blocks.extend(a.reshape(s.kglob * s.nglob, s.mglob))
    _write_rows(run_dir / "uvw0.txt", blocks)
    if s.wavemaker == "INT_IRR":
        write_irr2d(s, run_dir / "irr2d.txt")


def read_dump(path: Path) -> dict:
    """Reads a dump_NNNNN.bin written by the oracle's preview hook."""
    raw = Path(path).read_bytes()
    hdr = np.frombuffer(raw, dtype=np.int32, count=6)
    if hdr[0] != 20251202:
        raise ValueError(f"{path}: bad magic {hdr[0]}")
    mloc, nloc, kloc, step, icount = (int(v) for v in hdr[1:6])
    off = 24
    tim = np.frombuffer(raw, dtype=np.float64, count=2, offset=off)
    off += 16
    out = {"step": step, "icount": icount, "TIME": float(tim[0]), "dt": float(tim[1])}

    def take(name, shape, dtype=np.float64):
        nonlocal off
        cnt = int(np.prod(shape))
        a = np.frombuffer(raw, dtype=dtype, count=cnt, offset=off).reshape(shape)
        off += cnt * np.dtype(dtype).itemsize
        out[name] = a.copy()

    take("Eta", (nloc, mloc)); take("D", (nloc, mloc))
    take("Mask", (nloc, mloc), np.int32)
    for nm in ("U", "V", "W", "DU", "DV", "DW"):
        take(nm, (kloc, nloc, mloc))
    take("P", (kloc + 1, nloc, mloc)); take("Omega", (kloc + 1, nloc, mloc))
    take("Hfx0", (nloc, mloc + 1)); take("Hfy0", (nloc + 1, mloc))
    for nm in ("EtaxL", "EtaxR"):
        take(nm, (nloc, mloc + 1))
    for nm in ("EtayL", "EtayR"):
        take(nm, (nloc + 1, mloc))
    if off != len(raw):
        raise ValueError(f"{path}: {len(raw) - off} trailing bytes")
    # viscous runs of the reference's dump hook: the turbulence state written
    # beside dump_NNNNN.bin as turb_NNNNN.bin
    name = Path(path).name
    tpath = Path(path).with_name("turb_" + name[len("dump_"):]) if name.startswith("dump_") else None
    if tpath is not None and tpath.exists():
        t = np.fromfile(tpath, dtype=np.float64)
        names = ("Tke", "Eps", "DTke", "DEps", "CmuVt", "CmuHt")
        if t.size != len(names) * kloc * nloc * mloc:
            raise ValueError(f"{tpath}: unexpected size {t.size}")
        for q, nm in enumerate(names):
            out[nm] = t.reshape(len(names), kloc, nloc, mloc)[q].copy()
    return out
