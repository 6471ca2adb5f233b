"""Scenario ingest: a reference run directory (input.txt, depth.txt, eta0.txt,
uvw0.txt) read with the reference's own rules, into a Scenario the library
runs — the request format of the serving loop.

input.txt (src/mod_util.F:273-334, 341-385): 80-column records; '!' starts
a comment; the first record whose text before '=' is the key (blanks
trimmed) gives the value (text after '=', left-adjusted).  A value is
logical if it is the single character T or F, a string if it holds any
character other than 0-9 + - E e . and blank, else float if it has a '.',
else integer; a getter of the wrong type stops the run.  Keys and the
defaults of the optional ones follow src/initialize.F:141-555.

Arrays (src/initialize.F:1770-1836, 2991-3017): list-directed reads, one
READ per grid row, each starting on a new record and running over as many
records as the row needs.
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np

from .scenario import Scenario

RECORD = 80
COMMENT = "!"
_NUMCHARS = set("0123456789+-Ee. ")


class InputError(ValueError):
    """A run directory the reference would stop on, or one off this path."""


def _records(path) -> list:
    """The file's records; bytes are taken one character each (latin-1), as
    Fortran formatted input takes them, so non-ASCII comment text reads."""
    try:
        return Path(path).read_bytes().decode("latin-1").splitlines()
    except FileNotFoundError:
        raise InputError(f"{path}: file not found") from None


def read_input(path) -> dict:
    """Key -> value string, first occurrence winning, as Extract_Val_String
    finds them."""
    keys: dict = {}
    for raw in _records(path):
        line = raw[:RECORD]
        c = line.find(COMMENT)
        if c >= 0:
            line = line[:c]
        line = line.rstrip()
        eq = line.find("=")
        if eq < 0:
            continue
        name = line[:eq].strip()
        if name and name not in keys:
            keys[name] = line[eq + 1:].strip()
    return keys


def arg_type(val: str) -> str:
    """Check_Arg_Type, src/mod_util.F:341-378."""
    v = val.strip()
    if len(v) == 1 and v in "TF":
        return "logical"
    if any(ch not in _NUMCHARS for ch in v):
        return "string"
    return "float" if "." in v else "integer"


def _fortran_number(v: str) -> str:
    return v.strip().replace("d", "e").replace("D", "E")


class Input:
    """Typed getters over read_input, with the reference's type rules."""

    def __init__(self, path):
        self.path = Path(path)
        self.keys = read_input(path)

    def has(self, name: str) -> bool:
        return name in self.keys

    def _get(self, name: str, want: str) -> str:
        if name not in self.keys:
            raise InputError(f"{self.path}: {name} does not exist in the file")
        v = self.keys[name]
        t = arg_type(v)
        if t != want:
            raise InputError(f"{self.path}: {name} should be {want} but is {t}: {v!r}")
        return v

    def float(self, name: str, default=None) -> float:
        if default is not None and name not in self.keys:
            return float(default)
        return float(_fortran_number(self._get(name, "float")))

    def int(self, name: str, default=None) -> int:
        if default is not None and name not in self.keys:
            return int(default)
        return int(self._get(name, "integer"))

    def bool(self, name: str, default=None) -> bool:
        if default is not None and name not in self.keys:
            return bool(default)
        return self._get(name, "logical") == "T"

    def str(self, name: str, default=None) -> str:
        if name not in self.keys:
            if default is None:
                raise InputError(f"{self.path}: {name} does not exist in the file")
            return default
        return self.keys[name].strip()


def _tokens(line: str):
    """List-directed items of one record: blank or comma separated, r*c
    repeats; a '/' ends the record's input."""
    line = line.split("/")[0]
    for tok in re.split(r"[,\s]+", line.strip()):
        if not tok:
            continue
        if "*" in tok:
            r, c = tok.split("*", 1)
            for _ in range(int(r)):
                yield c
        else:
            yield tok


def read_irr2d(path):
    """The internal irregular wavemaker's spectrum file as src/initialize.F:
    835-878 reads it: NumFreq and NumDir, then one list-directed READ per
    value (frequencies, directions, Wave_Spc2d and Random_Phs, direction
    fastest); more than MaxNumFreq = 500 frequencies or MaxNumDir = 100
    directions stops the reference."""
    lines = _records(path)
    q = 0

    def read(n):
        nonlocal q
        vals = []
        while len(vals) < n:
            if q >= len(lines):
                raise InputError(f"{path}: end of file")
            vals.extend(_tokens(lines[q]))
            q += 1
        return vals[:n]

    nf, nd = (int(v) for v in read(2))
    if not (1 <= nf <= 500 and 1 <= nd <= 100):
        raise InputError(f"{path}: NumFreq = {nf}, NumDir = {nd} (1..500, 1..100)")
    one = lambda: float(_fortran_number(read(1)[0]))  # noqa: E731
    freq = np.array([one() for _ in range(nf)])
    dire = np.array([one() for _ in range(nd)])
    spc = np.array([one() for _ in range(nf * nd)]).reshape(nf, nd)
    phs = np.array([one() for _ in range(nf * nd)]).reshape(nf, nd)
    return freq, dire, spc, phs


def read_rows(path, nrows: int, ncols: int) -> np.ndarray:
    """nrows list-directed READs of ncols values each (a READ starts on a new
    record and takes as many records as it needs)."""
    lines = _records(path)
    out = np.empty((nrows, ncols))
    q = 0
    for r in range(nrows):
        vals = []
        while len(vals) < ncols:
            if q >= len(lines):
                raise InputError(f"{path}: end of file in row {r + 1} of {nrows}")
            vals.extend(float(_fortran_number(t)) for t in _tokens(lines[q]))
            q += 1
        out[r] = vals[:ncols]
    return out


# WAVEMAKER values the GPU path runs (the west wavemakers, with BC_X0 = 3,
# the internal irregular wavemaker, and the steady discharge FLUX_LR, with
# BC_X0 = BC_Xn = 3)
WAVEMAKERS = ("NONE", "LEF_SOL", "LEF_LIN", "LEF_STK", "LEF_CON", "INT_IRR", "FLUX_LR")
WEST_WAVEMAKERS = ("LEF_SOL", "LEF_LIN", "LEF_STK", "LEF_CON")


def _wavemaker(inp: "Input") -> str:
    """The wavemaker as update_wave_bc tells them apart: by the first seven
    characters (src/nhwave.F:5527-5545)."""
    v = inp.str("WAVEMAKER", "NONE").strip().upper()
    return "NONE" if v.startswith("NONE") else v[:7]


# options the GPU path runs; anything else is refused here, as nhw_create and
# the Fortran shim refuse it.  PX/PY (the MPI process grid) are read only by
# a parallel build and do not change the run: one GPU holds the whole domain.
# BAROTROPIC = F is read and runs as T: in a build without SALINITY,
# TEMPERATURE or SEDIMENT the density is Rho0 everywhere (src/initialize.F:
# 2561; eval_dens, src/nhwave.F:12732-12780, sets it only under those
# flags), so baropg_z (:13038-13179) integrates Rho - Rho0 = 0 and adds
# DRhoX = DRhoY = -0.0 to the momentum right-hand side (:643), which leaves
# every sum unchanged.
def _check_supported(inp: Input) -> None:
    bad = []
    for k, want in (("HLLC", False), ("PERIODIC_X", False), ("PERIODIC_Y", False),
                    ("WAVE_AVERAGE_ON", False), ("EXTERNAL_FORCING", False),
                    ("INITIAL_SALI", False), ("HOTSTART", False)):
        if inp.has(k) and inp.bool(k) != want:
            bad.append(f"{k} = {'T' if not want else 'F'}")
    wm = _wavemaker(inp)
    if wm not in WAVEMAKERS:
        bad.append(f"WAVEMAKER = {inp.str('WAVEMAKER')}")
    if inp.has("Iws") and inp.int("Iws") != 1:
        bad.append(f"Iws = {inp.int('Iws')}")
    for k in ("BC_X0", "BC_Xn", "BC_Y0", "BC_Yn"):
        v = inp.int(k)
        inflow = v == 3 and ((k == "BC_X0" and wm in WEST_WAVEMAKERS) or
                             (k in ("BC_X0", "BC_Xn") and wm == "FLUX_LR"))
        friction = v == 5 and inp.bool("VISCOUS_FLOW", False)   # wall friction side walls
        if v not in (1, 2) and not inflow and not friction:
            bad.append(f"{k} = {v}")
    if wm == "FLUX_LR" and (inp.int("BC_X0"), inp.int("BC_Xn")) != (3, 3):
        bad.append("WAVEMAKER = FLUX_LR without BC_X0 = BC_Xn = 3")
    # the viscous branch: constant viscosities (IVTURB = IHTURB = 1) or the
    # 3-D k-epsilon model (IVTURB = 10 with IHTURB >= 10, Kglob >= 3)
    if inp.bool("VISCOUS_FLOW", False):
        iv, ih = inp.int("IVTURB"), inp.int("IHTURB")
        if not ((iv == 1 and ih == 1) or (iv == 10 and ih >= 10)):
            bad.append(f"IVTURB = {iv}, IHTURB = {ih} with VISCOUS_FLOW = T")
        if inp.int("Kglob") > 32:
            bad.append(f"Kglob = {inp.int('Kglob')} with VISCOUS_FLOW = T")
        if iv == 10 and inp.int("Kglob") < 3:
            bad.append(f"Kglob = {inp.int('Kglob')} with the k-epsilon model")
    if inp.int("BC_Z0") not in (1, 2, 5):
        bad.append(f"BC_Z0 = {inp.int('BC_Z0')}")
    if bad:
        raise InputError("options off the GPU path: " + ", ".join(bad))


def scenario_from_run_dir(run_dir, name: str | None = None) -> Scenario:
    """The Scenario of a reference run directory."""
    d = Path(run_dir)
    inp = Input(d / "input.txt")
    _check_supported(inp)
    M, N, K = inp.int("Mglob"), inp.int("Nglob"), inp.int("Kglob")
    # bathymetry, src/initialize.F:1770-1836
    dtype = inp.str("DEPTH_TYPE").strip().upper()
    ana = inp.bool("ANA_BATHY")
    if dtype == "CELL_CENTER":
        hcg = np.full((N, M), 0.5) if ana else read_rows(d / "depth.txt", N, M)
    elif dtype == "CELL_GRID":
        hg = np.full((N + 1, M + 1), 30.0) if ana else read_rows(d / "depth.txt", N + 1, M + 1)
        mp = (hg >= -1000.0).astype(float)  # Maskp: permanent dry points
        num = hg[:-1, :-1] * mp[:-1, :-1] + hg[:-1, 1:] * mp[:-1, 1:] + hg[1:, :-1] * mp[1:, :-1] + \
            hg[1:, 1:] * mp[1:, 1:]
        den = mp[:-1, :-1] + mp[:-1, 1:] + mp[1:, :-1] + mp[1:, 1:] + float(np.float32(1.0e-16))
        hcg = num / den
    else:
        raise InputError(f"DEPTH_TYPE = {dtype} is not CELL_CENTER or CELL_GRID")
    # initial surface and velocities, src/initialize.F:2991-3017
    if inp.bool("INITIAL_EUVW"):
        etag = read_rows(d / "eta0.txt", N, M)
        uvw = read_rows(d / "uvw0.txt", 3 * K * N, M).reshape(3, K, N, M)
        ug, vg, wg = uvw[0], uvw[1], uvw[2]
    else:
        etag = np.zeros((N, M))
        ug = vg = wg = None
    # boundary forcing (src/initialize.F:537, 601-613, 946-957): the wave
    # keys are read only with a wavemaker, the sponge widths only with
    # SPONGE_ON
    wm = _wavemaker(inp)
    wave = {}
    if wm != "NONE":
        wave = dict(amp_wave=inp.float("AMP"), per_wave=inp.float("PER"), dep_wave=inp.float("DEP"),
                    theta_wave=inp.float("THETA"), cur_wave=inp.float("CUR"))
    if wm.startswith("INT"):  # src/initialize.F:629-634
        wave["source_box"] = tuple(inp.float(k) for k in ("Xsource_West", "Xsource_East", "Ysource_Suth",
                                                           "Ysource_Nrth"))
    if wm == "INT_IRR":
        f, di, sp, ph = read_irr2d(d / "irr2d.txt")
        wave.update(irr_freq=f, irr_dire=di, irr_spc=sp, irr_phs=ph)
    sponge = None
    if inp.bool("SPONGE_ON", False):
        sponge = tuple(inp.float(f"Sponge_{s}_Width") for s in ("West", "East", "South", "North"))
    # viscous keys, src/initialize.F:357, 385-393
    visc_kw = {}
    if inp.bool("VISCOUS_FLOW", False):
        visc_kw = dict(viscous=True, cvs=inp.float("Cvs"), chs=inp.float("Chs"), schmidt=inp.float("Schmidt"),
                       visc_number=inp.float("VISCOUS_NUMBER"), rng=inp.bool("RNG"),
                       turb=10 if inp.int("IVTURB") == 10 else 1)
    # bottom friction, src/initialize.F:466-468
    if inp.int("BC_Z0") == 5:
        visc_kw.update(ibot=inp.int("Ibot"), cd0=inp.float("Cd0"), zob=inp.float("Zob"))
    order = inp.str("TIME_ORDER", "THIRD").upper()
    time_order = 3 if order.startswith("THI") else 2 if order.startswith("SEC") else 1
    return Scenario(
        name=name or inp.str("TITLE", d.name), mglob=M, nglob=N, kglob=K, dx=inp.float("DX"),
        dy=inp.float("DY"), hcg=hcg, etag=etag, ug=ug, vg=vg, wg=wg, cfl=inp.float("CFL"),
        min_dep=inp.float("MinDep"), dt_ini=inp.float("DT_INI"), dt_min=inp.float("DT_MIN"),
        dt_max=inp.float("DT_MAX"), froude_cap=inp.float("FROUDE_CAP"), visc=inp.float("VISCOSITY"),
        slat=inp.float("slat"), wind_u=inp.float("WindU", 0.0), wind_v=inp.float("WindV", 0.0),
        time_order=time_order, non_hydro=inp.bool("NON_HYDRO"),
        bc=(inp.int("BC_X0"), inp.int("BC_Xn"), inp.int("BC_Y0"), inp.int("BC_Yn"), inp.int("BC_Z0")),
        itmax=inp.int("ITMAX"), tol=inp.float("TOL"), ivgrd=inp.int("IVGRD"), grd_r=inp.float("GRD_R"),
        steps=inp.int("SIM_STEPS"), total_time=inp.float("TOTAL_TIME"), wavemaker=wm, sponge=sponge,
        tramp=inp.float("TRAMP", 0.0), meta={"run_dir": str(d)}, **wave, **visc_kw)
