"""Host-side driver of one scenario on the GPU, mirroring the reference's
program flow (src/nhwave.F:53-333): set-up on the host (generate_grid,
read_bathymetry, the loaded part of `initial`), then the device-resident time
loop through the C ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import NhwConfig, NhwFields, NhwStats, NhwStatValues, check
from .scenario import Scenario, bathymetry_fields, initial_fields, vertical_grid

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)


def field_shape(s: Scenario, name: str) -> tuple:
    m, n, k = s.mloc, s.nloc, s.kloc
    if name in ("Hfx", "Hfx0", "EtaxL", "EtaxR"):
        return (n, m + 1)
    if name in ("Hfy", "Hfy0", "EtayL", "EtayR"):
        return (n + 1, m)
    if name in ("U", "V", "W", "DU", "DV", "DW", "DU0", "DV0", "DW0", "CmuVt", "Tke", "Eps", "DTke", "DEps"):
        return (k, n, m)
    if name in ("P", "Omega", "Fz", "Gz", "Hz"):
        return (k + 1, n, m)
    if name in ("Ex", "Fx", "Gx", "Hx"):
        return (k, n, m + 1)
    if name in ("Ey", "Fy", "Gy", "Hy"):
        return (k, n + 1, m)
    return (n, m)


INT_FIELDS = ("Mask", "Mask_Struct", "Mask9")


def make_config(s: Scenario, precision: int = capi.NHW_FP64,
                solver: int = capi.NHW_SOLVER_MG) -> NhwConfig:
    return NhwConfig(
        mglob=s.mglob, nglob=s.nglob, kglob=s.kglob, dx=s.dx, dy=s.dy, cfl=s.cfl,
        min_dep=s.min_dep, dt_ini=s.dt_ini, dt_min=s.dt_min, dt_max=s.dt_max,
        froude_cap=s.froude_cap, froude_cap_on=int(s.froude_cap_on), visc=s.visc, fcor=s.fcor,
        wind_u=s.wind_u, wind_v=s.wind_v, time_order=s.time_order, non_hydro=int(s.non_hydro),
        hllc=0, bc_x0=s.bc[0], bc_xn=s.bc[1], bc_y0=s.bc[2], bc_yn=s.bc[3], bc_z0=s.bc[4],
        itmax=s.itmax, tol=s.tol, precision=precision, solver=solver,
        wavemaker=capi.WAVEMAKERS[s.wavemaker], amp_wave=s.amp_wave, per_wave=s.per_wave,
        dep_wave=s.dep_wave, theta_wave=s.theta_wave, cur_wave=s.cur_wave, tramp=s.tramp,
        sponge_on=int(s.sponge is not None),
        **dict(zip(("sponge_w", "sponge_e", "sponge_s", "sponge_n"), s.sponge or (0.0,) * 4)),
        viscous=int(s.viscous), rng=int(s.rng), cvs=s.cvs, chs=s.chs, schmidt=s.schmidt,
        visc_number=s.visc_number, ibot=s.ibot, cd0=s.cd0, zob=s.zob, turb=s.turb)


class Model:
    """One scenario resident on the current GPU."""

    def __init__(self, s: Scenario, precision: int = capi.NHW_FP64,
                 solver: int = capi.NHW_SOLVER_MG, handle: C.c_void_p | None = None):
        """Creates a context for `s` (or, with `handle`, fills an existing
        one that something else owns, such as a batch member) and runs the
        host set-up and the device tail of `initial`."""
        self.s = s
        self.lib = capi.load()
        self.owned = handle is None
        self.ctx = C.c_void_p()
        if handle is None:
            cfg = make_config(s, precision, solver)
            check(self.lib.nhw_create(C.byref(cfg), C.byref(self.ctx)), "nhw_create")
        else:
            self.ctx = handle
        dsig, sig, sigc = vertical_grid(s.kglob, s.ivgrd, s.grd_r)
        self.dsig, self.sig, self.sigc = dsig, sig, sigc
        check(self.lib.nhw_set_grid(self.ctx, dsig.ctypes.data_as(_DP), sig.ctypes.data_as(_DP),
                                    sigc.ctypes.data_as(_DP)), "nhw_set_grid")
        host = bathymetry_fields(s)
        host.update(initial_fields(s))
        self.upload(host)
        check(self.lib.nhw_initial(self.ctx), "nhw_initial")
        if s.wavemaker == "INT_IRR":
            arr = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
                   for a in (s.irr_freq, s.irr_dire, s.irr_spc, s.irr_phs)]
            check(self.lib.nhw_set_internal_wavemaker(self.ctx, *(float(v) for v in s.source_box), len(arr[0]),
                                                      len(arr[1]), *(a.ctypes.data_as(_DP) for a in arr)),
                  "nhw_set_internal_wavemaker")

    # -- transfers ----------------------------------------------------------
    def _fields_struct(self, arrays: dict) -> NhwFields:
        f = NhwFields()
        for name, a in arrays.items():
            want = np.int32 if name in INT_FIELDS else np.float64
            shp = field_shape(self.s, name)
            if a.shape != shp or a.dtype != want or not a.flags.c_contiguous:
                raise ValueError(f"{name}: need C-contiguous {want.__name__}{shp}, got {a.dtype}{a.shape}")
            setattr(f, name, a.ctypes.data_as(_IP if name in INT_FIELDS else _DP))
        return f

    def upload(self, arrays: dict) -> None:
        arrs = {k: np.ascontiguousarray(v, dtype=np.int32 if k in INT_FIELDS else np.float64)
                for k, v in arrays.items()}
        f = self._fields_struct(arrs)
        check(self.lib.nhw_upload(self.ctx, C.byref(f)), "nhw_upload")

    def download(self, *names) -> dict:
        out = {n: np.empty(field_shape(self.s, n), dtype=np.int32 if n in INT_FIELDS else np.float64)
               for n in names}
        f = self._fields_struct(out)
        check(self.lib.nhw_download(self.ctx, C.byref(f)), "nhw_download")
        return out

    def field(self, name: str) -> np.ndarray:
        return self.download(name)[name]

    # -- output and restart ---------------------------------------------------
    def statistics(self) -> dict:
        """The reference's `statistics` quantities, reduced on the device."""
        v = NhwStatValues()
        check(self.lib.nhw_statistics(self.ctx, C.byref(v)), "nhw_statistics")
        return {n: getattr(v, n) for n, _ in NhwStatValues._fields_}

    def snapshot_begin(self, *names) -> dict:
        """Queues a snapshot of `names` as they are now; the returned arrays
        hold it after snapshot_wait() (stepping may continue meanwhile)."""
        out = {n: np.empty(field_shape(self.s, n), dtype=np.int32 if n in INT_FIELDS else np.float64)
               for n in names}
        self._snap = (out, self._fields_struct(out))
        check(self.lib.nhw_snapshot_begin(self.ctx, C.byref(self._snap[1])), "nhw_snapshot_begin")
        return out

    def snapshot_wait(self) -> None:
        check(self.lib.nhw_snapshot_wait(self.ctx), "nhw_snapshot_wait")

    def checkpoint_save(self, path) -> None:
        check(self.lib.nhw_checkpoint_save(self.ctx, str(path).encode()), "nhw_checkpoint_save")

    def checkpoint_load(self, path) -> None:
        check(self.lib.nhw_checkpoint_load(self.ctx, str(path).encode()), "nhw_checkpoint_load")

    # -- time loop ------------------------------------------------------------
    def step(self, n: int = 1) -> None:
        check(self.lib.nhw_step(self.ctx, int(n)), "nhw_step")

    def run(self, total_time: float | None = None, sim_steps: int | None = None) -> int:
        """The reference's loop (src/nhwave.F:119): steps while TIME <
        TOTAL_TIME and RUN_STEP < SIM_STEPS, the scenario's by default;
        returns the number of steps taken."""
        s = self.s
        tt = s.total_time if total_time is None else total_time
        ns = s.steps if sim_steps is None else sim_steps
        taken = C.c_int32()
        check(self.lib.nhw_run(self.ctx, float(tt), int(ns), C.byref(taken)), "nhw_run")
        return taken.value

    def poisson_system(self):
        """generate_coef_rhs on the current state: (coef (15, neq), rhs (neq,)),
        rows i fastest, then j, then k over the interior."""
        n = self.s.mglob * self.s.nglob * self.s.kglob
        coef, rhs = np.empty((15, n)), np.empty(n)
        check(self.lib.nhw_generate_coef_rhs(self.ctx, coef.ctypes.data_as(_DP), rhs.ctypes.data_as(_DP)),
              "nhw_generate_coef_rhs")
        return coef, rhs

    def poisson_apply(self, x: np.ndarray) -> np.ndarray:
        """A x with the multigrid solver's matrix-free operator."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        check(self.lib.nhw_poisson_apply(self.ctx, x.ctypes.data_as(_DP), y.ctypes.data_as(_DP)),
              "nhw_poisson_apply")
        return y

    def call(self, name: str, *args) -> None:
        check(getattr(self.lib, name)(self.ctx, *args), name)

    def time(self):
        t, dt, st = C.c_double(), C.c_double(), C.c_int32()
        check(self.lib.nhw_get_time(self.ctx, C.byref(t), C.byref(dt), C.byref(st)), "nhw_get_time")
        return t.value, dt.value, st.value

    def stats(self) -> NhwStats:
        st = NhwStats()
        check(self.lib.nhw_solver_stats(self.ctx, C.byref(st)), "nhw_solver_stats")
        return st

    def sync(self) -> None:
        check(self.lib.nhw_sync(self.ctx), "nhw_sync")

    def close(self) -> None:
        if self.ctx and self.owned:
            self.lib.nhw_destroy(self.ctx)
        self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
