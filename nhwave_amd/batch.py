"""Batched (ensemble) form: many scenarios of uneven grid sizes advanced
together on one GPU, one kernel launch per step of the time step for the
whole ensemble (nhw_batch_* in include/nhwave_amd.h)."""
from __future__ import annotations

import ctypes as C

from . import capi
from .capi import check
from .model import Model, make_config
from .scenario import Scenario


class Batch:
    def __init__(self, scenarios: list[Scenario], precision: int = capi.NHW_FP64, autotune: bool = True,
                 graphs: bool = True):
        self.lib = capi.load()
        self.scenarios = list(scenarios)
        n = len(self.scenarios)
        cfgs = (capi.NhwConfig * n)(*[make_config(s, precision) for s in self.scenarios])
        self.h = C.c_void_p()
        check(self.lib.nhw_batch_create(cfgs, n, C.byref(self.h)), "nhw_batch_create")
        self.members: list[Model] = []
        for i, s in enumerate(self.scenarios):
            ctx = C.c_void_p()
            check(self.lib.nhw_batch_member(self.h, i, C.byref(ctx)), "nhw_batch_member")
            self.members.append(Model(s, handle=ctx))
        check(self.lib.nhw_batch_prepare(self.h, int(autotune)), "nhw_batch_prepare")
        check(self.lib.nhw_batch_graphs(self.h, int(graphs)), "nhw_batch_graphs")

    def step(self, n: int = 1) -> None:
        check(self.lib.nhw_batch_step(self.h, int(n)), "nhw_batch_step")

    def tuning(self) -> list[dict]:
        out = []
        for kind in range(7):
            name = C.c_char_p()
            thr = C.c_int32()
            ms = (C.c_double * 3)()
            check(self.lib.nhw_batch_tuning(self.h, kind, C.byref(name), C.byref(thr), ms), "nhw_batch_tuning")
            out.append(dict(kernel=name.value.decode(), threads=thr.value,
                            ms={64: ms[0], 128: ms[1], 256: ms[2]}))
        return out

    def cells(self) -> int:
        return sum(s.mglob * s.nglob * s.kglob for s in self.scenarios)

    def close(self) -> None:
        for m in self.members:
            m.ctx = C.c_void_p()
        if self.h:
            self.lib.nhw_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
