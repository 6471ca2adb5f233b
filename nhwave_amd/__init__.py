"""nhwave_amd: MI355X (gfx950) implementation of the NHWAVE time-stepping hot
path (fluxes, eval_duvw, update_mask, poisson_solver and the per-step glue)
behind the C ABI declared in include/nhwave_amd.h."""
from .scenario import Scenario  # noqa: F401

__all__ = ["Scenario", "capi", "model", "scenario"]
