"""The CFL-safety policy (the contract of specmc/core/cfl_policy.py:54-74)."""
import numpy as np

__all__ = ["CFL_SAFETY_DEFAULT", "CFL_SAFETY_MAX", "clamp_safety", "cfl_dt"]

CFL_SAFETY_DEFAULT = 40.0
CFL_SAFETY_MAX = 80.0


def clamp_safety(safety, lo=0.0):
    """Clamp a requested safety factor into [lo, CFL_SAFETY_MAX]."""
    return float(np.clip(float(safety), float(lo), CFL_SAFETY_MAX))


def cfl_dt(model, that, safety=CFL_SAFETY_DEFAULT, dtmax=4.0e-5, dtmin=1e-10,
           clamp=True):
    """dt = min(dtmax, model.cfl_dt(that, safety=clamp(safety), ...))."""
    fac = clamp_safety(safety) if clamp else float(safety)
    return float(min(dtmax, model.cfl_dt(that, safety=fac, dtmax=dtmax,
                                         dtmin=dtmin)))
