"""Many independent runs of one shape advancing together on one GPU.

``Ensemble`` keeps every run's spectral state resident in HBM and advances all
active runs with one launch sequence per step, each run with its own CFL dt
(the production policy of cfl_policy.py:64-74 and api.py:169-190 applied per
run).  Runs are retired when their own ``t >= tstop``.
"""
import numpy as np
import torch

from . import _numerics as NUM
from .api import model_spec
from .core.cfl_policy import clamp_safety
from .engine import ANTISYM_TOL, Engine, KIND_CONST, KIND_DEPTH, KIND_COUPLED, antisym_part

_KIND = {"const": KIND_CONST, "eta_z": KIND_DEPTH, "eta_T": KIND_COUPLED,
         "arrhenius": KIND_COUPLED}


def run_params(cfg):
    """Per-run physics and Stokes route of a ConvectionConfig as the engine's
    parameter dict."""
    return dict(Ra=float(cfg.Ra), deta_T=float(cfg.deta_T), Ts=float(cfg.arrhenius_Ts),
                arrhenius_V=float(cfg.arrhenius_V), reference_anchor=str(cfg.reference_anchor),
                eta_min=float(cfg.eta_min), eta_max=float(cfg.eta_max),
                eta_smooth=float(cfg.eta_smooth),
                internal_heating=float(cfg.internal_heating),
                dissipation_number=float(cfg.dissipation_number),
                precond=str(cfg.precond).strip().lower(), lag=int(cfg.lag))


class Ensemble:
    """B runs sharing one (kind, Nx, Nz, Lx); ``cfgs`` are ConvectionConfigs."""

    def __init__(self, cfgs, thetas, tstops, safety=None, dtmax=None, prec=0,
                 **engine_kw):
        specs = [model_spec(c) for c in cfgs]
        kinds = {_KIND[s["model"]] for s in specs}
        shapes = {(int(c.Nx), int(c.Nz), float(c.Lx)) for c in cfgs}
        if len(kinds) != 1 or len(shapes) != 1:
            raise ValueError("an Ensemble holds one model kind and one grid shape")
        (Nx, Nz, Lx), = shapes
        self.kind = kinds.pop()
        self.cfgs = list(cfgs)
        self.eng = Engine(self.kind, Nx, Nz, Lx, [run_params(c) for c in cfgs],
                          prec=prec, **engine_kw)
        self.B = len(cfgs)
        if self.kind != KIND_CONST and any(antisym_part(t) > ANTISYM_TOL for t in thetas):
            raise ValueError("a %s ensemble needs mirror-symmetric initial states (real "
                             "x-mode coefficients)" % self.kind)
        self.eng.set_theta(np.stack([np.asarray(t) for t in thetas]))
        self.tstop = np.broadcast_to(np.asarray(tstops, dtype=np.float64), (self.B,)).copy()
        self.t = np.zeros(self.B)
        self.steps = np.zeros(self.B, dtype=np.int64)
        self.safety = clamp_safety(NUM.CFL_SAFETY if safety is None else safety)
        self.dtmax = NUM.DTMAX if dtmax is None else float(dtmax)
        self.done = self.t >= self.tstop

    def active_mask(self):
        return torch.as_tensor((~self.done).astype(np.int32), device=self.eng.dev)

    def step(self):
        """One step of every unfinished run; -> number of run-steps taken."""
        live = ~self.done
        if not np.any(live):
            return 0
        act = None if np.all(live) else self.active_mask()
        dt = self.eng.cfl(self.safety, self.dtmax, NUM.DTMIN, act)
        dt = np.minimum(self.dtmax, dt)
        self.eng.step(dt, act)
        self.t[live] += dt[live]
        self.steps[live] += 1
        self.done = self.t >= self.tstop
        return int(live.sum())

    def march(self, max_steps=None):
        n = 0
        while not np.all(self.done):
            n += self.step()
            if max_steps is not None and n >= max_steps:
                break
        return n

    def states(self):
        return self.eng.get_theta()

    def budgets(self):
        """Energy budget of every run's last step, on the device (see
        ``Engine.energy_budget``); -> dict of (B,) arrays."""
        return self.eng.energy_budget()

    def diagnostics(self):
        return self.eng.diagnostics()
