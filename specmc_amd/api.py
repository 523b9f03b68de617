"""Public API: config -> device model -> march (the surface of specmc/api.py).

``build_model(cfg)`` returns the device-backed class of ``cfg.viscosity_model``
and checks it, ``march`` / ``march_adaptive`` run the reference's loops
unchanged (api.py:147-190), and the introspection helpers keep their
meaning.  Both Stokes routes of the coupled models run on the device:
``precond='bfbt'`` and ``precond='legacy'`` (the lagged-LU route, the
ConvectionConfig default).
"""
from __future__ import annotations


import numpy as np

from . import _numerics as NUM
from .config import (ConvectionConfig, normalize_compressibility,
                     normalize_viscosity_model)
from .core.cfl_policy import cfl_dt as _policy_dt

__all__ = ["MODEL_CLASSES", "numerics", "model_spec", "model_kwargs",
           "build_model", "describe", "production_dt", "load_checkpoint",
           "march", "march_adaptive"]

_BUILDERS = {"const": "RBC", "eta_z": "RBCImplicitTheta", "eta_T": "RBCCoupled",
             "arrhenius": "RBCCoupled"}


def _classes():
    from . import models
    return {k: getattr(models, v) for k, v in _BUILDERS.items()}


class _LazyClasses(dict):
    def __missing__(self, key):
        self.update(_classes())
        return dict.__getitem__(self, key)

    def items(self):
        self.update(_classes())
        return dict.items(self)


MODEL_CLASSES = _LazyClasses()


def numerics():
    return NUM.as_dict()


def _kwargs(cfg, vm):
    base = dict(Ra=float(cfg.Ra), Nx=int(cfg.Nx), Nz=int(cfg.Nz), Lx=float(cfg.Lx),
                internal_heating=float(cfg.internal_heating),
                dissipation_number=float(cfg.dissipation_number))
    if vm == "const":
        return base
    base.update(deta_T=float(cfg.deta_T), eta_min=float(cfg.eta_min),
                eta_max=float(cfg.eta_max), eta_smooth=float(cfg.eta_smooth),
                symmetrize=bool(cfg.symmetrize))
    if vm == "eta_z":
        return base
    base.update(arrhenius_Ts=float(cfg.arrhenius_Ts), arrhenius_V=float(cfg.arrhenius_V),
                reference_anchor=str(cfg.reference_anchor),
                precond=str(cfg.precond).strip().lower(), lag=int(cfg.lag),
                K=int(cfg.K), max_K=int(cfg.max_K),
                defect_tol=(None if float(cfg.defect_tol) <= 0.0 else float(cfg.defect_tol)))
    return base


def model_spec(cfg: ConvectionConfig):
    """-> the construction recipe for ``cfg`` (api.py:73-94)."""
    if not isinstance(cfg, ConvectionConfig):
        raise TypeError("expected ConvectionConfig, got %r" % (type(cfg),))
    cfg.validate()
    vm = normalize_viscosity_model(cfg.viscosity_model)
    comp = normalize_compressibility(cfg.compressibility)
    if vm not in _BUILDERS:
        raise NotImplementedError("viscosity_model=%r is not implemented" % (vm,))
    return dict(model=vm, compressibility=comp, builder=_BUILDERS[vm],
                cls=MODEL_CLASSES[vm], kwargs=_kwargs(cfg, vm))


def model_kwargs(cfg):
    return dict(model_spec(cfg)["kwargs"])


def build_model(cfg: ConvectionConfig, precision="fp64"):
    """Build the device model for ``cfg`` and check its class (api.py:102-110).

    ``precision`` is an extension of this package: "fp64" (default) runs every
    kernel in FP64; "mixed" factors the dense rescue systems in FP32 under the
    FP64 defect correction (the engine's ``PRECISIONS``)."""
    spec = model_spec(cfg)
    kw = dict(spec["kwargs"])
    if precision != "fp64":
        kw["precision"] = precision
    model = spec["cls"](**kw)
    if type(model) is not spec["cls"]:
        raise AssertionError("internal mapping error: viscosity_model=%r built %s, "
                             "expected %s" % (spec["model"], type(model).__name__,
                                              spec["cls"].__name__))
    return model


def describe(cfg):
    spec = model_spec(cfg)
    return dict(config=cfg.canonical(), model=spec["model"],
                model_class=spec["cls"].__name__, builder=spec["builder"],
                ctor_kwargs=dict(spec["kwargs"]), numerics=numerics())


def production_dt(model, that):
    return float(_policy_dt(model, that, safety=NUM.CFL_SAFETY, dtmax=NUM.DTMAX))


def load_checkpoint(path):
    """-> ``(theta_coefficients, meta)`` from a state file (api.py's
    load_checkpoint delegating to io/checkpoint.py:82-103): FileNotFoundError
    for a missing file, ValueError for a file without ``that``; meta holds
    path, keys, t, n, fac (and whether the file carries a device snapshot,
    see ``checkpoint.load_run``)."""
    from . import checkpoint
    return checkpoint.load(path)


def march(model, that, dt, nsteps, keep_states=False, on_step=None):
    """Fixed-dt march (api.py:147-166)."""
    dt = float(dt)
    nsteps = int(nsteps)
    s = np.array(that, copy=True)
    states = np.empty((nsteps,) + s.shape, dtype=complex) if keep_states else None
    for n in range(nsteps):
        s = model.step(s, dt)
        if states is not None:
            states[n] = s
        if on_step is not None:
            on_step(n + 1, (n + 1) * dt, s)
    info = dict(steps=nsteps, t=nsteps * dt, dt=dt)
    if states is not None:
        info["states"] = states
    return s, info


def march_adaptive(model, that, tstop, safety=None, dtmax=None, on_step=None):
    """CFL-limited march to ``tstop`` (api.py:169-190)."""
    safety = NUM.CFL_SAFETY if safety is None else float(safety)
    dtmax = NUM.DTMAX if dtmax is None else float(dtmax)
    s = np.array(that, copy=True)
    t = 0.0
    n = 0
    dt = float("nan")
    while t < float(tstop):
        dt = float(_policy_dt(model, s, safety=safety, dtmax=dtmax))
        s = model.step(s, dt)
        t += dt
        n += 1
        if on_step is not None:
            on_step(n, t, s)
    return s, dict(steps=n, t=t, dt=dt, safety=safety, dtmax=dtmax)
