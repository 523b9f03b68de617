"""The reference hydrostatic state and the calibration of the pressure-dependent
(pV) Arrhenius viscosity law (reference_state.py:142-357, arrhenius.py:283-325).

The law is

    ln eta(T, z) = (E + V p0_norm(z)) / (T + Ts) + C,   C = -E / Ts,
    p0_norm(z)   = 1 - z,

so eta = 1 at the cold top (T = 0, z = 1).  It is evaluated in the
cancellation-free form ``-E T / (Ts (T + Ts)) + V p0_norm / (T + Ts)`` by the
device's viscosity-law kernel.  ``calibrate_pv`` fixes (E, V, Ts, C) from
eight dimensional constants and one anchor: ``"Ea"`` takes the activation
energy from the constants (the contrast is an output), ``"D"`` takes the
contrast ``deta_T`` as the input (the implied activation energy is an
output).  The record carries the same keys as the reference's.
"""
from __future__ import annotations

import math
from dataclasses import asdict, dataclass

import numpy as np

__all__ = ["ReferenceConstants", "REFERENCE_DEFAULTS", "PV_ANCHORS", "p0_norm", "V_tilde",
           "Ts_tilde", "E_tilde", "infer_E_a", "infer_V_a", "calibrate_pv"]


@dataclass(frozen=True)
class ReferenceConstants:
    """The dimensional constants of the calibration (reference_state.py:142-174):
    mantle density, gravity, layer depth, temperature drop, surface
    temperature, activation volume and energy, gas constant."""
    rho0: float = 3300.0
    g: float = 9.81
    d: float = 2.89e6
    DeltaT: float = 3000.0
    T_surf: float = 273.0
    V_a: float = 6e-6
    E_a: float = 5.3e5
    R: float = 8.314462618

    def as_dict(self):
        return asdict(self)


REFERENCE_DEFAULTS = ReferenceConstants()
PV_ANCHORS = ("Ea", "D")


def _checked(c):
    if not isinstance(c, ReferenceConstants):
        raise TypeError("expected ReferenceConstants, got %r" % (type(c),))
    for name in ("rho0", "g", "d", "DeltaT", "T_surf", "E_a", "R"):
        v = float(getattr(c, name))
        if not (math.isfinite(v) and v > 0.0):
            raise ValueError("ReferenceConstants.%s must be positive and finite (got %r)"
                             % (name, getattr(c, name)))
    va = float(c.V_a)
    if not (math.isfinite(va) and va >= 0.0):
        raise ValueError("ReferenceConstants.V_a must be finite and >= 0 (got %r)" % (c.V_a,))
    return c


def p0_norm(z):
    """Dimensionless reference hydrostatic pressure 1 - z (reference_state.py:208-216)."""
    return 1.0 - np.asarray(z, dtype=float)


def V_tilde(constants=None):
    """rho0 g d V_a / (R DeltaT) (reference_state.py:230-238)."""
    c = _checked(constants or REFERENCE_DEFAULTS)
    return c.rho0 * c.g * c.d * c.V_a / (c.R * c.DeltaT)


def Ts_tilde(constants=None):
    """T_surf / DeltaT (reference_state.py:241-248)."""
    c = _checked(constants or REFERENCE_DEFAULTS)
    return c.T_surf / c.DeltaT


def E_tilde(constants=None):
    """E_a / (R DeltaT) (reference_state.py:251-254)."""
    c = _checked(constants or REFERENCE_DEFAULTS)
    return c.E_a / (c.R * c.DeltaT)


def infer_V_a(V, constants=None):
    c = _checked(constants or REFERENCE_DEFAULTS)
    return float(V) * c.R * c.DeltaT / (c.rho0 * c.g * c.d)


def infer_E_a(E, constants=None):
    c = _checked(constants or REFERENCE_DEFAULTS)
    return float(E) * c.R * c.DeltaT


def calibrate_pv(constants=None, anchor="Ea", deta_T=None, V=None):
    """-> the calibration record of the pV law for one anchor
    (reference_state.py:277-357).  ``V`` overrides the physical V-tilde.
    ValueError for an unknown anchor, a missing or <= 1 ``deta_T`` under
    anchor "D", or a negative / non-finite ``V``."""
    c = _checked(constants or REFERENCE_DEFAULTS)
    a = str(anchor)
    if a not in PV_ANCHORS:
        raise ValueError("anchor must be one of %r (got %r)" % (PV_ANCHORS, anchor))
    V = V_tilde(c) if V is None else float(V)
    if not (math.isfinite(V) and V >= 0.0):
        raise ValueError("V-tilde must be finite and >= 0 (got %r)" % (V,))
    Ts = Ts_tilde(c)
    if a == "Ea":
        E = E_tilde(c)
        E_eff = E - Ts * V
        # ln eta at the hot bottom (T = 1, p0_norm = 1) is -E_eff/(Ts (1+Ts))
        ln_D = E_eff / (Ts * (1.0 + Ts))
        D = math.exp(ln_D)
        Ea = c.E_a
    else:
        if deta_T is None:
            raise ValueError('anchor="D" needs deta_T (the frozen contrast)')
        D = float(deta_T)
        if not (math.isfinite(D) and D > 1.0):
            raise ValueError("deta_T must exceed 1 under anchor='D' (got %r)" % (deta_T,))
        ln_D = math.log(D)
        E_eff = ln_D * Ts * (1.0 + Ts)
        E = E_eff + Ts * V
        Ea = infer_E_a(E, c)
    C = -E / Ts
    hot = (E + V) / (1.0 + Ts) + C
    return dict(anchor=a, law="pV(L) reference hydrostatic", E=float(E), V=float(V),
                Ts=float(Ts), C=float(C), E_eff=float(E_eff), ln_D=float(ln_D), D=float(D),
                Ea=float(Ea), V_a=float(c.V_a), R=float(c.R), rho0=float(c.rho0),
                g=float(c.g), d=float(c.d), DeltaT=float(c.DeltaT), T_surf=float(c.T_surf),
                E_a=float(c.E_a), ratio_VE=float(V / E), p0_ref=float(c.rho0 * c.g * c.d),
                contrast_log10=float(ln_D / math.log(10.0)), ln_eta_hot_end=float(hot),
                hot_end_residual=float(abs(hot + ln_D)))
