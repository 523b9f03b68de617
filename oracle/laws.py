"""Viscosity laws, the eta window and the heat sources (oracle copy).

TEST INFRASTRUCTURE: restatement used only as the checker (see ``oracle``).
"""
import numpy as np

CLIP_RTOL = 1e-12          # eta_bounds.py:199


def eta_exp_T(T, deta_T):
    """``exp(-ln(deta_T) T)`` (viscosity.py:157-165)."""
    return np.exp(-np.log(float(deta_T)) * np.asarray(T, dtype=float))


def eta_depth(z, deta_T):
    """Depth-only law on ``T = 1 - z`` (viscosity.py:168-189)."""
    return eta_exp_T(1.0 + (0.0 - 1.0) * np.asarray(z, dtype=float), deta_T)


def arrhenius_E(deta_T, Ts):
    """``E = ln(deta_T) Ts (1 + Ts)`` (arrhenius.py:208-223)."""
    return float(np.log(float(deta_T)) * float(Ts) * (1.0 + float(Ts)))


def arrhenius_exponent(T, deta_T, Ts):
    """Anchored exponent ``-ln(D)(1+Ts) T/(T+Ts)`` (arrhenius.py:231-264)."""
    T = np.asarray(T, dtype=float)
    s = float(Ts)
    L = float(np.log(float(deta_T)))
    return -L * (1.0 + s) * T / (T + s)


# reference-state constants of the pV calibration (reference_state.py:142-177):
# rho0, g, d, DeltaT, T_surf, V_a, E_a, R
PV_CONSTANTS = dict(rho0=3300.0, g=9.81, d=2.89e6, DeltaT=3000.0, T_surf=273.0, V_a=6e-6,
                    E_a=5.3e5, R=8.314462618)


def calibrate_pv(anchor, V, deta_T=None, c=PV_CONSTANTS):
    """(E, V, Ts) of the pV law (reference_state.py:277-357): Ts = T_surf /
    DeltaT; anchor 'Ea': E = E_a / (R DeltaT); anchor 'D': E = ln(deta_T)
    Ts (1 + Ts) + Ts V."""
    Ts = c["T_surf"] / c["DeltaT"]
    if anchor == "Ea":
        E = c["E_a"] / (c["R"] * c["DeltaT"])
    else:
        E = float(np.log(float(deta_T))) * Ts * (1.0 + Ts) + Ts * float(V)
    return float(E), float(V), float(Ts)


def arrhenius_exponent_V(T, p0n, E, V, Ts):
    """ln eta of the pV law, -E T / (Ts (T + Ts)) + V p0n / (T + Ts)
    (arrhenius.py:283-325), with p0n = 1 - z (reference_state.py:208-216)."""
    T = np.asarray(T, dtype=float)
    s = float(Ts)
    return -float(E) * T / (s * (T + s)) + float(V) * np.asarray(p0n, dtype=float) / (T + s)


def effective_bounds(lo, hi):
    """Dead-band window (eta_bounds.py:238-241)."""
    return float(lo) * (1.0 - CLIP_RTOL), float(hi) * (1.0 + CLIP_RTOL)


def clip_eta(eta, lo, hi):
    """Hard clip with the 1e-12 dead-band; the input comes back untouched when
    nothing is outside the effective window (eta_bounds.py:360-379)."""
    eta = np.asarray(eta, dtype=float)
    elo, ehi = effective_bounds(lo, hi)
    n = int(np.count_nonzero((eta < elo) | (eta > ehi)))
    if n == 0:
        return eta, 0
    return np.clip(eta, lo, hi), n


def smoothstep5(t):
    t = np.clip(np.asarray(t, dtype=float), 0.0, 1.0)
    return t * t * t * (t * (6.0 * t - 15.0) + 10.0)


def smooth_clip_eta(eta, lo, hi, w):
    """C^2 clamp in ln(eta) (eta_bounds.py:411-480); ``w <= 0`` is the hard clip."""
    w = float(w)
    if w <= 0.0:
        return clip_eta(eta, lo, hi)
    ylo, yhi = float(np.log(lo)), float(np.log(hi))
    y = np.log(np.asarray(eta, dtype=float))
    n = int(np.count_nonzero(((y > ylo) & (y < ylo + w))
                             | ((y > yhi - w) & (y < yhi))))
    Y = ylo + (y - ylo) * smoothstep5((y - ylo) / w)
    yo = yhi - (yhi - Y) * smoothstep5((yhi - Y) / w)
    return np.exp(yo), n


def n_clipped(eta, lo, hi):
    elo, ehi = effective_bounds(lo, hi)
    eta = np.asarray(eta)
    return int(np.count_nonzero((eta < elo) | (eta > ehi)))


def conductive_theta(z, H):
    """Motionless steady state with internal heating (heating.py:168-171)."""
    z = np.asarray(z, dtype=float)
    return 0.5 * float(H) * z * (1.0 - z)
