"""Fixed numerical constants of the stepper (the values of specmc/_numerics.py:67-148).

Not part of the public interface: the config cannot change them.  They are
kept under the reference's names because drivers log them by name
(``api.numerics()``).
"""
from __future__ import annotations

STOKES_MODE = "defect"
PARTNER_SIGN = -1.0
ZFORM = "reduced"
ZSOLVER = "tau"
USE_LU = True
DEALIAS = True
DEALIAS_Z = False
STOKES_RTOL = 1e-8
ETA_DEALIAS = 2
ETA_DRIFT_TOL = 0.1
ETA_T_CLIP = (0.0, 2.0)
THETA_SCHEME = "implicit_bdf2"
THETA_TIME_SCHEME = "bdf2"
THETA_EXTRAPOLATE = True
IMPLICIT_BUOY = False
THETA_RTOL = 1e-10
THETA_MAXITER = 200
THETA_RESTART = 40
CFL_SAFETY = 40.0
CFL_MAX = 80.0
CFL_FALLBACK = 20.0
DTMAX = 4.0e-5
DTMIN = 1e-10
TSTOP = 0.25
AMP_LIMIT = 5.0
STEADY_TOL = 1e-3
STEADY_WINDOW = 0.02
CKPT_DT = 0.01
CKPT_KEEP_SPACING = 0.05
DEFECT_ADAPTIVE = True
#: Krylov sizes of the BFBT Stokes route (coupled_model.py:139)
STOKES_RESTART = 80
STOKES_MAXITER = 800


def as_dict():
    """-> every constant as a plain ``{name: value}`` dict."""
    out = {}
    for name, value in globals().items():
        if name.startswith("_") or not name.isupper():
            continue
        if isinstance(value, (bool, int, float, str, tuple, list, type(None))):
            out[name] = value
    return out
