"""The energy-budget diagnostics, restated (oracle copy).

TEST INFRASTRUCTURE: restatement used only as the checker (see ``oracle``).

* ``theta_mean`` / ``nu_pair``  -- heating.py:269-276 (volume mean of the
  grid field; the boundary fluxes 1 - dtheta_0/dz at z = 0, 1);
* ``heating_balance``           -- heating.py:279-306 (storage, fluxes,
  heating, residual of the discrete energy budget);
* ``phi_integral``              -- dissipation.py:365-392 (<Phi> of the
  tau-zeroed spectral image of the padded dissipation field);
* ``pi_theta``                  -- dissipation.py:423-444 (Ra <u_z theta>
  through the same padded synthesis and truncation, dealias=True);
* ``dissipation_balance``       -- dissipation.py:540-571 (the heating
  budget less (Di/Ra) <Phi>).

The models are the oracle's (``oracle.models``): ``grid``, ``vol_mean``, the
basis ``b`` with its padded synthesis / truncation, ``stokes`` and the eta the
last Stokes solve consumed (``dissipation_eta``).
"""
import numpy as np

from .models import phi_padded


def theta_mean(m, that):
    """<theta> = vol_mean(grid(that)) (heating.py:275-276)."""
    return float(m.vol_mean(m.grid(that)))


def nu_pair(m, that):
    """(Nu_bot, Nu_top) = 1 - dtheta_0/dz at z = 0 and 1 (heating.py:269-272)."""
    dth = m.D @ np.asarray(that)[0]
    return float(1.0 - (m.V[0] @ dth).real), float(1.0 - (m.V[-1] @ dth).real)


def heating_balance(m, np1, n, nm1, a0, b0, b1, H=0.0, advection=None):
    """heating.py:279-306."""
    nu_bot, nu_top = nu_pair(m, np1)
    storage = (a0 * theta_mean(m, np1) - b0 * theta_mean(m, n)
               + (b1 * theta_mean(m, nm1) if nm1 is not None else 0.0))
    adv = 0.0 if advection is None else float(advection)
    resid = storage - (nu_bot - nu_top) - float(H) - adv
    return dict(storage=storage, nu_bot=nu_bot, nu_top=nu_top, heating=float(H),
                advection=adv, residual=resid, theta_mean=theta_mean(m, np1),
                T_mean=0.5 + theta_mean(m, np1))


def phi_integral(m, psi, eta_grid=None):
    """<Phi> of the source's tau-zeroed coefficients (dissipation.py:365-392)."""
    b = m.b
    c = (b.Vi @ b.trunc_pad(phi_padded(b, psi, eta_grid)).T).T
    c[:, -2:] = 0.0
    return float(m.vol_mean(m.grid(c)))


def pi_theta(m, psi, that):
    """Ra <u_z theta> with u_z and theta synthesised on the padded grid, the
    product truncated back (dissipation.py:423-444, dealias=True)."""
    b = m.b
    uz = b.synth_pad(-1j * m.kx[:, None] * np.asarray(psi))
    th = b.synth_pad(np.asarray(that))
    cf = (b.Vi @ b.trunc_pad(uz * th).T).T
    return float(m.Ra * m.vol_mean(m.grid(cf)))


def dissipation_balance(m, np1, n, nm1, a0, b0, b1, H=0.0, Di=0.0, psi=None, eta=None,
                        Ra=None):
    """dissipation.py:540-571: psi and eta default to the Stokes solution of
    ``np1`` and the eta that solve consumed."""
    base = heating_balance(m, np1, n, nm1, a0, b0, b1, H=H)
    if float(Di) == 0.0:
        base["dissipation"] = 0.0
        return base
    if psi is None:
        psi, _ = m.stokes(np1)
    if eta is None:
        eta = m.dissipation_eta()
    Ra_v = float(m.Ra if Ra is None else Ra)
    diss = (float(Di) / Ra_v) * phi_integral(m, psi, eta)
    out = dict(base)
    out["dissipation"] = diss
    out["residual"] = base["residual"] - diss
    return out
