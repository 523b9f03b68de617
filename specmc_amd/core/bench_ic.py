"""Benchmark initial conditions (the contract of specmc/core/bench_ic.py).

States are Chebyshev/Fourier coefficients of ``theta = T - (1 - z)`` on the
model grid (mirror-doubled periodic carrier of length ``Lx``).
"""
import numpy as np

__all__ = ["analytic_bl_field", "analytic_bl_theta", "conductive_theta",
           "ic_theta", "IC_KINDS", "seeded_theta", "heated_conductive_theta"]

IC_KINDS = ("aspect_file", "analytic_bl", "conductive")


def analytic_bl_field(model, z1=0.102367, z2=0.897633, amp=0.1, m=1, n=1):
    """Boundary-layer profile + perturbation on the grid (bench_ic.py:143-159)."""
    X, Z = np.meshgrid(np.asarray(model.x, float), np.asarray(model.z, float),
                       indexing="ij")
    T = np.where(Z < z1, 0.5 + 0.5 * (z1 - Z) / z1,
                 np.where(Z > z2, 0.5 * (1.0 - Z) / (1.0 - z2), 0.5))
    T = T + amp * np.cos(m * np.pi * X) * np.sin(n * np.pi * Z)
    g = T - (1.0 - Z)
    g[:, 0] = 0.0
    g[:, -1] = 0.0
    return g


def analytic_bl_theta(model, **kw):
    return model.coeffs(analytic_bl_field(model, **kw))


def conductive_theta(model, amp=1e-3, m=1, n=1):
    """Conductive profile + ``amp cos(m pi x) sin(n pi z)`` (bench_ic.py:166-168)."""
    return model.seed(amp=amp, m=m, n=n)


def ic_theta(model, kind, path=None, amp=1e-3, zinterp="pchip"):
    kind = str(kind).lower()
    if kind in ("analytic_bl", "analytic", "bl"):
        return analytic_bl_theta(model)
    if kind in ("conductive", "cond", "old"):
        return conductive_theta(model, amp=amp)
    if kind in ("aspect_file", "aspect", "file"):
        raise FileNotFoundError("the ASPECT ascii initial field "
                                "initial_temperature_case2a.txt is not shipped")
    raise ValueError("unknown initial condition %r (use one of %s)"
                     % (kind, IC_KINDS))


def heated_conductive_theta(z, H):
    """theta of the motionless steady state with internal heating H
    (heating.py:168-171): ``H z (1 - z) / 2``."""
    z = np.asarray(z, dtype=float)
    return 0.5 * float(H) * z * (1.0 - z)


def seeded_theta(model, H=0.0, seed=2, sigma=1e-2, nm=8):
    """Heated conductive profile plus seeded low modes
    ``sum_{m<8, n=1..8} a_mn cos(k_m x) sin(n pi z)``,
    ``a_mn ~ N(0, sigma^2)/(m^2 + n^2)`` from ``default_rng(seed)``."""
    rng = np.random.default_rng(seed)
    X, Z = np.meshgrid(model.x, model.z, indexing="ij")
    g = heated_conductive_theta(Z, H)
    for m in range(nm):
        for n in range(1, nm + 1):
            a = rng.normal(0.0, sigma) / float(m * m + n * n)
            g = g + a * np.cos(model.kx[m] * X) * np.sin(n * np.pi * Z)
    return model.coeffs(g)
