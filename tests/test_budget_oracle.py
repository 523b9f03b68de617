"""The oracle's energy-budget functions against closed forms (CPU).

Each check fixes a term independently of the restated formula: a volume mean
against the exact integral of a Chebyshev polynomial, Ra <u_z theta> and <Phi>
of single-mode fields against their integrals worked out by hand, and the
budget of a steady internally heated conductive state, which closes to zero
only with the right signs of the flux and heating terms.
"""
import numpy as np
from numpy.polynomial import chebyshev as C

import oracle
from oracle import budget as OB


def _model(Nz=24, H=0.0, vm="const", **kw):
    return oracle.build_oracle(dict(viscosity_model=vm, Ra=1e4, Nx=16, Nz=Nz, Lx=2.0,
                                    internal_heating=H, **kw))


def _int01(q):
    """Exact integral over z in [0, 1] of sum_j q_j T_j(2z - 1)."""
    ci = C.chebint(q, lbnd=-1)
    return 0.5 * float(C.chebval(1.0, ci))


def _dz(q):
    """Coefficients of d/dz of sum_j q_j T_j(2z - 1)."""
    return 2.0 * C.chebder(q)


def test_basis_convention():
    m = _model()
    assert np.allclose(m.V, C.chebvander(2.0 * m.z - 1.0, m.Nz - 1), atol=1e-13)


def test_theta_mean_is_the_exact_integral():
    m = _model()
    rng = np.random.default_rng(0)
    that = np.zeros((m.Nk, m.Nz), dtype=complex)
    q = rng.normal(size=m.Nz)
    that[0] = q
    that[1:] = rng.normal(size=(m.Nk - 1, m.Nz))      # no x-mean
    assert abs(OB.theta_mean(m, that) - _int01(q)) < 1e-13


def test_pi_theta_single_modes():
    # theta = 2 f(z) cos(kx), psi = -2 g(z) sin(kx) (coefficients f and i g at
    # mode 1): u_z = -psi_x = 2 k g cos(kx), so <u_z theta> = 2 k int f g dz
    m = _model()
    rng = np.random.default_rng(1)
    f, g = rng.normal(size=10), rng.normal(size=12)     # deg f + deg g <= Nz - 1
    that = np.zeros((m.Nk, m.Nz), dtype=complex)
    psi = np.zeros_like(that)
    that[1, :10] = f
    psi[1, :12] = 1j * g
    k = m.kx[1]
    want = m.Ra * 2.0 * k * _int01(C.chebmul(f, g))
    got = OB.pi_theta(m, psi, that)
    assert abs(got - want) <= 1e-12 * abs(want)


def test_phi_integral_single_mode():
    # psi = -2 g sin(kx): e_xx = -e_zz = psi_xz = -2 k g' cos, e_xz =
    # -(g'' + k^2 g) sin, so <Phi> = int 8 k^2 g'^2 + 2 (g'' + k^2 g)^2 dz
    # (degree <= Nz - 3: the zeroed tau rows hold nothing)
    m = _model()
    rng = np.random.default_rng(2)
    g = rng.normal(size=9)                              # degree 8
    psi = np.zeros((m.Nk, m.Nz), dtype=complex)
    psi[1, :9] = 1j * g
    k = m.kx[1]
    g1, g2 = _dz(g), _dz(_dz(g))
    s = C.chebadd(g2, k * k * g)
    want = _int01(C.chebadd(8.0 * k * k * C.chebmul(g1, g1), 2.0 * C.chebmul(s, s)))
    got = OB.phi_integral(m, psi, None)
    assert abs(got - want) <= 1e-12 * abs(want)
    # eta = 2 everywhere doubles it
    eta = np.full((m.Nx, m.Nz), 2.0)
    assert abs(OB.phi_integral(m, psi, eta) - 2.0 * want) <= 1e-12 * abs(want)


def test_steady_heated_conduction_closes_the_budget():
    # theta_0 = (H/2) z (1 - z): Nu_bot = 1 - H/2, Nu_top = 1 + H/2, no storage
    H = 1.5
    m = _model(H=H)
    that = np.zeros((m.Nk, m.Nz), dtype=complex)
    zc = np.zeros(m.Nz)
    # z (1 - z) = (1 - x^2) / 4 with x = 2z - 1: (T0 - T2) / 8
    zc[0], zc[2] = 1.0 / 8.0, -1.0 / 8.0
    that[0] = 0.5 * H * zc
    dt, r = 1e-3, 0.7
    a0, b0, b1 = (1 + 2 * r) / ((1 + r) * dt), (1 + r) / dt, r * r / ((1 + r) * dt)
    out = OB.heating_balance(m, that, that, that, a0, b0, b1, H=H)
    assert abs(out["nu_bot"] - (1.0 - 0.5 * H)) < 1e-12
    assert abs(out["nu_top"] - (1.0 + 0.5 * H)) < 1e-12
    assert abs(out["storage"]) < 1e-9 and abs(out["residual"]) < 1e-9
    assert abs(out["theta_mean"] - H / 12.0) < 1e-14
    # a heating term of the wrong sign, or a flux difference reversed, would
    # leave 2 H
    bad = out["storage"] - (out["nu_top"] - out["nu_bot"]) - H
    assert abs(bad + 2 * H) < 1e-9


def test_dissipation_balance_adds_the_source():
    m = _model(vm="eta_T", deta_T=1e2, dissipation_number=0.3)
    th = m.seed(amp=3e-2)
    s1 = oracle.march(m, th, 2e-5, 2)[0]
    d0 = OB.heating_balance(m, s1, th, None, 1.0, 1.0, 0.0, H=0.0)
    d1 = OB.dissipation_balance(m, s1, th, None, 1.0, 1.0, 0.0, H=0.0, Di=0.3)
    psi, _ = m.stokes(s1)
    phi = OB.phi_integral(m, psi, m.dissipation_eta())
    assert phi > 0.0
    assert abs(d1["dissipation"] - 0.3 / m.Ra * phi) < 1e-15
    assert abs((d0["residual"] - d1["residual"]) - d1["dissipation"]) < 1e-12
