"""CPU checks of the algebra the device kernels implement.

A numpy twin of the device pipeline (the same half-grid DFT matrices and
z-padded product transforms of ``specmc_amd.basis``) is compared with the
oracle's restatement of the reference operators.  These pin the design
choices before any kernel runs: the mirror half grids reproduce numpy's
rfft/irfft, the z-padded product is the exact truncated Chebyshev product,
and the pseudo-spectral Schur matvec equals the matrix-form one (both halves
of the one-sided convolution, partner sign -1, flipped cross term) on the
stream-function class of vectors.
"""
import numpy as np
import pytest

import oracle
from oracle.models import _SchurOps, cheb_mult_matrix
from specmc_amd.basis import ShapeOps, cheb_D, cheb_V, cgl


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max()
                 / max(np.abs(np.asarray(b)).max(), 1e-300))


@pytest.fixture(scope="module", params=[(16, 24, 2.0), (32, 48, 4.0)])
def shape(request):
    return ShapeOps(*request.param)


def test_basis_matches_numpy_polynomial():
    N = 20
    z = cgl(N)
    assert rel(cheb_V(z, N), np.polynomial.chebyshev.chebvander(2 * z - 1, N - 1)) == 0.0
    # oracle's own recurrence and closed-form derivative agree with numpy
    assert rel(oracle.spectral.cheb_vander(2 * z - 1, N), cheb_V(z, N)) < 1e-15
    assert rel(oracle.spectral.cheb_deriv(N), cheb_D(N)) < 1e-15


def test_half_grid_dfts_match_numpy_fft(shape):
    o = shape
    rng = np.random.default_rng(0)
    c = rng.normal(size=(o.Nk, o.Nz))
    s = rng.normal(size=(o.Nk, o.Nz))
    b = oracle.spectral.Basis(o.Nx, o.Nz, o.Lx)
    # padded synthesis of even (real) and odd (imaginary) coefficients
    ge = b.synth_pad(c)
    go = b.synth_pad(1j * s)
    Xe, Xo = (o.V @ c.T).T, (o.V @ s.T).T
    assert rel(o.Ce_p @ Xe, ge[:o.nrp]) < 1e-13
    assert rel(o.So_p @ Xo, go[:o.nrp]) < 1e-13
    # analysis of a full even / odd padded field back to Nk modes
    assert rel(o.Ae_p @ ge[:o.nrp], b.trunc_pad(ge).real) < 1e-13
    assert rel(o.Ao_p @ go[:o.nrp], b.trunc_pad(go).imag) < 1e-13
    # unpadded grid / coeffs, Nyquist counted once
    gu = b.grid(c)
    assert rel(o.Ce_u @ Xe, gu[:o.nru]) < 1e-13
    assert rel(o.So_u @ Xo, b.grid(1j * s)[:o.nru]) < 1e-13
    assert rel(o.Vi @ (o.Ae_u @ gu[:o.nru]).T, b.coeffs(gu).real.T) < 1e-12


def test_padded_z_product_is_exact_chebyshev_product(shape):
    o = shape
    assert rel(o.Tp @ o.Vp, np.eye(o.Nz)) < 1e-13
    rng = np.random.default_rng(1)
    a = rng.normal(size=o.Nz)
    f = rng.normal(size=o.Nz)
    want = cheb_mult_matrix(a, o.Nz) @ f
    got = o.Tp @ ((o.Vp @ a) * (o.Vp @ f))
    assert rel(got, want) < 1e-12
    # the Tp closed form is the first Nz rows of the inverse Vandermonde
    Vq = np.polynomial.chebyshev.chebvander(2 * o.zp - 1, o.Q - 1)
    assert rel(o.Tp, np.linalg.inv(Vq)[:o.Nz]) < 1e-10


def _twin_schur_matvec(o, eta_c, v):
    """Pseudo-spectral Schur matvec on s-parts (the device pipeline)."""
    G = o.G
    w = np.einsum("kij,kj->ki", G, v)
    k = o.kx[:, None]
    etaS = o.Ce_p @ (eta_c @ o.Vp.T)                    # (nrp, Q)
    fields = [(v, True, 1.0), (w, True, 2.0 * k ** 2),
              ((o.D2 @ w.T).T, True, 1.0), ((o.D @ w.T).T, False, k)]
    t = []
    for f, odd, sc in fields:
        Y = (sc * f) @ o.Vp.T                            # (Nk, Q)
        g = (o.So_p if odd else o.Ce_p) @ Y
        A = (o.Ao_p if odd else o.Ae_p) @ (g * etaS)
        t.append(A @ o.Tp.T)                             # (Nk, Nz)
    D, D2 = o.D, o.D2
    out = (D2 @ t[0].T).T - k ** 2 * t[0]
    out = out + (D2 @ t[1].T).T
    out = out + 2.0 * k ** 2 * t[2]
    out = out - 4.0 * k * (D @ t[3].T).T
    out[:, -2] = v @ o.V0
    out[:, -1] = v @ o.V1
    return out


def test_pseudospectral_schur_matvec_equals_matrix_form(shape):
    o = shape
    m = oracle.build_oracle(dict(viscosity_model="eta_T", Nx=o.Nx, Nz=o.Nz,
                                 Lx=o.Lx, precond="bfbt", deta_T=1e3))
    X, Z = np.meshgrid(m.x, m.z, indexing="ij")
    th = m.coeffs(0.2 * np.cos(2 * np.pi * X / o.Lx) * np.sin(np.pi * Z)
                  + 0.05 * np.cos(4 * np.pi * X / o.Lx) * np.sin(2 * np.pi * Z))
    eta = m.eta_of(th.real)
    ops = _SchurOps(m, eta)
    rng = np.random.default_rng(3)
    s = rng.normal(size=(o.Nk, o.Nz)) * np.exp(-np.arange(o.Nz) / 6.0)
    s[0] = 0.0
    want = ops.matvec(1j * s).reshape(o.Nk, o.Nz)
    got = _twin_schur_matvec(o, ops.eta_c.real, s)
    assert rel(got, want.imag) < 1e-11
    assert np.abs(want.real).max() <= 1e-12 * np.abs(want.imag).max()


def _twin_schur_apply(o, eta_c, v):
    """The three launches of smx_schur_apply on the host: per-mode composite
    synthesis, x-products, combination against Mc with the tau rows."""
    W, Mc = o.schur_apply_ops()
    Q, k = o.Q, o.kx[:, None]
    yz = np.einsum("kn,knc->kc", v, W).reshape(o.Nk, 3, Q)
    etaS = o.Ce_p @ (eta_c @ o.Vp.T)
    tz = np.empty_like(yz)
    for f in range(3):
        odd = f < 2
        g = (o.So_p if odd else o.Ce_p) @ yz[:, f]
        tz[:, f] = (o.Ao_p if odd else o.Ae_p) @ (g * etaS)
    a = np.concatenate([tz[:, 0], k ** 2 * tz[:, 1], -4.0 * k * tz[:, 2]], 1)
    out = a @ Mc
    out[:, -2] = v @ o.V0
    out[:, -1] = v @ o.V1
    return out


def test_three_launch_schur_apply_equals_matrix_form(shape):
    """The composite operators of the fused apply (Wz, Mc) give the
    reference's matrix-form Schur matvec (lagged_stokes.py:202-220)."""
    o = shape
    m = oracle.build_oracle(dict(viscosity_model="eta_T", Nx=o.Nx, Nz=o.Nz,
                                 Lx=o.Lx, precond="bfbt", deta_T=1e3))
    X, Z = np.meshgrid(m.x, m.z, indexing="ij")
    th = m.coeffs(0.2 * np.cos(2 * np.pi * X / o.Lx) * np.sin(np.pi * Z)
                  + 0.05 * np.cos(4 * np.pi * X / o.Lx) * np.sin(2 * np.pi * Z))
    ops = _SchurOps(m, m.eta_of(th.real))
    rng = np.random.default_rng(5)
    s = rng.normal(size=(o.Nk, o.Nz)) * np.exp(-np.arange(o.Nz) / 6.0)
    # the eight-launch form it replaces, operation for operation (every row)
    assert rel(_twin_schur_apply(o, ops.eta_c.real, s),
               _twin_schur_matvec(o, ops.eta_c.real, s)) < 1e-12
    s[0] = 0.0            # the k = 0 row of a Stokes unknown is zero
    want = ops.matvec(1j * s).reshape(o.Nk, o.Nz).imag
    got = _twin_schur_apply(o, ops.eta_c.real, s)
    assert rel(got, want) < 1e-11


def test_pseudospectral_theta_advection_equals_reference(shape):
    o = shape
    m = oracle.build_oracle(dict(viscosity_model="eta_z", Nx=o.Nx, Nz=o.Nz,
                                 Lx=o.Lx, deta_T=10.0))
    th0 = m.seed(amp=3e-2).real
    psi, _ = m.stokes(th0)
    m.prepare(th0)
    rng = np.random.default_rng(4)
    v = rng.normal(size=(o.Nk, o.Nz)) * np.exp(-np.arange(o.Nz) / 5.0)
    want = m.advect(v)
    ps = psi.imag
    ux = o.So_p @ ((o.D @ ps.T).T @ o.V.T)
    uz = o.Ce_p @ ((o.kx[:, None] * ps) @ o.V.T)
    vx = o.So_p @ ((o.kx[:, None] * v) @ o.V.T)
    vz = o.Ce_p @ ((o.D @ v.T).T @ o.V.T)
    got = (o.Ae_p @ (-(ux * vx + uz * vz))) @ o.Vi.T
    assert rel(got, want.real) < 1e-12
    assert np.abs(want.imag).max() <= 1e-12 * np.abs(want.real).max()
