"""Isoviscous states with an x-antisymmetric part: the oracle's behaviour the
device is checked against, and the host-side rules (no GPU).

RBC.step (rbc_galerkin.py:396-436) evolves any complex state.  The oracle's
step on such a state is pinned through x-translation: on the periodic domain
the equations commute with x -> x + d, and the padded-grid product is exact
for the retained modes, so step(shift_d(a)) = shift_d(step(a)) for every d.
A shifted mirror state is a general state, and the mirror step is pinned by
the reference's golden vectors, so a wrong sign or a dropped parity term in
the general path breaks this identity.  Grid diagnostics are invariant under
shifts by whole grid cells.
"""
import numpy as np
import pytest

import oracle
from golden_io import rel, seeded_theta
from specmc_amd.engine import ANTISYM_TOL, KIND_CONST, antisym_part


def shift(m, c, d):
    """Coefficients of theta(x - d): mode k times exp(-i k d)."""
    return np.asarray(c) * np.exp(-1j * m.kx[:, None] * d)


@pytest.mark.parametrize("fields", [
    dict(viscosity_model="const", Ra=2e4, Nx=32, Nz=17),
    dict(viscosity_model="const", Ra=1e4, Nx=16, Nz=24, internal_heating=0.5,
         dissipation_number=0.3),
])
def test_oracle_step_commutes_with_x_shift(fields):
    m = oracle.build_oracle(fields)
    a = seeded_theta(m, fields.get("internal_heating", 0.0), 5, sigma=3e-2)
    dt = 1e-3
    # (a shift by Lx / 2 maps mirror states to mirror states)
    for d in (0.3137, 1.3, m.Lx / m.Nx * 3):
        b = shift(m, a, d)
        assert antisym_part(b) > 1e-3
        got, want = m.step(b, dt), shift(m, m.step(a, dt), d)
        assert rel(got, want) < 1e-13, d
    # a planted sign error in the general path (the x-odd theta_x flipped)
    # breaks the identity: what the check can see
    b = shift(m, a, 0.3137)
    psi, _ = m.stokes(b)
    G = m.b.synth_pad
    ux, uz = G(m.b.dz(psi)), G(-1j * m.kx[:, None] * psi)
    thx_bad = G(1j * m.kx[:, None] * np.real(b) - m.kx[:, None] * np.imag(b) * -1)
    N_bad = -(ux * thx_bad + uz * G(m.b.dz(b))) + uz
    N_ok = m.explicit_rhs(b, psi)
    assert rel(N_bad, N_ok) > 1e-3


def test_oracle_diagnostics_invariant_under_grid_shift():
    m = oracle.build_oracle(dict(viscosity_model="const", Ra=2e4, Nx=32, Nz=17))
    a = seeded_theta(m, 0.0, 7, sigma=3e-2)
    b = shift(m, a, 5 * m.Lx / m.Nx)
    da, db = m.diagnostics(a), m.diagnostics(b)
    for k in da:
        assert abs(da[k] - db[k]) <= 1e-12 * max(abs(da[k]), 1e-12), k
    assert abs(m.cfl_dt(a) - m.cfl_dt(b)) <= 1e-14 * m.cfl_dt(a)


def test_antisym_part():
    assert antisym_part(np.ones((3, 4))) == 0.0
    c = np.ones((3, 4), dtype=complex)
    assert antisym_part(c) == 0.0
    c[1, 2] += 1e-12j
    assert 0.0 < antisym_part(c) <= ANTISYM_TOL
    c[2, 3] = 0.5j
    assert antisym_part(c) == pytest.approx(0.5)


def test_server_refuses_antisymmetric_variable_viscosity_request():
    from specmc_amd import ConvectionConfig
    from specmc_amd.serve import Request, Server
    srv = Server(slots=2, mem_budget=1 << 30)
    cfg = ConvectionConfig(viscosity_model="eta_T", deta_T=1e2, Ra=1e4, Nx=16, Nz=24)
    m = oracle.build_oracle(dict(viscosity_model="const", Ra=1e4, Nx=16, Nz=24))
    th = shift(m, m.seed(amp=1e-2), 0.3)
    with pytest.raises(ValueError, match="mirror-symmetric"):
        srv.submit([Request(0, cfg, th, 0.01)])
    assert not srv.pending
    # the isoviscous model takes it, and so does a variable-viscosity model
    # given a mirror state
    srv.submit([Request(1, ConvectionConfig(viscosity_model="const", Ra=1e4, Nx=16, Nz=24),
                        th, 0.01),
                Request(2, cfg, np.real(th), 0.01)])
    assert [r.rid for r in srv.pending] == [1, 2]
    assert Server.key_of(srv.pending[0].cfg)[0] == KIND_CONST
