"""Isoviscous states with an x-antisymmetric part on the device (MI355X).

RBC.step, cfl_dt, diagnostics, stokes and velocity (rbc_galerkin.py:342-471)
take any complex state; the device carries one as two real channels.  Here
the FP64 device path must match the oracle to 1e-10 relative over a short
march (with and without heating and dissipation, the energy budget
included), commute with x-translation as the oracle does, resume exactly
from a snapshot, and step the mirror runs of the same engine as before.  The
variable-viscosity models project such a state in ``step`` (symmetrize=True,
convection_modes.py:198, implicit_theta.py:527) and refuse it elsewhere.
"""
import warnings

import numpy as np
import pytest
import torch

import oracle
from oracle import budget as OB
from golden_io import rel

pytestmark = pytest.mark.gpu

TOL = 1e-10

CASES = {
    "plain_32x25": dict(viscosity_model="const", Ra=2e4, Nx=32, Nz=25),
    "H_Di_16x24": dict(viscosity_model="const", Ra=1e4, Nx=16, Nz=24, internal_heating=0.5,
                       dissipation_number=0.1),
}


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    from specmc_amd import runtime
    return runtime.require_device()


def general_theta(model, H, seed, sigma=3e-2, nm=8):
    """A heated conductive profile plus cos and sin x-modes: real grid, complex
    coefficients with an O(sigma) imaginary part."""
    rng = np.random.default_rng(seed)
    X, Z = np.meshgrid(model.x, model.z, indexing="ij")
    g = 0.5 * H * Z * (1.0 - Z)
    for m in range(nm):
        for n in range(1, nm + 1):
            a, b = rng.normal(0.0, sigma, 2) / float(m * m + n * n)
            g = g + (a * np.cos(model.kx[m] * X) + b * np.sin(model.kx[m] * X)) \
                * np.sin(n * np.pi * Z)
    return model.coeffs(g)


def shift(m, c, d):
    return np.asarray(c) * np.exp(-1j * m.kx[:, None] * d)


def _cmp_diag(got, want):
    for k in want:
        assert abs(got[k] - want[k]) <= TOL * max(abs(want[k]), 1e-8), (k, got[k], want[k])


@pytest.mark.parametrize("name", sorted(CASES))
def test_general_state_matches_oracle(dev, name):
    from specmc_amd import ConvectionConfig, api
    from specmc_amd.engine import antisym_part
    fields = CASES[name]
    print("general state", name, flush=True)
    H, Di = fields.get("internal_heating", 0.0), fields.get("dissipation_number", 0.0)
    mo = oracle.build_oracle(fields)
    md = api.build_model(ConvectionConfig(**fields))
    th = general_theta(mo, H, 11)
    # the k = 0 mode's imaginary part is invisible on the grid and diffuses
    th[0] = th[0] + 1j * 1e-3 * np.sin(np.arange(mo.Nz))
    assert antisym_part(th) > 1e-2
    psi_o, om_o = mo.stokes(th)
    psi_d, om_d = md.stokes(th)
    assert rel(psi_d, psi_o) < TOL and rel(om_d, om_o) < TOL
    for g, w in zip(md.velocity(psi_d), mo.velocity(psi_o)):
        assert rel(g, w) < TOL
    assert abs(md.cfl_dt(th) - mo.cfl_dt(th)) <= TOL * mo.cfl_dt(th)
    _cmp_diag(md.diagnostics(th), mo.diagnostics(th))
    dt = 2e-4
    so, sd = [th], [th]
    for i in range(4):
        so.append(mo.step(so[-1], dt))
        sd.append(md.step(sd[-1], dt))
        e = rel(sd[-1], so[-1])
        print("  step %d rel %.2e  Im part %.3e" % (i + 1, e, antisym_part(so[-1])), flush=True)
        assert e < TOL
    _cmp_diag(md.diagnostics(sd[-1]), mo.diagnostics(so[-1]))
    # the energy budget of the last step
    got = md.energy_budget()
    want = OB.dissipation_balance(mo, so[4], so[3], None, 1.0 / dt, 1.0 / dt, 0.0, H=H, Di=Di)
    psi, _ = mo.stokes(so[4])
    want["pi_theta"] = OB.pi_theta(mo, psi, so[4])
    scale = max(abs(want["storage"]), abs(want["nu_bot"]), abs(want["nu_top"]), H, 1.0)
    for k in ("storage", "nu_bot", "nu_top", "dissipation", "residual", "theta_mean"):
        assert abs(got[k] - want[k]) <= TOL * scale, (k, got[k], want[k])
    assert abs(got["pi_theta"] - want["pi_theta"]) <= TOL * abs(want["pi_theta"])


def test_general_step_commutes_with_x_shift(dev):
    """step(shift(a)) = shift(step(a)) on the device for a mirror state a:
    the general path against the mirror path of the same kernels' library."""
    from specmc_amd import ConvectionConfig, api
    fields = dict(CASES["H_Di_16x24"], Nx=32)
    cfg = ConvectionConfig(**fields)
    m1, m2 = api.build_model(cfg), api.build_model(cfg)
    a = m1.seed(amp=3e-2) + m1.seed(amp=1e-2, m=2, n=2)
    dt, d = 2e-4, 0.3137
    x1, x2 = a, shift(m1, a, d)
    for _ in range(3):
        x1, x2 = m1.step(x1, dt), m2.step(x2, dt)
        assert rel(x2, shift(m1, x1, d)) < 1e-12
    assert not m1.eng.general and m2.eng.general


def test_general_snapshot_resume_is_exact(dev, tmp_path):
    from specmc_amd import ConvectionConfig, api, checkpoint
    cfg = ConvectionConfig(**CASES["H_Di_16x24"])
    a, b = api.build_model(cfg), api.build_model(cfg)
    th = general_theta(a, 0.5, 4)
    dt = 2e-4
    sa = th
    for _ in range(4):
        sa = a.step(sa, dt)
    sb = th
    for _ in range(2):
        sb = b.step(sb, dt)
    p = checkpoint.save_run(str(tmp_path), b, sb, 2 * dt, n=2)
    c = api.build_model(cfg)
    sc, _ = checkpoint.load_run(p, c)
    for _ in range(2):
        sc = c.step(sc, dt)
    assert np.array_equal(sc, sa)


def test_mirror_runs_keep_their_path_beside_a_general_run(dev):
    from specmc_amd import ConvectionConfig, api
    from specmc_amd.batch import Ensemble
    cfg = ConvectionConfig(**CASES["plain_32x25"])
    m = api.build_model(cfg)
    a = m.seed(amp=3e-2)
    g = general_theta(m, 0.0, 8)
    ens = Ensemble([cfg, cfg], [a, g], 2e-3)
    ens.march()
    st = ens.states()
    assert ens.eng.general
    ref = []
    for t0 in (a, g):
        mm = api.build_model(cfg)
        s, _ = api.march_adaptive(mm, t0, 2e-3)
        ref.append(s)
    assert rel(st[0], ref[0]) < 1e-12 and np.abs(np.imag(st[0])).max() == 0.0
    assert rel(st[1], ref[1]) < 1e-12


def test_variable_viscosity_step_projects(dev):
    from specmc_amd import ConvectionConfig, api
    fields = dict(viscosity_model="eta_z", deta_T=1e2, Ra=1e4, Nx=16, Nz=24)
    m1 = api.build_model(ConvectionConfig(**fields))
    m2 = api.build_model(ConvectionConfig(**fields))
    th = general_theta(m1, 0.0, 2)
    dt = 2e-4
    s1 = m1.step(th, dt)
    s2 = m2.step(np.real(th), dt)
    assert np.array_equal(s1, s2)
    assert m1.sym_removed > 1e-3 and m2.sym_removed < 1e-12
    with pytest.raises(ValueError, match="mirror-symmetric"):
        m1.cfl_dt(th)
    with pytest.raises(ValueError, match="mirror-symmetric"):
        m1.diagnostics(th)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)
        m3 = api.build_model(ConvectionConfig(symmetrize=False, **fields))
    with pytest.raises(ValueError, match="mirror-symmetric"):
        m3.step(th, dt)


@pytest.mark.parametrize("concurrent", [False, True])
def test_served_general_requests_match_solo(dev, concurrent):
    """Isoviscous requests with and without an antisymmetric part, served
    through two slots (the third request reuses a retired run's slot), give
    their solo marches' states, step counts and diagnostics."""
    from specmc_amd import ConvectionConfig, api
    from specmc_amd.serve import Request, Server
    cfg = ConvectionConfig(**CASES["H_Di_16x24"])
    m = api.build_model(cfg)
    ths = [general_theta(m, 0.5, 21), m.seed(amp=2e-2), general_theta(m, 0.5, 22)]
    tstops = [2e-3, 1.5e-3, 1e-3]

    def reqs():
        return [Request("g%d" % i, cfg, th, ts, arrival=0.0)
                for i, (th, ts) in enumerate(zip(ths, tstops))]

    solo = {}
    for r in reqs():
        mm = api.build_model(cfg)
        s, info = api.march_adaptive(mm, r.theta0, r.tstop)
        solo[r.rid] = (s, info["steps"], mm.diagnostics(s))
    srv = Server(slots=2, concurrent=concurrent, graph_rounds=0)
    srv.submit(reqs())
    done = srv.run(max_wall=120)
    assert len(done) == 3
    for r in done:
        s, steps, d = solo[r.rid]
        err = rel(r.theta, s)
        print("  %s: steps %d/%d theta rel %.2e" % (r.rid, r.steps, steps, err), flush=True)
        assert r.steps == steps
        assert err < 1e-12
        for k in ("Nu_top", "Nu_energy", "urms", "umax"):
            assert abs(r.diag[k] - d[k]) <= 1e-12 * max(abs(d[k]), 1.0), (r.rid, k)
