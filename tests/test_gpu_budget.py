"""The energy budget on the device against the oracle's (MI355X).

After three steps of a march, ``model.energy_budget()`` must reproduce the
oracle's heating / dissipation balance of the same states (storage, fluxes,
heating, dissipation, residual, <theta>) and Ra <u_z theta>, to 1e-10 of the
budget's scale, for every model kind, with heating and dissipation on; and
the budgets of runs batched together equal their solo budgets.
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import budget as OB

pytestmark = pytest.mark.gpu

CASES = {
    "const_H_Di": dict(viscosity_model="const", Ra=1e4, Nx=16, Nz=24, internal_heating=0.5,
                       dissipation_number=0.1),
    "eta_z_H": dict(viscosity_model="eta_z", deta_T=1e2, Ra=1e4, Nx=16, Nz=24,
                    internal_heating=1.0),
    "eta_T_H_Di": dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=16, Nz=24, Lx=2.0,
                       internal_heating=1.0, dissipation_number=0.2, precond="bfbt"),
    "legacy_Di": dict(viscosity_model="eta_T", deta_T=1e2, Ra=2e4, Nx=16, Nz=24, Lx=2.0,
                      dissipation_number=0.3),
}
KEYS = ("storage", "nu_bot", "nu_top", "heating", "dissipation", "residual", "theta_mean")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    from specmc_amd import runtime
    return runtime.require_device()


@pytest.mark.parametrize("name", sorted(CASES))
def test_budget_matches_oracle(dev, name):
    from specmc_amd import ConvectionConfig, api
    fields = CASES[name]
    print("budget", name, flush=True)
    mo = oracle.build_oracle(fields)
    md = api.build_model(ConvectionConfig(**fields))
    dt = 2e-5
    s = [mo.seed(amp=3e-2)]
    for _ in range(3):
        s.append(mo.step(s[-1], dt))
    sd, _ = api.march(md, s[0], dt, 3)
    got = md.energy_budget()
    H, Di = fields.get("internal_heating", 0.0), fields.get("dissipation_number", 0.0)
    if fields["viscosity_model"] == "const":
        want = OB.dissipation_balance(mo, s[3], s[2], None, 1.0 / dt, 1.0 / dt, 0.0, H=H, Di=Di)
    else:
        want = OB.dissipation_balance(mo, s[3], s[2], s[1], mo.a0, mo.b0, mo.b1, H=H, Di=Di)
    psi, _ = mo.stokes(s[3])
    want["pi_theta"] = OB.pi_theta(mo, psi, s[3])
    scale = max(abs(want["storage"]), abs(want["nu_bot"]), abs(want["nu_top"]), H, 1.0)
    print("  residual %.3e (oracle %.3e), dissipation %.6e, pi_theta %.6e" % (
        got["residual"], want["residual"], got["dissipation"], got["pi_theta"]), flush=True)
    for k in KEYS:
        assert abs(got[k] - want[k]) <= 1e-10 * scale, (k, got[k], want[k])
    assert abs(got["pi_theta"] - want["pi_theta"]) <= 1e-10 * abs(want["pi_theta"])
    assert abs(got["T_mean"] - (0.5 + want["theta_mean"])) <= 1e-12


def test_batched_budgets_equal_solo(dev):
    from specmc_amd import ConvectionConfig, api
    from specmc_amd.batch import Ensemble
    cfgs = [ConvectionConfig(**CASES["eta_T_H_Di"]),
            ConvectionConfig(**dict(CASES["eta_T_H_Di"], Ra=2e4, internal_heating=0.0)),
            ConvectionConfig(**dict(CASES["eta_T_H_Di"], deta_T=1e2, dissipation_number=0.0))]
    solo, th0 = [], []
    for c in cfgs:
        m = api.build_model(c)
        t0 = m.seed(amp=2e-2)
        th0.append(t0)
        api.march_adaptive(m, t0, 1e-4)
        solo.append(m.energy_budget())
    ens = Ensemble(cfgs, th0, 1e-4)
    ens.march()
    b = ens.budgets()
    for i, sb in enumerate(solo):
        for k, v in sb.items():
            assert abs(b[k][i] - v) <= 1e-12 * max(1.0, abs(v)), (i, k, b[k][i], v)
