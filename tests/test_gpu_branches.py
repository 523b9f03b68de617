"""Device path vs the reference's branch fixtures (MI355X).

Each fixture drives one decision of the coupled model that the trajectory
fixtures never reach: the forced BFBT rescue, a defect loop that fails twice,
the dense direct solve, the weak acceptance above the direct cap, a natural
rescue of the 64x96 run at its step-300 state, the smooth clip of the eta
window, the post-truncation positivity guard and the analytic boundary-layer
start.  Per step the route, the rescue / defect / refresh / clip counters,
the dt and the FGMRES iteration counts must equal the reference's; theta,
Nu and Vrms agree to 1e-10 relative (the FP64 bar).
"""
import numpy as np
import pytest
import torch

from golden_io import branch_names, initial_theta, load, rel
from test_oracle_golden import GUARD_PSI_TOL, check_counters

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")


def _model(meta):
    from specmc_amd import ConvectionConfig, api
    cfg = ConvectionConfig(**meta["fields"])
    spec = api.model_spec(cfg)
    kw = dict(spec["kwargs"])
    kw.update(meta.get("model_kw", {}))
    return spec["cls"](**kw)


def _counters(m, names):
    d = m.describe()
    return [float(d[c]) for c in names]


def _march_case(name):
    from specmc_amd import api
    g = load(name)
    meta = g["meta"]
    print("case", name, meta["fields"], meta.get("model_kw"), flush=True)
    m = _model(meta)
    th0 = initial_theta(m, g)
    assert rel(th0, g["theta0"]) < 1e-14
    names = meta["counter_names"]
    dts, sits, tits, counters, routes = [], [], [], [], []
    last = {"t": 0.0}

    def on_step(n, t, s):
        dts.append(t - last["t"])
        last["t"] = t
        sits.append(int(m.last_bfbt_iters))
        tits.append(int(m.theta.last_iters))
        counters.append(_counters(m, names))
        routes.append(m.last_stokes_route)

    th0 = np.real(th0)
    if meta["tstop"] is not None:
        s, info = api.march_adaptive(m, th0, meta["tstop"], on_step=on_step)
    else:
        s, info = api.march(m, th0, meta["dt"], meta["nsteps"], on_step=on_step)
    assert info["steps"] == meta["info"]["steps"]
    np.testing.assert_allclose(dts, g["dts"], rtol=TOL, atol=0)
    assert sits == [int(v) for v in g["stokes_iters"]]
    assert tits == [int(v) for v in g["theta_iters"]]
    assert routes == meta["routes"]
    check_counters(meta, np.asarray(counters), np.asarray(g["counters"]))
    err = rel(s, g["final"].real)
    print("theta rel err %.3e" % err, flush=True)
    assert err < TOL
    d = m.diagnostics(s)
    for k in ("Nu_top", "Nu_bot", "Nu_energy", "urms", "umax", "amp"):
        ref = meta["diagnostics"][k]
        assert abs(d[k] - ref) <= TOL * max(abs(ref), 1.0), (k, d[k], ref)


def _probe_case(name):
    g = load(name)
    meta = g["meta"]
    print("probe", name, meta["fields"], flush=True)
    m = _model(meta)
    th0 = np.real(initial_theta(m, g))
    eta = m.eta_of(th0)
    assert float(eta.min()) > 0.0
    assert rel(eta, g["eta"]) < 1e-12, rel(eta, g["eta"])
    psi, om = m.stokes(th0)
    assert rel(psi, g["psi"]) < GUARD_PSI_TOL, rel(psi, g["psi"])
    assert rel(om, g["omega"]) < GUARD_PSI_TOL, rel(om, g["omega"])
    assert [m.last_stokes_route] == meta["routes"]
    assert m.last_K == meta["last_K"]
    got = np.asarray([_counters(m, meta["counter_names"])])
    check_counters(meta, got, np.asarray(g["counters"]))


@pytest.mark.parametrize("name", branch_names())
def test_device_matches_reference_branches(name):
    _need_gpu()
    if load(name)["meta"].get("probe"):
        _probe_case(name)
    else:
        _march_case(name)
