"""The legacy Stokes route on the device (precond='legacy', the
ConvectionConfig default: a lagged dense Schur factorisation with Picard
defect correction against the true operator, refreshed on lag, eta drift or
a failed correction; coupled_model.py:509-581, 693-716) against the
oracle's legacy route, which tests/test_oracle_golden.py pins to the
reference's etaT_legacy_16x24 trajectory.

Three marches cover the three refresh triggers: the adaptive march (eta
drifts more than 10 % per step, so every step refreshes), a fixed small dt
(refresh every ``lag`` = 16 solves, up to 5 corrections per solve) and the
same with ``max_K=2`` (the correction misses with the older factor, is
redone after a refresh and passes).  The oracle's decisions in these
marches keep a margin of at least 1 % from their thresholds (drift against
0.1, correction against 1e-8), so the device takes the same branches.
FP64: per-step refresh / correction / failure counters and dt exactly,
theta, Nu and Vrms to 1e-10.  Mixed precision (FP32 lagged factor, FGMRES
in place of the Picard loop): Nu and Vrms to 1e-6 of the FP64 oracle.
"""
import functools

import numpy as np
import pytest
import torch

import oracle
from golden_io import rel

pytestmark = pytest.mark.gpu

TOL = 1e-10
BASE = dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=16, Nz=24, Lx=2.0)
MARCHES = {
    "drift": (dict(), ("adaptive", 1.2e-3)),
    "lag": (dict(), ("fixed", 2e-6, 40)),
    "retry": (dict(max_K=2, K=2), ("fixed", 2e-6, 40)),
}


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    from specmc_amd import runtime
    return runtime.require_device()


def _record(m, log):
    def cb(n, t, s):
        log.append((t, m.n_refresh, m.last_K, m.n_defect_failed, m.n_defect_double_fail))
    return cb


def _march(march, m, th0, cb, mod):
    if march[0] == "adaptive":
        return mod.march_adaptive(m, th0, march[1], on_step=cb)
    return mod.march(m, th0, march[1], march[2], on_step=cb)


@functools.lru_cache(maxsize=None)
def _oracle_run(name):
    extra, march = MARCHES[name]
    ref = oracle.build_oracle(dict(BASE, **extra))
    th0 = ref.seed(amp=3e-2)
    log = []
    s, info = _march(march, ref, th0, _record(ref, log), oracle)
    return th0, log, s, info, ref.diagnostics(s)


def _device_run(name, precision):
    from specmc_amd import ConvectionConfig, api
    extra, march = MARCHES[name]
    m = api.build_model(ConvectionConfig(**BASE, **extra), precision=precision)
    assert m.precond == "legacy"
    th0 = _oracle_run(name)[0]
    log = []
    s, info = _march(march, m, th0, _record(m, log), api)
    return m, log, s, info, m.diagnostics(s)


@pytest.mark.parametrize("name", sorted(MARCHES))
def test_legacy_route_matches_oracle(dev, name):
    print("legacy fp64", name, flush=True)
    _, rlog, sr, rinfo, dr = _oracle_run(name)
    m, log, s, info, d = _device_run(name, "fp64")
    assert info["steps"] == rinfo["steps"]
    np.testing.assert_allclose([r[0] for r in log], [r[0] for r in rlog], rtol=TOL, atol=0)
    assert [r[1:] for r in log] == [r[1:] for r in rlog]
    err = rel(s, np.real(sr))
    print("  %d steps, %d refreshes, %d failed corrections: theta rel %.2e"
          % (info["steps"], log[-1][1], log[-1][3], err), flush=True)
    assert err < TOL
    for k in ("Nu_top", "Nu_bot", "Nu_energy", "urms", "umax"):
        assert abs(d[k] - dr[k]) <= TOL * max(abs(dr[k]), 1.0), (k, d[k], dr[k])
    assert m.last_stokes_route == "legacy"
    if name == "retry":
        assert log[-1][3] > 0
    if name == "lag":
        assert log[-1][1] == 3 and max(r[2] for r in log) > 1


@pytest.mark.parametrize("name", ["drift", "lag"])
def test_legacy_route_mixed_precision(dev, name):
    print("legacy mixed", name, flush=True)
    _, rlog, sr, rinfo, dr = _oracle_run(name)
    m, log, s, info, d = _device_run(name, "mixed")
    err = rel(s, np.real(sr))
    enu = abs(d["Nu_top"] - dr["Nu_top"]) / abs(dr["Nu_top"])
    evr = abs(d["urms"] - dr["urms"]) / abs(dr["urms"])
    dsc = m.describe()
    print("  %d steps, %d refreshes: theta %.2e Nu %.2e Vrms %.2e"
          % (info["steps"], dsc["n_refresh"], err, enu, evr), flush=True)
    assert abs(info["steps"] - rinfo["steps"]) <= 1
    assert dsc["n_defect_double_fail"] == 0
    assert err < 1e-8
    assert enu < 1e-6 and evr < 1e-6


def test_legacy_batched_matches_solo(dev):
    """Legacy runs of different physics in one Ensemble (own lag counters,
    own frozen factors, refreshed on their own schedule) reproduce their
    solo marches."""
    from specmc_amd import ConvectionConfig, api
    from specmc_amd.batch import Ensemble
    base = dict(Nx=16, Nz=24, Lx=2.0)
    cfgs = [ConvectionConfig(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, **base),
            ConvectionConfig(viscosity_model="eta_T", deta_T=1e2, Ra=2e4, lag=5,
                             internal_heating=1.0, **base),
            ConvectionConfig(viscosity_model="arrhenius", deta_T=1e3, arrhenius_Ts=1.0,
                             arrhenius_E=2.0 * np.log(1e3), Ra=1e4, **base)]
    tstops = (3e-4, 2e-4, 2.5e-4)
    solo, th0s = [], []
    for c, ts in zip(cfgs, tstops):
        m = api.build_model(c)
        th0 = m.seed(amp=2e-2)
        th0s.append(th0)
        s, info = api.march_adaptive(m, th0, ts)
        solo.append((s, info["steps"], m.n_refresh))
    ens = Ensemble(cfgs, th0s, tstops)
    ens.march()
    st = ens.states()
    for i, (s, steps, nref) in enumerate(solo):
        e = rel(st[i], np.real(s))
        print("run %d: %d steps, %d refreshes, rel %.2e" % (i, steps, nref, e), flush=True)
        assert int(ens.steps[i]) == steps
        assert int(ens.eng.n_refresh[i]) == nref
        assert e < 1e-12
