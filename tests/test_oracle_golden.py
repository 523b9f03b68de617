"""The oracle against the reference's own trajectories (CPU).

The fixtures come from tests/golden/make_golden.py, which ran the reference
package in the development container.  Every step's dt, every Stokes / theta
FGMRES iteration count and the final state must match.
"""
import numpy as np
import pytest

import oracle
from golden_io import load, names, rel, seeded_theta

SMALL = [n for n in names() if not n.startswith(("c1_full", "oracle_"))]


def _run(name):
    g = load(name)
    meta = g["meta"]
    m = oracle.build_oracle(meta["fields"])
    th0 = g["theta0"]
    ic = meta["ic"]
    if ic["kind"] == "conductive":
        mine = m.seed(amp=ic["amp"], m=ic.get("m", 1), n=ic.get("n", 1))
    else:
        mine = seeded_theta(m, float(meta["fields"].get("internal_heating",
                                                          0.0)), ic["seed"])
    assert rel(mine, th0) < 1e-14
    dts, sits, tits, states = [], [], [], []
    last = {"t": 0.0}

    def on_step(n, t, s):
        dts.append(t - last["t"])
        last["t"] = t
        sits.append(int(getattr(m, "last_bfbt_iters", 0)))
        tits.append(int(getattr(m, "last_iters", 0)))
        if n % meta["keep_every"] == 0:
            states.append(np.array(s))

    if meta["tstop"] is not None:
        s, info = oracle.march_adaptive(m, th0, meta["tstop"], on_step=on_step)
    else:
        s, info = oracle.march(m, th0, meta["dt"], meta["nsteps"],
                               on_step=on_step)
    return g, m, s, info, dts, sits, tits, states


@pytest.mark.parametrize("name", SMALL)
def test_oracle_matches_reference_trajectory(name):
    g, m, s, info, dts, sits, tits, states = _run(name)
    meta = g["meta"]
    assert info["steps"] == meta["info"]["steps"]
    np.testing.assert_allclose(dts, g["dts"], rtol=1e-11, atol=0)
    assert list(sits) == [int(v) for v in g["stokes_iters"]]
    assert list(tits) == [int(v) for v in g["theta_iters"]]
    assert rel(s, g["final"]) < 1e-11, rel(s, g["final"])
    for a, b in zip(states, g["states"]):
        assert rel(a, b) < 1e-11
    d = m.diagnostics(s)
    for k in ("Nu_top", "Nu_bot", "Nu_energy", "urms", "umax", "amp"):
        ref = meta["diagnostics"][k]
        assert abs(d[k] - ref) <= 1e-10 * max(abs(ref), 1.0), (k, d[k], ref)


def test_golden_generator_in_sync():
    """The seeded IC used by the tests is the one the fixtures were made with."""
    import importlib.util
    import os
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                     "make_golden.py")
    spec = importlib.util.spec_from_file_location("make_golden", p)
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    m = oracle.build_oracle(dict(viscosity_model="const", Nx=16, Nz=24,
                                 Lx=4.0))
    assert rel(mg.seeded_theta(m, 1.5, 2), seeded_theta(m, 1.5, 2)) == 0.0
