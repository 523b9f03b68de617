"""FGMRES rounds replayed from HIP graphs (MI355X): the same trajectories,
iteration counts and batch results as the eager launches.

Each solve replays its captured graph many times with the program's eager
work (fg_init, the polls, the rest of the step) and host synchronisations in
between, so every replay is checked against the eager trajectory.
Tolerance: the FP64 bar, 1e-10 relative on theta against the reference
fixtures, identical iteration counts; bit-identical to the eager run.
"""
import numpy as np
import pytest
import torch

from golden_io import load, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    from specmc_amd import runtime
    return runtime.require_device()


def _march(cls, meta, kw, graph_rounds):
    from specmc_amd import api
    m = cls(graph_rounds=graph_rounds, **kw)
    ic = meta["ic"]
    if ic["kind"] == "conductive":
        th0 = m.seed(amp=ic["amp"], m=ic.get("m", 1), n=ic.get("n", 1))
    else:
        from golden_io import seeded_theta
        th0 = seeded_theta(m, float(meta["fields"].get("internal_heating", 0.0)), ic["seed"])
    sits, tits = [], []

    def on_step(n, t, s):
        sits.append(int(getattr(m, "last_bfbt_iters", 0)))
        tits.append(int(m.theta.last_iters))

    if meta["tstop"] is not None:
        s, info = api.march_adaptive(m, th0, meta["tstop"], on_step=on_step)
    else:
        s, info = api.march(m, th0, meta["dt"], meta["nsteps"], on_step=on_step)
    return s, info, sits, tits


@pytest.mark.parametrize("name", ["etaz_16x24", "etaT_bfbt_16x24", "arrh_bfbt_16x24",
                                  "c1_short_64x96"])
@pytest.mark.parametrize("G", [1, 4])
def test_graph_rounds_match_reference_trajectory(dev, name, G):
    from specmc_amd import ConvectionConfig, api
    from specmc_amd.models import RBCCoupled, RBCImplicitTheta
    g = load(name)
    meta = g["meta"]
    cfg = ConvectionConfig(**meta["fields"])
    cls = RBCImplicitTheta if meta["fields"]["viscosity_model"] == "eta_z" else RBCCoupled
    kw = api.model_kwargs(cfg)
    s, info, sits, tits = _march(cls, meta, kw, G)
    s0, _, _, _ = _march(cls, meta, kw, 0)
    err = rel(s, g["final"].real)
    print("graph G=%d %s: theta rel %.3e vs ref, %.3e vs eager" % (G, name, err, rel(s, s0)),
          flush=True)
    assert info["steps"] == meta["info"]["steps"]
    assert err < 1e-10
    assert np.array_equal(s, s0)
    if meta["fields"]["viscosity_model"] != "eta_z":
        assert sits == [int(v) for v in g["stokes_iters"]]
    assert tits == [int(v) for v in g["theta_iters"]]


def test_graph_ensemble_matches_eager(dev):
    """Batched runs with graph replays (and runs retiring mid-way) give the
    eager batch's states."""
    from specmc_amd import ConvectionConfig
    from specmc_amd.batch import Ensemble
    from specmc_amd.core import bench_ic
    from specmc_amd.workloads import _Grid
    cfgs = [ConvectionConfig(viscosity_model="eta_T", deta_T=10.0 ** (1 + q % 3), Ra=1e4 * (1 + q),
                             Nx=32, Nz=48, Lx=2.0, precond="bfbt") for q in range(6)]
    g = _Grid(32, 48, 2.0)
    th = [bench_ic.seeded_theta(g, H=0.0, seed=q, sigma=2e-2).real for q in range(6)]
    tst = np.array([3e-4, 2e-4, 3e-4, 1e-4, 2.5e-4, 3e-4])
    out = []
    for G in (0, 3):
        ens = Ensemble(cfgs, th, tst, graph_rounds=G)
        ens.march()
        out.append((ens.states(), ens.steps.copy(), ens.eng.ssol.rounds))
    print("rounds eager %d, graph %d" % (out[0][2], out[1][2]), flush=True)
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][0], out[1][0])
