"""The uneven-grid ensemble and the served stream at their real shapes
(MI355X), cut down to a few steps per run.

* The 256-run ensemble draw (grids {32x24, 64x48, 64x96, 128x96}, eta_z /
  eta_T / Arrhenius, Ra 1e4-1e6, deta_T 1-1e4, H 0-2, Lx 2 or 4): a subset
  that holds every grid and every law is served in mixed precision by
  concurrent buckets with two slots each; every run equals its solo mixed
  march (identical step counts, theta and diagnostics to 1e-12), and the FP64
  solo marches of the members up to 64x48 equal the oracle's to 1e-10.
* The heated 2:1 case in mixed precision against the reference's march of
  its first steps.
* The Poisson stream draw (grids up to 256x192, Arrhenius with the window
  [1e-5, 1], Ra 1e4-1e7, Di 0-0.5): the first 128x192 and 256x192 requests
  and two small ones, served the same way, equal their solo marches.
"""
import numpy as np
import pytest
import torch

import oracle
from golden_io import rel

pytestmark = pytest.mark.gpu

TSTOP = 8e-5            # a few steps per run at the production dtmax of 4e-5


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    from specmc_amd import runtime
    return runtime.require_device()


def _law(cfg):
    if cfg.viscosity_model == "arrhenius":
        return "arrhenius"
    return cfg.viscosity_model


def _ensemble_subset():
    """First requests of the ensemble draw until every grid and every law has
    appeared, at most two per grid."""
    from specmc_amd.workloads import ensemble_requests
    grids, laws, per_grid, out = set(), set(), {}, []
    for r in ensemble_requests(n=256, seed=3, tstop=TSTOP):
        g = (r.cfg.Nx, r.cfg.Nz)
        if per_grid.get(g, 0) >= 2 or (g in grids and _law(r.cfg) in laws):
            continue
        out.append(r)
        per_grid[g] = per_grid.get(g, 0) + 1
        grids.add(g)
        laws.add(_law(r.cfg))
        if len(grids) == 4 and len(laws) == 3 and len(out) >= 6:
            break
    assert grids == {(32, 24), (64, 48), (64, 96), (128, 96)}
    assert laws == {"eta_z", "eta_T", "arrhenius"}
    return out


STREAM_STEPS = 3          # steps per stream request (Ra up to 1e7: dt ~ 1e-7)


def _stream_subset():
    """The first request of each of four grids of the stream draw, each cut
    to STREAM_STEPS steps: its tstop becomes the time its solo march reaches
    after them (the served run accumulates the same dt sequence, so it
    retires after the same steps)."""
    from specmc_amd.workloads import stream_requests
    reqs = stream_requests(n=64, seed=4, rate=1.0)
    pick = []
    for want in ((128, 192), (256, 192), (32, 24), (64, 48)):
        r = next(q for q in reqs if (q.cfg.Nx, q.cfg.Nz) == want)
        r.tstop = 1.0
        pick.append(r)
    for i, r in enumerate(pick):
        r.arrival = 0.2 * i
    return pick


class _Stop(Exception):
    pass


def _solo(r, precision, max_steps=None):
    """Solo march of ``r`` to its tstop, or for ``max_steps`` steps (then
    ``r.tstop`` is set to the time reached)."""
    from specmc_amd import api
    m = api.build_model(r.cfg, precision=precision)
    last = {}

    def cb(n, t, s):
        last.update(n=n, t=t, s=s)
        if max_steps is not None and n >= max_steps:
            raise _Stop()

    try:
        s, info = api.march_adaptive(m, r.theta0, r.tstop, on_step=cb)
        steps = info["steps"]
    except _Stop:
        s, steps = last["s"], last["n"]
        r.tstop = last["t"]
    print("  solo %s %dx%d Ra=%.3g deta=%.3g: %d steps, t=%.3g" % (
        r.rid, r.cfg.Nx, r.cfg.Nz, r.cfg.Ra, r.cfg.deta_T, steps, r.tstop), flush=True)
    return np.real(s), steps, m.diagnostics(s)


def _serve_and_compare(reqs, precision, max_steps=None):
    from specmc_amd.serve import Server
    solo = {r.rid: _solo(r, precision, max_steps) for r in reqs}
    srv = Server(slots=2, concurrent=True, graph_rounds=2, precision=precision)
    srv.submit(reqs)
    done = srv.run(max_wall=600)
    st = srv.stats()
    print("served %d requests, %d run-steps in %.1fs, buckets %s" % (
        st["requests"], st["run_steps"], st["wall_s"], st["buckets"]), flush=True)
    assert len(done) == len(reqs)
    for r in done:
        s, steps, d = solo[r.rid]
        err = rel(r.theta, s)
        print("  %s %s %dx%d Ra=%.3g deta=%.3g: steps %d/%d theta rel %.2e" % (
            r.rid, r.cfg.viscosity_model, r.cfg.Nx, r.cfg.Nz, r.cfg.Ra, r.cfg.deta_T,
            r.steps, steps, err), flush=True)
        assert r.steps == steps
        assert err < 1e-12
        for k in ("Nu_top", "Nu_energy", "urms"):
            assert abs(r.diag[k] - d[k]) <= 1e-12 * max(abs(d[k]), 1.0), (r.rid, k)


def test_ensemble_subset_served_equals_solo(dev):
    _serve_and_compare(_ensemble_subset(), "mixed")


def test_ensemble_small_members_match_oracle(dev):
    """FP64 device marches of the ensemble members up to 64x48 against the
    oracle's marches of the same requests."""
    n = 0
    for r in _ensemble_subset():
        if r.cfg.Nx * r.cfg.Nz > 64 * 48:
            continue
        s, steps, d = _solo(r, "fp64")
        m = oracle.build_oracle(r.cfg)
        so, info = oracle.march_adaptive(m, np.asarray(r.theta0, dtype=complex), r.tstop)
        do = m.diagnostics(so)
        err = rel(s, np.real(so))
        print("  %s %s %dx%d: steps %d/%d theta rel %.2e" % (
            r.rid, r.cfg.viscosity_model, r.cfg.Nx, r.cfg.Nz, steps, info["steps"], err),
            flush=True)
        assert steps == info["steps"]
        assert err < 1e-10
        for k in ("Nu_top", "Nu_energy", "urms"):
            assert abs(d[k] - do[k]) <= 1e-10 * max(abs(do[k]), 1.0), (r.rid, k)
        n += 1
    assert n >= 2


def test_heated_2to1_mixed_matches_reference(dev):
    """The heated 2:1 case (64x96, Lx = 4, Ra = 1.25e4, H = 1.5) in mixed
    precision against the reference's own march of its first steps
    (c2_heated_64x96): the same steps, Nu and Vrms within the mixed bar
    of 1e-6 relative."""
    from specmc_amd import ConvectionConfig, api
    from golden_io import load, seeded_theta
    g = load("c2_heated_64x96")
    meta = g["meta"]
    cfg = ConvectionConfig(**meta["fields"])
    m = api.build_model(cfg, precision="mixed")
    th0 = seeded_theta(m, float(meta["fields"]["internal_heating"]), meta["ic"]["seed"])
    s, info = api.march_adaptive(m, np.real(th0), meta["tstop"])
    assert info["steps"] == meta["info"]["steps"]
    d = m.diagnostics(s)
    err = rel(s, g["final"].real)
    print("mixed heated 2:1: %d steps, theta rel %.2e" % (info["steps"], err), flush=True)
    assert err < 1e-6
    for k in ("Nu_top", "Nu_energy", "urms"):
        ref = meta["diagnostics"][k]
        assert abs(d[k] - ref) <= 1e-6 * abs(ref), (k, d[k], ref)


def test_stream_subset_served_equals_solo(dev):
    _serve_and_compare(_stream_subset(), "mixed", max_steps=STREAM_STEPS)
