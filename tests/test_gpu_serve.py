"""Continuous batching on the device (MI355X): requests served in a mix --
different shapes, models and physics, staggered arrivals, slots reused after
earlier runs retire -- give the results of their solo marches.

Tolerance: FP64, 1e-12 relative on theta and the diagnostics, identical step
counts (the bar of the batch-vs-solo test of the ensemble).
"""
import numpy as np
import pytest
import torch

from golden_io import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    from specmc_amd import runtime
    return runtime.require_device()


def _requests():
    from specmc_amd import ConvectionConfig
    from specmc_amd.core import bench_ic
    from specmc_amd.serve import Request
    from specmc_amd.workloads import _Grid
    base = dict(Lx=2.0, precond="bfbt")
    specs = [
        ("a", dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=16, Nz=24), 0.0, 3e-4, 0.0),
        ("b", dict(viscosity_model="arrhenius", deta_T=1e3, arrhenius_Ts=1.0,
                   arrhenius_E=2.0 * np.log(1e3), Ra=2e4, Nx=16, Nz=24), 1.0, 2e-4, 0.0),
        ("c", dict(viscosity_model="eta_T", deta_T=1e2, Ra=1e4, Nx=16, Nz=24,
                   dissipation_number=0.3), 0.5, 2.5e-4, 0.0),
        ("d", dict(viscosity_model="eta_z", deta_T=1e2, Ra=1e4, Nx=16, Nz=24), 0.0, 2e-4, 0.0),
        ("e", dict(viscosity_model="const", Ra=1e4, Nx=16, Nz=24), 0.0, 2e-4, 0.0),
        ("f", dict(viscosity_model="eta_T", deta_T=1e3, Ra=3e4, Nx=32, Nz=24), 0.2, 2e-4, 0.3),
        ("g", dict(viscosity_model="eta_T", deta_T=3e2, Ra=1e4, Nx=16, Nz=24), 0.0, 1.5e-4, 0.6),
        ("h", dict(viscosity_model="eta_z", deta_T=1e3, Ra=2e4, Nx=16, Nz=24), 1.5, 1e-4, 0.6),
    ]
    out = []
    for i, (rid, f, H, tstop, arr) in enumerate(specs):
        cfg = ConvectionConfig(internal_heating=H, **base, **f)
        g = _Grid(cfg.Nx, cfg.Nz, cfg.Lx)
        th = bench_ic.seeded_theta(g, H=H, seed=10 + i, sigma=2e-2).real
        out.append(Request(rid, cfg, th, tstop, arr))
    return out


@pytest.mark.parametrize("concurrent", [False, True])
def test_served_mix_matches_solo(dev, concurrent):
    from specmc_amd import api
    from specmc_amd.serve import Server
    reqs = _requests()
    solo = {}
    for r in reqs:
        m = api.build_model(r.cfg)
        s, info = api.march_adaptive(m, r.theta0, r.tstop)
        solo[r.rid] = (s, info["steps"], m.diagnostics(s))
    # two slots per bucket: later arrivals reuse the slots of retired runs;
    # concurrent: the buckets step in worker threads on their own streams,
    # with graph-replayed FGMRES rounds
    srv = Server(slots=2, concurrent=concurrent, graph_rounds=2 if concurrent else 0)
    srv.submit(_requests())
    done = srv.run(max_wall=300)
    st = srv.stats()
    print("served %d requests, %d run-steps in %.2fs, p50 %.3fs p99 %.3fs, buckets %s" % (
        st["requests"], st["run_steps"], st["wall_s"], st["p50_s"], st["p99_s"],
        st["buckets"]), flush=True)
    assert len(done) == len(reqs)
    for r in done:
        s, steps, d = solo[r.rid]
        err = rel(r.theta, np.real(s))
        print("  %s: steps %d/%d theta rel %.2e latency %.3fs" % (r.rid, r.steps, steps, err,
                                                                  r.latency), flush=True)
        assert r.steps == steps
        assert err < 1e-12
        for k in ("Nu_top", "Nu_energy", "urms"):
            assert abs(r.diag[k] - d[k]) <= 1e-12 * max(abs(d[k]), 1.0), (r.rid, k)
        assert r.latency is not None and r.latency >= 0.0


def test_admission_by_memory_budget(dev):
    """A budget that holds one slot of the larger shape but no two buckets
    of different shapes: the second shape waits until the idle bucket of the
    first is released, a request of the first shape that arrives later waits
    for the second bucket, and every request completes with its solo
    result."""
    from specmc_amd import api
    from specmc_amd.serve import Server, run_bytes
    reqs = [r for r in _requests() if r.rid in ("a", "f", "g")]
    key_bytes = [run_bytes(Server.key_of(r.cfg)[0], r.cfg.Nx, r.cfg.Nz) for r in reqs]
    budget = max(key_bytes)
    assert min(key_bytes) + max(key_bytes) > budget
    srv = Server(slots=4, mem_budget=budget)
    peak = []
    srv.submit(reqs)
    done = srv.run(max_wall=300, on_retire=lambda r: peak.append(len(srv.buckets)))
    assert len(done) == 3
    assert max(peak) == 1
    for r in done:
        m = api.build_model(r.cfg)
        s, info = api.march_adaptive(m, r.theta0, r.tstop)
        assert r.steps == info["steps"]
        assert rel(r.theta, np.real(s)) < 1e-12
