"""Device snapshots resume a march bit for bit (MI355X).

For each model kind and Stokes route a run marches N1 + N2 steps straight
through; a second run marches N1 steps, is saved with ``checkpoint.save_run``,
restored into a fresh model with ``checkpoint.load_run`` and marches N2 more.
The final states, the dt sequences and the solver counters must be
identical: the snapshot carries the BDF2 history and, on the lagged-LU
route, the frozen factor.  A served run preempted by a higher-priority
request and resumed from its snapshot must equal its solo march.
"""
import numpy as np
import pytest
import torch

from golden_io import rel

pytestmark = pytest.mark.gpu

CASES = {
    "const": (dict(viscosity_model="const", Ra=1e4, Nx=16, Nz=24), "fp64", {}),
    "eta_z": (dict(viscosity_model="eta_z", deta_T=1e2, Ra=1e4, Nx=16, Nz=24), "fp64", {}),
    "bfbt": (dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=16, Nz=24,
                  precond="bfbt"), "fp64", {}),
    "legacy": (dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=16, Nz=24,
                    internal_heating=1.0, dissipation_number=0.2), "fp64", {}),
    "legacy_mixed": (dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=16, Nz=24),
                     "mixed", {}),
    # every Stokes solve goes through the mixed rescue's reused FP32 factor
    "rescue_mixed": (dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=16, Nz=24,
                          precond="bfbt"), "mixed", dict(stokes_maxiter=2)),
}


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    from specmc_amd import runtime
    return runtime.require_device()


def _model(name):
    from specmc_amd import ConvectionConfig, api
    fields, precision, extra = CASES[name]
    cfg = ConvectionConfig(**fields)
    if not extra:
        return api.build_model(cfg, precision=precision)
    from specmc_amd.models import RBCCoupled
    kw = api.model_kwargs(cfg)
    kw.update(extra)
    return RBCCoupled(precision=precision, **kw)


def _march(m, s, tstop):
    """-> (final state, info, [(dt, state) per step])."""
    from specmc_amd import api
    rec, last = [], [0.0]

    def cb(n, t, x):
        rec.append((t - last[0], np.array(np.real(x))))
        last[0] = t
    s, info = api.march_adaptive(m, s, tstop, on_step=cb)
    return s, info, rec


@pytest.mark.parametrize("name", sorted(CASES))
def test_snapshot_resume_is_exact(dev, name, tmp_path):
    from specmc_amd import checkpoint
    print("snapshot resume", name, flush=True)
    t1, t2 = 1.2e-4, 2.4e-4
    a = _model(name)
    th0 = a.seed(amp=3e-2)
    _, _, ra = _march(a, th0, t2)
    b = _model(name)
    sb, ib, rb = _march(b, th0, t1)
    p = checkpoint.save_run(str(tmp_path), b, sb, ib["t"], n=ib["steps"])
    c = _model(name)
    s1, meta = checkpoint.load_run(p, c)
    assert meta["snapshot"] and meta["n"] == ib["steps"]
    assert np.array_equal(np.real(s1), np.real(sb))
    _, _, rc = _march(c, s1, t2 - ib["t"])
    n1 = len(rb)
    steps = n1 + len(rc)
    print("  %d + %d steps (straight %d)" % (n1, len(rc), len(ra)), flush=True)
    # the split run stops on its own clock; it can differ by one step at t2
    assert abs(steps - len(ra)) <= 1 and len(rc) >= 2
    k = min(steps, len(ra))
    for q in range(n1, k):
        dt_a, x_a = ra[q]
        dt_c, x_c = rc[q - n1]
        # dt is recorded as a difference of the march clocks (rounding only)
        assert abs(dt_a - dt_c) <= 1e-12 * dt_a, (q, dt_a, dt_c)
        assert np.array_equal(x_a, x_c), (q, float(np.abs(x_a - x_c).max()))
    if name not in ("const", "eta_z"):
        ea, ec = a.eng, c.eng
        if k == len(ra) == steps:
            for key in ("n_refresh", "n_rescue", "n_rescue_fp32", "bfbt_iters", "n_defect"):
                assert int(getattr(ea, key)[0]) == int(getattr(ec, key)[0]), key
        if name in ("legacy", "legacy_mixed", "rescue_mixed"):
            assert int(ec.n_refresh[0]) >= 1


@pytest.mark.parametrize("concurrent", [False, True])
def test_preempted_request_resumes_exactly(dev, concurrent):
    """Two slots, two long low-priority runs, then a short high-priority one:
    it preempts a running request, which later resumes from its snapshot;
    every result equals its solo march."""
    from specmc_amd import ConvectionConfig, api
    from specmc_amd.serve import Request, Server
    base = dict(viscosity_model="eta_T", deta_T=1e3, Nx=16, Nz=24, Lx=2.0, precond="bfbt")
    # the low-priority runs take ~50 steps, so they still run at t = 0.2 s.
    # Both stay at Ra = 1e4: at Ra = 2e4 this 16x24 run enters the BFBT
    # rescue regime near step 21, where its FP64 dt sequence depends on
    # roundoff (with one block-inversion kernel it reaches tstop in 58 steps,
    # with another its dt falls below 1e-6) -- a property of the run, not of
    # serving, which this test is about
    specs = [("lo1", dict(Ra=1e4), 2e-2, 2e-3, 0.0, 0), ("lo2", dict(Ra=1e4), 3e-2, 2e-3, 0.0, 0),
             ("hi", dict(Ra=1.5e4), 2e-2, 1e-4, 0.2, 5)]
    reqs = []
    for rid, f, amp, tstop, arr, pri in specs:
        cfg = ConvectionConfig(**base, **f)
        m = api.build_model(cfg)
        reqs.append(Request(rid, cfg, m.seed(amp=amp), tstop, arr, priority=pri))
    srv = Server(slots=2, concurrent=concurrent)
    srv.submit(reqs)
    done = {r.rid: r for r in srv.run(max_wall=300)}
    st = srv.stats()
    print("served (concurrent=%s): preempted %d, %s" % (
        concurrent, st["preempted"], {k: (r.steps, r.preemptions) for k, r in done.items()}),
        flush=True)
    assert len(done) == 3
    assert st["preempted"] >= 1 and sum(r.preemptions for r in done.values()) >= 1
    for r in done.values():
        m = api.build_model(r.cfg)
        s, info = api.march_adaptive(m, r.theta0, r.tstop)
        assert r.steps == info["steps"], r.rid
        assert rel(r.theta, np.real(s)) < 1e-12, r.rid
