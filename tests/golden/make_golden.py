"""Write the trajectory fixtures under tests/golden/ from the reference package.

Run in the development container only (the reference is not shipped anywhere):

    python tests/golden/make_golden.py [--only NAME ...] [--c1-full | --branches]

What it records, and why these need the reference itself: whole-trajectory
outputs of ``api.march_adaptive`` / ``api.march`` -- the combined effect of
the CFL policy, memo hits, FGMRES iteration paths and stopping, BDF2 history,
the backward-Euler first step and the mirror projections.  No closed form
pins those; every per-function piece is pinned elsewhere by closed forms.

Each fixture holds inputs (config fields as JSON, initial state, march
arguments) and outputs (states, dt sequence, iteration counts, diagnostics).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(OUT))


def _ref():
    sys.path.insert(0, REF)
    import specmc                       # noqa: F401
    from specmc import api, ConvectionConfig
    from specmc.core import bench_ic
    return api, ConvectionConfig, bench_ic


# per-step counters of the coupled model recorded by the branch fixtures: the
# rescue routing (coupled_model.py:613-681), the defect loop (523-574), the
# eta law and its positivity guard (333-496)
COUNTERS = ("n_bfbt_rescue", "n_bfbt_rescue_direct", "n_bfbt_rescue_weak",
            "n_defect_failed", "n_defect_double_fail", "n_refresh", "n_defect",
            "n_clip_T", "n_T_out_of_range", "n_clip_eta", "n_smooth_eta", "n_clip_eta_proj",
            "n_memo_hits", "n_stokes", "last_K")


def guard_theta(model, H, amp):
    """Heated conductive profile plus ``amp cos(pi x) sin(pi z)``: the state of
    the reference's positivity-guard self-test (coupled_model.py:927-953)."""
    X, Z = np.meshgrid(model.x, model.z, indexing="ij")
    return model.coeffs(0.5 * H * Z * (1.0 - Z)
                        + amp * np.cos(np.pi * X) * np.sin(np.pi * Z))


def seeded_theta(model, H, rng_seed, sigma=1e-2, nm=8):
    """Conductive profile with internal heating plus seeded low modes:
    ``sum_{m<8, n=1..8} a_mn cos(k_m x) sin(n pi z)``,
    ``a_mn ~ N(0, sigma^2) / (m^2 + n^2)``; returned as coefficients."""
    rng = np.random.default_rng(rng_seed)
    X, Z = np.meshgrid(model.x, model.z, indexing="ij")
    g = 0.5 * H * Z * (1.0 - Z)
    for m in range(nm):
        for n in range(1, nm + 1):
            a = rng.normal(0.0, sigma) / float(m * m + n * n)
            g = g + a * np.cos(model.kx[m] * X) * np.sin(n * np.pi * Z)
    return model.coeffs(g)


def run_case(name, fields, ic, tstop=None, nsteps=None, dt=None,
             keep_every=1, model_kw=None):
    """``model_kw``: constructor arguments beyond the config's (the model
    recipe's kwargs plus these), e.g. the Stokes budget of the reference's
    forced-rescue self-test (coupled_model.py:955-972)."""
    api, Cfg, bench_ic = _ref()
    cfg = Cfg(**fields)
    if model_kw:
        spec = api.model_spec(cfg)
        kw = dict(spec["kwargs"])
        kw.update(model_kw)
        m = spec["cls"](**kw)
    else:
        m = api.build_model(cfg)
    H = float(fields.get("internal_heating", 0.0))
    if ic["kind"] == "conductive":
        th0 = bench_ic.conductive_theta(m, amp=ic["amp"], m=ic.get("m", 1),
                                        n=ic.get("n", 1))
    elif ic["kind"] == "analytic_bl":
        th0 = bench_ic.analytic_bl_theta(m)
    elif ic["kind"] == "guard":
        th0 = guard_theta(m, H, ic["amp"])
    elif ic["kind"] == "state":
        with np.load(os.path.join(ROOT, ic["file"]), allow_pickle=False) as z:
            th0 = np.asarray(z[ic["key"]], dtype=complex)
    else:
        th0 = seeded_theta(m, H, ic["seed"])
    states, dts, stokes_its, theta_its, ts = [], [], [], [], []
    counters, routes = [], []
    last = {"t": 0.0}

    def on_step(n, t, s):
        dts.append(t - last["t"])
        last["t"] = t
        ts.append(t)
        stokes_its.append(int(getattr(m, "last_bfbt_iters", 0)))
        th = getattr(m, "theta", None)
        theta_its.append(int(th.last_iters) if th is not None else 0)
        if model_kw is not None or ic["kind"] in ("guard", "state", "analytic_bl"):
            counters.append([float(getattr(m, c, 0)) for c in COUNTERS])
            routes.append(str(getattr(m, "last_stokes_route", "")))
        if n % keep_every == 0:
            states.append(np.array(s))

    t0 = time.time()
    if tstop is not None:
        s, info = api.march_adaptive(m, th0, tstop, on_step=on_step)
    else:
        s, info = api.march(m, th0, dt, nsteps, on_step=on_step)
    wall = time.time() - t0
    diag = m.diagnostics(s)
    desc = {}
    if hasattr(m, "describe"):
        try:
            d = m.describe()
            desc = {k: v for k, v in d.items()
                    if isinstance(v, (int, float, str, bool)) or v is None}
        except Exception as e:                      # pragma: no cover
            desc = {"describe_error": repr(e)}
    meta = dict(name=name, fields=fields, ic=ic, tstop=tstop, nsteps=nsteps,
                dt=dt, keep_every=keep_every, info=info, diagnostics=diag,
                describe=desc, wall_s=wall, numpy=np.__version__,
                scipy=__import__("scipy").__version__)
    extra = {}
    if counters:
        meta.update(model_kw=model_kw or {}, counter_names=list(COUNTERS),
                    routes=routes,
                    theta_nonconv=int(getattr(getattr(m, "theta", None),
                                              "n_nonconv", 0)))
        extra["counters"] = np.asarray(counters)
    np.savez_compressed(os.path.join(OUT, name + ".npz"),
                        theta0=np.asarray(th0), final=np.asarray(s),
                        states=np.asarray(states), dts=np.asarray(dts),
                        ts=np.asarray(ts), stokes_iters=np.asarray(stokes_its),
                        theta_iters=np.asarray(theta_its),
                        meta=np.array(json.dumps(meta, default=float)), **extra)
    print("%-18s steps=%4d t=%.6g wall=%.1fs Nu_top=%.10f urms=%.10f"
          % (name, info["steps"], info["t"], wall, diag["Nu_top"],
             diag["urms"]), flush=True)


CASES = {
    # config 0: the isoviscous CPU-runnable case (25 steps at dtmax)
    "c0_const_32x24": dict(
        fields=dict(viscosity_model="const", Ra=1e4, Nx=32, Nz=24, Lx=2.0),
        ic=dict(kind="conductive", amp=3e-2), tstop=1e-3),
    "const_h_di_16x24": dict(
        fields=dict(viscosity_model="const", Ra=2e4, Nx=16, Nz=24, Lx=2.0,
                    internal_heating=1.0, dissipation_number=0.2),
        ic=dict(kind="conductive", amp=5e-2), nsteps=6, dt=2e-5),
    "etaz_16x24": dict(
        fields=dict(viscosity_model="eta_z", deta_T=100.0, Ra=3e4, Nx=16,
                    Nz=24, Lx=2.0, internal_heating=0.5,
                    dissipation_number=0.1),
        ic=dict(kind="conductive", amp=3e-2), tstop=3e-4),
    "etaT_bfbt_16x24": dict(
        fields=dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=16, Nz=24,
                    Lx=2.0, precond="bfbt"),
        ic=dict(kind="conductive", amp=3e-2), tstop=4e-4),
    "etaT_legacy_16x24": dict(
        fields=dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=16, Nz=24,
                    Lx=2.0),
        ic=dict(kind="conductive", amp=3e-2), tstop=2e-4),
    "arrh_bfbt_16x24": dict(
        fields=dict(viscosity_model="arrhenius", arrhenius_Ts=1.0,
                    deta_T=1e3, Ra=1e4, Nx=16, Nz=24, Lx=4.0,
                    internal_heating=1.5, dissipation_number=0.2,
                    eta_min=1e-2, eta_max=1e2, precond="bfbt"),
        ic=dict(kind="seeded", seed=2), nsteps=6, dt=1e-5),
    "etaT_bfbt_32x48_h": dict(
        fields=dict(viscosity_model="eta_T", deta_T=1e3, Ra=1.25e4, Nx=32,
                    Nz=48, Lx=4.0, internal_heating=1.5, precond="bfbt"),
        ic=dict(kind="seeded", seed=2), tstop=4e-4),
    # the heated 2:1 case at its real shape: 64x96, Lx = 4, Ra = 1.25e4,
    # H = 1.5, eta_T 1e3, conductive + seeded low modes; its first steps
    "c2_heated_64x96": dict(
        fields=dict(viscosity_model="eta_T", deta_T=1e3, Ra=1.25e4, Nx=64, Nz=96,
                    Lx=4.0, internal_heating=1.5, precond="bfbt"),
        ic=dict(kind="seeded", seed=2), tstop=3e-4),
    # 64x96 eta_T shape, the first steps of the benchmark run
    "c1_short_64x96": dict(
        fields=dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=64,
                    Nz=96, Lx=2.0, precond="bfbt"),
        ic=dict(kind="conductive", amp=3e-2), tstop=1.5e-4),
}

# the branches of the rescue, the eta window and guard, and the other initial
# condition: each case drives one decision of the reference that no trajectory
# above reaches, and records the counters behind it at every step
_ETAT16 = dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=16, Nz=24, Lx=2.0,
               precond="bfbt")
BRANCHES = {
    # one-iteration budget: every Stokes solve is rejected and rescued on the
    # lagged route at this eta (coupled_model.py:955-972 self-test [6])
    "rescue_maxiter1_16x24": dict(
        fields=dict(_ETAT16), model_kw=dict(stokes_maxiter=1),
        ic=dict(kind="conductive", amp=3e-2), tstop=2e-4),
    # a defect tolerance no correction meets: the defect fails, the factor is
    # refreshed and it fails again (double failure), then the true residual
    # decides (coupled_model.py:636-645)
    "rescue_defectfail_16x24": dict(
        fields=dict(_ETAT16, defect_tol=1e-30, max_K=1, K=1),
        model_kw=dict(stokes_maxiter=1),
        ic=dict(kind="conductive", amp=3e-2), tstop=2e-4),
    # an acceptance tolerance below roundoff: the lagged result misses it and
    # the dense direct solve replaces it (Nk*Nz = 216 <= 4096;
    # coupled_model.py:646-654)
    "rescue_direct_16x24": dict(
        fields=dict(_ETAT16), model_kw=dict(stokes_maxiter=3, stokes_rtol=1e-20),
        ic=dict(kind="conductive", amp=3e-2), nsteps=4, dt=2e-5),
    # the same above the direct cap (Nk*Nz = 65*64 = 4160 > 4096): the lagged
    # result is accepted and counted weak (coupled_model.py:646-648)
    "rescue_weak_128x64": dict(
        fields=dict(_ETAT16, Nx=128, Nz=64), model_kw=dict(stokes_maxiter=3,
                                                           stokes_rtol=1e-20),
        ic=dict(kind="conductive", amp=3e-2), nsteps=2, dt=2e-5),
    # the C^2 smooth clip of the eta window (eta_bounds.py:411-480) with the
    # window inside the law's range, so both bands are crossed
    "smooth_clip_16x24": dict(
        fields=dict(_ETAT16, eta_min=1e-2, eta_max=0.5, eta_smooth=0.6), model_kw={},
        ic=dict(kind="conductive", amp=3e-2), tstop=2e-4),
    # a heated state whose truncated eta undershoots through zero: the
    # positivity guard re-applies the window (coupled_model.py:451-496,
    # self-test [5] at 927-953).  Recorded as one eta_of and one Stokes solve,
    # not a march: the theta solve of this state ends its 200-iteration budget
    # at a relative residual of 0.5, which leaves the stepped state to
    # roundoff.  The Stokes operator is badly conditioned (its defect
    # corrections sit near 1e-8 from roundoff alone), so the lagged route with
    # a defect tolerance of 1e-6 keeps the solve's decisions far from their
    # thresholds
    "guard_h10_16x24": dict(
        fields=dict(_ETAT16, internal_heating=10.0, eta_min=1e-6, eta_max=1e6,
                    precond="legacy", defect_tol=1e-6),
        ic=dict(kind="guard", amp=None), probe=True),
    # the analytic boundary-layer start of bench_ic (bench_ic.py:149-163)
    "analytic_bl_32x48": dict(
        fields=dict(_ETAT16, Nx=32, Nz=48),
        ic=dict(kind="analytic_bl"), tstop=2e-4),
    # the 64x96 benchmark run at its step-300 state, inside the phase where
    # BFBT exhausts its budget and every solve is rescued
    "c1_s300_64x96": dict(
        fields=dict(_ETAT16, Nx=64, Nz=96),
        ic=dict(kind="state", file="bench_states_c1_64x96.npz", key="s300"),
        nsteps=2, dt=4e-6),
}


def run_probe(name, fields, ic, model_kw=None, probe=True):
    """One ``eta_of`` and one ``stokes`` of the initial state: the eta field
    on the Nx grid, psi and omega, the counters and the route."""
    api, Cfg, bench_ic = _ref()
    m = api.build_model(Cfg(**fields))
    H = float(fields.get("internal_heating", 0.0))
    th0 = guard_theta(m, H, ic["amp"])
    # the unguarded field from a second instance, so that the counters below
    # are those of exactly one eta_of and one stokes
    raw = api.build_model(Cfg(**fields))._eta_padded_truncated(th0)
    eta = m.eta_of(th0)
    psi, om = m.stokes(th0)
    counters = [[float(getattr(m, c, 0)) for c in COUNTERS]]
    meta = dict(name=name, fields=fields, ic=ic, probe=True, model_kw=model_kw or {},
                counter_names=list(COUNTERS), routes=[str(m.last_stokes_route)],
                raw_min=float(raw.min()), last_K=int(m.last_K),
                last_defect_res=float(m.last_defect_res),
                numpy=np.__version__, scipy=__import__("scipy").__version__)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), theta0=np.asarray(th0),
                        eta=np.asarray(eta), eta_raw=np.asarray(raw),
                        psi=np.asarray(psi), omega=np.asarray(om),
                        counters=np.asarray(counters),
                        meta=np.array(json.dumps(meta, default=float)))
    print("%-18s eta raw min %.3e -> %.3e, n_clip_eta_proj %d, route %s"
          % (name, raw.min(), eta.min(), m.n_clip_eta_proj, m.last_stokes_route),
          flush=True)


def guard_amp(fields):
    """The smallest amplitude of the self-test's scan (0.2 .. 1.2) whose
    truncated eta undershoots through zero (coupled_model.py:940-946)."""
    api, Cfg, _ = _ref()
    m = api.build_model(Cfg(**fields))
    H = float(fields["internal_heating"])
    for amp in (0.2, 0.4, 0.6, 0.8, 1.0, 1.2):
        raw = m._eta_padded_truncated(guard_theta(m, H, amp))
        if float(raw.min()) <= 0.0:
            return amp
    raise RuntimeError("no amplitude of the scan makes the guard fire")


LONG = {
    # the 64x96 eta_T run in full (CFL-limited: on the order of 10^3 steps)
    "c1_full_64x96": dict(
        fields=dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=64,
                    Nz=96, Lx=2.0, precond="bfbt"),
        ic=dict(kind="conductive", amp=3e-2), tstop=5e-3, keep_every=50),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*")
    ap.add_argument("--c1-full", action="store_true")
    ap.add_argument("--branches", action="store_true")
    a = ap.parse_args()
    cases = dict(LONG) if a.c1_full else dict(BRANCHES) if a.branches else dict(CASES)
    for name, spec in cases.items():
        if a.only and name not in a.only:
            continue
        if spec["ic"]["kind"] == "guard" and spec["ic"]["amp"] is None:
            spec = dict(spec, ic=dict(spec["ic"], amp=guard_amp(spec["fields"])))
        if spec.get("probe"):
            run_probe(name, **spec)
        else:
            run_case(name, **spec)


if __name__ == "__main__":
    main()
