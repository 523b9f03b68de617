"""Write the trajectory fixtures under tests/golden/ from the reference package.

Run in the development container only (the reference is not shipped anywhere):

    python tests/golden/make_golden.py [--only NAME ...] [--c1-full]

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


def _ref():
    sys.path.insert(0, REF)
    import specmc                       # noqa: F401
    from specmc import api, ConvectionConfig
    from specmc.core import bench_ic
    return api, ConvectionConfig, bench_ic


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
             keep_every=1):
    api, Cfg, bench_ic = _ref()
    cfg = Cfg(**fields)
    m = api.build_model(cfg)
    if ic["kind"] == "conductive":
        th0 = bench_ic.conductive_theta(m, amp=ic["amp"], m=ic.get("m", 1),
                                        n=ic.get("n", 1))
    else:
        th0 = seeded_theta(m, float(fields.get("internal_heating", 0.0)),
                           ic["seed"])
    states, dts, stokes_its, theta_its, ts = [], [], [], [], []
    last = {"t": 0.0}

    def on_step(n, t, s):
        dts.append(t - last["t"])
        last["t"] = t
        ts.append(t)
        stokes_its.append(int(getattr(m, "last_bfbt_iters", 0)))
        th = getattr(m, "theta", None)
        theta_its.append(int(th.last_iters) if th is not None else 0)
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
    np.savez_compressed(os.path.join(OUT, name + ".npz"),
                        theta0=np.asarray(th0), final=np.asarray(s),
                        states=np.asarray(states), dts=np.asarray(dts),
                        ts=np.asarray(ts), stokes_iters=np.asarray(stokes_its),
                        theta_iters=np.asarray(theta_its),
                        meta=np.array(json.dumps(meta, default=float)))
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
    # 64x96 eta_T shape, the first steps of the benchmark run
    "c1_short_64x96": dict(
        fields=dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=64,
                    Nz=96, Lx=2.0, precond="bfbt"),
        ic=dict(kind="conductive", amp=3e-2), tstop=1.5e-4),
}

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
    a = ap.parse_args()
    cases = dict(LONG) if a.c1_full else dict(CASES)
    for name, spec in cases.items():
        if a.only and name not in a.only:
            continue
        run_case(name, **spec)


if __name__ == "__main__":
    main()
