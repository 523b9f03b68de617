"""Benchmark: batched steps/s at 64x96 eta_T (deta_T = 1e3) vs the plain FP64
single-run device stepper, with p50/p99 run latency.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]

Workload (config 1 of the scenario set): eta_T viscosity, deta_T = 1e3,
Ra = 1e4, Nx = 64, Nz = 96, Lx = 2 (the unit box on its mirror-doubled carrier),
BFBT Stokes route, theta0 = conductive + amp cos(pi x) sin(pi z).  A "step"
is one run-step (CFL dt + Stokes solve + BDF2 theta solve) of one run.

* plain   -- one run (amp = 3e-2) on the FP64 device stepper, unbatched;
* batched -- B concurrent runs (amp ~ U[2.5e-2, 3.5e-2], seeded) of the same
  shape advancing together, each with its own dt.

Prints one JSON line.  ``value`` is the batched aggregate run-steps/s.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOAD = dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=64, Nz=96, Lx=2.0,
                precond="bfbt")


def _configs(B, seed=1):
    from specmc_amd import ConvectionConfig
    rng = np.random.default_rng(seed)
    amps = rng.uniform(2.5e-2, 3.5e-2, size=B)
    return [ConvectionConfig(**WORKLOAD) for _ in range(B)], amps


def _ensemble(B, amps, prec=0):
    from specmc_amd.batch import Ensemble
    from specmc_amd.core import bench_ic
    from specmc_amd import ConvectionConfig
    from specmc_amd.basis import shape_ops
    cfgs = [ConvectionConfig(**WORKLOAD) for _ in range(B)]
    o = shape_ops(64, 96, 2.0)

    class _M:
        x, z, kx, Nx, V, Vi = o.x, o.z, o.kx, o.Nx, o.V, o.Vi

        def coeffs(self, g):
            A = np.fft.rfft(np.asarray(g), axis=0) / self.Nx
            return (self.Vi @ A.T).T

        def seed(self, amp=1e-3, m=1, n=1):
            X, Z = np.meshgrid(self.x, self.z, indexing="ij")
            return self.coeffs(amp * np.cos(m * np.pi * X) * np.sin(n * np.pi * Z))
    th = [bench_ic.conductive_theta(_M(), amp=float(a)) for a in amps]
    return Ensemble(cfgs, th, np.full(B, np.inf), prec=prec)


def _timed(ens, steps, warmup):
    import torch
    for _ in range(warmup):
        ens.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n = 0
    for _ in range(steps):
        n += ens.step()
    torch.cuda.synchronize()
    return n, time.perf_counter() - t0


def cpu_baseline(nsteps=4):
    """The oracle (numpy/scipy FP64 restatement) on the host cores: the first
    ``nsteps`` steps of the plain config-1 run."""
    import oracle
    m = oracle.build_oracle(WORKLOAD)
    s = m.seed(amp=3e-2)
    for _ in range(2):
        s = m.step(s, oracle.models.policy_dt(m, s))
    t0 = time.perf_counter()
    for _ in range(nsteps):
        s = m.step(s, oracle.models.policy_dt(m, s))
    el = time.perf_counter() - t0
    try:
        import threadpoolctl
        cores = max(int(p.get("num_threads", 1)) for p in threadpoolctl.threadpool_info())
    except Exception:
        cores = os.cpu_count() or 1
    return dict(value=nsteps / el, unit="steps/s", cores=cores, kind="port",
                sample="%d steps (after 2 warm-up steps) of the config-1 run "
                       "(64x96 eta_T deta=1e3, amp 3e-2) on the oracle" % nsteps)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: every run of this benchmark fits one MI355X "
              "and runs shard by run across GPUs without communication; run it with "
              "--gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    from specmc_amd import runtime
    runtime.require_device()
    # plain FP64 single-run stepper
    plain = _ensemble(1, np.array([3e-2]))
    n1, t1 = _timed(plain, a.plain_steps, a.warmup)
    plain_rate = n1 / t1
    del plain
    torch.cuda.empty_cache()
    # batched form
    _, amps = _configs(a.batch)
    ens = _ensemble(a.batch, amps)
    n, el = _timed(ens, a.steps, a.warmup)
    value = n / el
    out = dict(metric="batched steps/s at 64x96 eta_T deta=1e3 vs plain FP64 GPU",
               value=value, unit="run-steps/s", n_gpus=1, steps=a.steps,
               warmup=a.warmup, ms_per_step=1e3 * el / a.steps, higher_is_better=True,
               scaling="weak", vs_baseline=None, dtype="f64", data="synthetic",
               config=dict(workload="64x96 eta_T deta_T=1e3 Ra=1e4 bfbt, seeded amp",
                           batch=a.batch, Nx=64, Nz=96, precision="fp64"),
               plain_fp64_steps_per_s=plain_rate,
               speedup_vs_plain=value / plain_rate)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
