"""Benchmark: batched steps/s at 64x96 eta_T (deta_T = 1e3) vs the plain FP64
single-run device stepper, with p50/p99 run latency.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
    python bench.py --workload heated-2to1|ensemble-256|stream   # the other scenarios
    python bench.py --make-states            # record the 64x96 trajectory once

Workload (the headline): eta_T viscosity, deta_T = 1e3, Ra = 1e4, Nx = 64,
Nz = 96, Lx = 2 (the unit box on its mirror-doubled carrier), BFBT Stokes
route, theta0 = conductive + amp cos(pi x) sin(pi z), marched to tstop = 5e-3.
A "step" is one run-step (CFL dt + coupled Stokes solve + BDF2 theta solve).

``value`` -- the full-run leg: B requests with seeded amp ~ U[2.5e-2, 3.5e-2]
(one of them at 3e-2, the run the FP64 oracle's full march recorded) arrive
as a Poisson stream and are served in mixed precision by one bucket of B
slots, each from t = 0 to its tstop; value = sum of run-steps / wall time from
the first arrival to the last retirement, p50/p99 = time to solution.  It is
compared with ``plain_full`` -- one run (amp 3e-2) marched alone to tstop on
the FP64 device stepper, unbatched, with the same graph settings.

The cost of a step changes by ~50x along the run (the Stokes FGMRES needs ~10
iterations at first, then hits its 800-iteration budget and goes through the
dense rescue route, then settles near 400).  The phase legs break the rate
down: ``--make-states`` recorded phase checkpoints and the per-step log of the
whole run in ``bench_states_c1_64x96.npz``, and from each checkpoint
``--steps`` steps (after ``--warmup``) are timed of

* plain   -- one run on the FP64 device stepper, unbatched;
* batched -- B runs (the checkpoint with seeded 1e-3 relative perturbations)
  advancing together in FP64, each with its own dt, with the launch clock on
  the per-mode preconditioner apply, the Arnoldi step and the dense solves
  (the rooflines);
* mixed   -- the same B runs in mixed precision (FP32 dense rescue factor
  under FP64 FGMRES refinement, reused across steps; after a rejected BFBT
  solve a run goes straight to that factor for up to 16 solves, then probes
  BFBT again with a 100-iteration budget); its errors against the batched
  FP64 leg (same inputs) are reported;

weighted by each phase's share of the recorded run.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
STATES = os.path.join(ROOT, "bench_states_c1_64x96.npz")

WORKLOAD = dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=64, Nz=96, Lx=2.0,
                precond="bfbt")
TSTOP = 5e-3
CHECKPOINTS = (100, 300, 525)
# timed steps of the rescue phase at least: one carry period of the mixed
# rescue plus its BFBT attempt (17), which also holds one refresh of the reused
# factor at the default lag (ConvectionConfig.lag = 16)
HARD_MIN_STEPS = 17


def make_states(path=STATES):
    """March the plain FP64 stepper through the 64x96 eta_T run and record checkpoints."""
    from specmc_amd import ConvectionConfig, api
    from specmc_amd.core import bench_ic
    m = api.build_model(ConvectionConfig(**WORKLOAD))
    th0 = bench_ic.conductive_theta(m, amp=3e-2)
    log, saved = [], {}
    t0 = time.perf_counter()

    def cb(n, t, s):
        log.append((n, t, m.last_bfbt_iters, m.theta.last_iters,
                    m.last_stokes_route != "bfbt", time.perf_counter() - t0))
        if n in CHECKPOINTS:
            saved["s%d" % n] = np.array(s)
        if n % 50 == 0:
            print("step %d t=%.4e bfbt=%d theta=%d route=%s" % (
                n, t, m.last_bfbt_iters, m.theta.last_iters, m.last_stokes_route),
                flush=True)

    s, info = api.march_adaptive(m, th0, TSTOP, on_step=cb)
    wall = time.perf_counter() - t0
    d = m.diagnostics(s)
    np.savez_compressed(path, log=np.array(log), final=s, wall=wall,
                        diag=json.dumps(d), **saved)
    print(json.dumps(dict(steps=info["steps"], t=info["t"], wall=wall, diag=d)), flush=True)


# roofline kernels: the FP32-factor dense solve of the rescue route dominates
# the full runs (the value); the per-mode preconditioner apply of both FGMRES
# solves (k_inv_apply_pairs_pf) dominates the BFBT phases, and is the roofline
# when the full runs are not measured; the Arnoldi step is reported beside
ROOF_KERNEL = "pc_apply"
CLOCKED = ("pc_apply", "fg_post", "dense_solve")
PEAK_HBM = 8.0e12            # MI355X HBM3E, bytes/s
TRAFFIC_FILES = {
    "dense_solve_full": os.path.join(ROOT, "profiles", "ludm_solve_c1_b256_traffic.json"),
    "dense_solve_fixed": os.path.join(ROOT, "profiles", "ludm_solve_c1_b256_traffic.json"),
    "pc_apply": os.path.join(ROOT, "profiles", "inv_apply_c1_b256_traffic.json"),
    "fg_post": os.path.join(ROOT, "profiles", "fg_post_c1_b256_traffic.json"),
}
KERNEL_SYMBOLS = {"pc_apply": "k_inv_apply_pairs_pf", "fg_post": "k_fg_post",
                  "dense_solve": "k_ludm_solve"}
AMP_RANGE = (2.5e-2, 3.5e-2)   # seeded amplitudes of the full-run requests
ORACLE_FULL = os.path.join(ROOT, "tests", "golden", "oracle_c1_full_64x96.npz")


def legs_of(a):
    return a.legs.split(",")


def _tune(z, a, legs):
    """The autotuner's GEMM pass (autotune.tune_gemms) on a throwaway batch
    of each batch size the legs run, from the first checkpoint: one step is
    recorded, every GEMM shape it issued is timed under each kernel choice,
    and the fastest goes into the library's table for the timed legs."""
    import torch
    from specmc_amd import autotune
    sizes = sorted({1 if "plain" in legs else None, a.batch} - {None})
    s0 = z["s%d" % CHECKPOINTS[0]]
    rep = {}
    for nb in sizes:
        t0 = time.perf_counter()
        ens = _ensemble([s0] * nb, pc=a.pc, ortho=a.ortho)
        g = autotune.tune_gemms(ens.eng, ens.step)
        del ens
        torch.cuda.empty_cache()
        per_step = sum(v["us"] * v["calls_per_step"] for v in g.values())
        ref = sum(v["candidates"].get("mfma32x32", v["us"]) * v["calls_per_step"]
                  for v in g.values())
        rep[str(nb)] = dict(shapes=len(g), seconds=round(time.perf_counter() - t0, 2),
                            gemm_us_per_step=round(per_step, 1),
                            gemm_us_per_step_mfma32x32=round(ref, 1),
                            kernels=sorted({v["kernel"] for v in g.values()}))
    print("autotune: %s" % json.dumps(rep), file=sys.stderr, flush=True)
    return rep


def _ensemble(states, prec=0, pc=None, ortho=None, precision="fp64", graph_rounds=0):
    from specmc_amd import ConvectionConfig
    from specmc_amd.batch import Ensemble
    B = len(states)
    kw = dict(graph_rounds=graph_rounds)
    if pc is not None:
        kw["pc_apply"] = pc
    if ortho is not None:
        kw["ortho"] = ortho
    return Ensemble([ConvectionConfig(**WORKLOAD) for _ in range(B)], states,
                    np.full(B, np.inf), prec=prec, precision=precision, **kw)


def _diag_errors(ens, ref_states, ref_diag):
    """Relative errors of the states and of Nu_top / Vrms of ``ens`` against
    FP64 results of the same inputs: max over runs."""
    got = ens.states()
    d = ens.diagnostics()
    eth = max(float(np.linalg.norm(got[q] - ref_states[q]) / np.linalg.norm(ref_states[q]))
              for q in range(len(got)))
    enu = float(np.max(np.abs(d[:, 0] - ref_diag[:, 0]) / np.abs(ref_diag[:, 0])))
    evr = float(np.max(np.abs(d[:, 3] - ref_diag[:, 3]) / np.abs(ref_diag[:, 3])))
    return dict(theta_l2=eth, Nu=enu, Vrms=evr)


def unit_bytes(eng, kernel, factor_bytes=8):
    """Algorithmic HBM bytes per unit each clocked kernel counts.

    * fg_post: one vector of length n = Nk*Nz (an Arnoldi step j reads z, w
      and V_0..V_j and writes Z_j and V_{j+1}: j + 5 vectors; a restart reads
      the residual and writes V_0);
    * pc_apply: one system, the Nz x Nz inverse block read once, its
      right-hand side read and its solution written: (Nz^2 + 2 Nz) * 8;
    * dense_solve: one n x n factor read once (``factor_bytes`` per entry),
      its row order, and the FP64 right-hand side and solution."""
    n, N = eng.n, eng.Nz
    if kernel == "fg_post":
        return n * eng.esize
    if kernel == "pc_apply":
        return (N * N + 2 * N) * eng.esize
    return n * n * factor_bytes + n * 4 + 2 * n * 8


def _timed(ens, steps, warmup, clock=None):
    import torch
    for _ in range(warmup):
        ens.step()
    torch.cuda.synchronize()
    if clock is not None:
        clock.start()
    t0 = time.perf_counter()
    n = 0
    for _ in range(steps):
        n += ens.step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    if clock is not None:
        clock.stop()
    return n, el


def phase_weights(log):
    """Share of the recorded run's steps that each checkpoint represents
    (nearest checkpoint by step-cost class)."""
    it = log[:, 2]
    rescue = log[:, 4] > 0
    early = (~rescue) & (np.arange(len(log)) < np.argmax(rescue)) if rescue.any() else ~rescue
    hard = rescue | (it >= 700)
    late = ~(early | hard)
    w = np.array([early.sum(), hard.sum(), late.sum()], dtype=float)
    return w / w.sum()


def cpu_baseline(z, weights):
    """The oracle (numpy/scipy FP64 restatement of the reference's algorithm)
    on the host cores: one step (CFL dt with its Stokes solve, then the step)
    from each phase checkpoint, combined with the phase weights of the device
    legs into a whole-run rate."""
    import oracle
    per = {}
    for name, ck in zip(("early", "hard", "late"), CHECKPOINTS):
        m = oracle.build_oracle(WORKLOAD)
        s = np.array(z["s%d" % ck], dtype=complex)
        t0 = time.perf_counter()
        s = m.step(s, oracle.models.policy_dt(m, s))
        per[name] = 1.0 / (time.perf_counter() - t0)
        print("cpu baseline %s: %.3f steps/s (route %s)" % (name, per[name],
                                                            m.last_stokes_route),
              file=sys.stderr, flush=True)
    value = 1.0 / sum(wi / per[p] for wi, p in zip(weights, ("early", "hard", "late")))
    try:
        import threadpoolctl
        cores = max(int(p.get("num_threads", 1)) for p in threadpoolctl.threadpool_info())
    except Exception:
        cores = os.cpu_count() or 1
    return dict(value=value, unit="steps/s", cores=cores, kind="port", phases=per,
                sample="one step of the 64x96 eta_T run from each of its step-%d/%d/%d "
                       "checkpoints on the oracle, weighted like the device legs"
                       % CHECKPOINTS)


def _theta0(amp):
    from specmc_amd.workloads import bench_64x96_theta0
    return np.real(bench_64x96_theta0(amp=amp))


def _oracle_full():
    """The FP64 oracle's march of the amp = 3e-2 run to tstop (final state and
    diagnostics), if the fixture is present."""
    if not os.path.exists(ORACLE_FULL):
        return None
    with np.load(ORACLE_FULL, allow_pickle=False) as f:
        return np.real(f["final"]), json.loads(str(f["diag"])), json.loads(str(f["info"]))


def _vs_oracle(theta, diag, ref):
    if ref is None or theta is None or diag is None:
        return None
    th, d, _ = ref
    return dict(theta_l2=float(np.linalg.norm(theta - th) / np.linalg.norm(th)),
                Nu=abs(diag["Nu_top"] - d["Nu_top"]) / abs(d["Nu_top"]),
                Nu_energy=abs(diag["Nu_energy"] - d["Nu_energy"]) / abs(d["Nu_energy"]),
                Vrms=abs(diag["urms"] - d["urms"]) / abs(d["urms"]))


def full_leg(B, precision, graph_rounds, arrive_s, seed=8, clock=None):
    """B requests of the 64x96 run with seeded amplitudes, Poisson arrivals
    over about ``arrive_s`` seconds, each marched from t = 0 to tstop by one
    bucket of B slots (continuous admission and retirement).  A warm-up of B
    two-step requests first sets kernel attributes, workspaces and graph
    captures up; the timed window runs from the first arrival to the last
    retirement."""
    from specmc_amd import ConvectionConfig
    from specmc_amd.serve import Request, Server
    from specmc_amd import _numerics as NUM
    rng = np.random.default_rng(seed)
    amps = rng.uniform(AMP_RANGE[0], AMP_RANGE[1], B)
    amps[0] = 3e-2
    cfg = ConvectionConfig(**WORKLOAD)
    srv = Server(slots=B, precision=precision, diagnostics=True, graph_rounds=graph_rounds)
    srv.submit([Request("warm%d" % i, cfg, _theta0(amps[i]), 2 * NUM.DTMAX, 0.0)
                for i in range(B)])
    srv.run()
    nwarm = len(srv.done)
    reqs = [Request(i, cfg, _theta0(amps[i]), TSTOP) for i in range(B)]
    if clock is not None:
        # the dense solves of the rescue run eagerly (outside every graph):
        # their launch clock leaves the graph-replayed rounds as they are
        clock.start()
    t_first = srv.now() + 0.05
    arr = t_first + np.concatenate([[0.0], np.cumsum(rng.exponential(arrive_s / B, B - 1))])
    for r, a in zip(reqs, arr):
        r.arrival = float(a)
    srv.submit(reqs)
    srv.run(progress=lambda p: print("full runs: %s" % json.dumps(p), file=sys.stderr,
                                     flush=True))
    if clock is not None:
        clock.stop()
    done = srv.done[nwarm:]
    wall = max(r.finished for r in done) - t_first
    # every slot held one timed request from its admission to the end, so the
    # engine's per-slot counters are those of the timed requests
    eng = next(iter(srv.buckets.values())).eng
    counters = dict(bfbt_iters=int(eng.bfbt_iters.sum()),
                    bfbt_rejections=int(eng.n_rescue.sum()),
                    rescues_carried=int(eng.n_rescue_carry.sum()),
                    bfbt_probes=int(eng.n_bfbt_probe.sum()),
                    rescue_fgmres_iters=int(eng.rescue_iters.sum()),
                    factor_refreshes=int(eng.n_refresh.sum()),
                    refresh_batches=int(eng.n_refresh_batches),
                    rescues_fp64_redone=int(eng.n_rescue_fp32_redo.sum()))
    steps = int(sum(r.steps for r in done))
    lat = np.array([r.latency for r in done])
    r0 = next(r for r in done if r.rid == 0)
    return dict(requests=len(done), run_steps=steps, wall_s=wall, steps_per_s=steps / wall,
                p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                steps_per_run=[int(min(r.steps for r in done)),
                               int(max(r.steps for r in done))],
                arrival_window_s=float(arr[-1] - arr[0]), amp_range=list(AMP_RANGE),
                slots=B, precision=precision, graph_rounds=graph_rounds, counters=counters,
                amp3e2_vs_fp64_oracle=_vs_oracle(r0.theta, r0.diag, _oracle_full()))


DENSE_ACTIVE = 71   # matrices per launch of the full runs' FP32-factor solves (their mean)


def dense_leg(nmat, active=DENSE_ACTIVE, launches=20, clock=None):
    """The FP32-factor dense solve alone (smx_dense_lu_solve_mw prec 2) on
    ``nmat`` factors of the 64x96 Schur size (n = 3168, random matrices:
    bytes and time do not depend on the values), ``active`` of them solved per
    launch (a seeded draw), ``launches`` launches; with ``clock`` armed, each
    launch is timed.  The kernel's traffic is measured on this leg."""
    import torch
    from specmc_amd import runtime
    L = runtime.load()
    n = (WORKLOAD["Nx"] // 2 + 1) * WORKLOAD["Nz"]
    st = torch.cuda.current_stream().cuda_stream
    g = torch.Generator(device="cuda").manual_seed(7)
    A = torch.randn(nmat, n, n, device="cuda", dtype=torch.float32, generator=g)
    A.diagonal(dim1=1, dim2=2).add_(4.0 * n ** 0.5)
    piv = torch.zeros(nmat, n, dtype=torch.int32, device="cuda")
    info = torch.zeros(nmat, dtype=torch.int32, device="cuda")
    L.dense_lu(1, nmat, n, A.data_ptr(), n * n, piv.data_ptr(), info.data_ptr(), st)
    perm = torch.zeros(nmat, n, dtype=torch.int32, device="cuda")
    L.dense_lu_perm(nmat, n, piv.data_ptr(), perm.data_ptr(), None, st)
    b = torch.randn(nmat, n, device="cuda", dtype=torch.float64, generator=g)
    x = torch.zeros_like(b)
    ws = torch.zeros(L.dense_lu_solve_mw_bytes(2, nmat, n), dtype=torch.uint8, device="cuda")
    rng = np.random.default_rng(8)
    m = np.zeros(nmat, dtype=np.int32)
    m[rng.choice(nmat, size=min(active, nmat), replace=False)] = 1
    mask = torch.as_tensor(m, device="cuda")

    def run():
        L.dense_lu_solve_mw(2, nmat, n, A.data_ptr(), n * n, perm.data_ptr(), b.data_ptr(),
                            x.data_ptr(), mask.data_ptr(), ws.data_ptr(), ws.numel(), st)
    run()
    torch.cuda.synchronize()
    if clock is not None:
        clock.start()
    t0 = time.perf_counter()
    for _ in range(launches):
        run()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if clock is not None:
        clock.stop()
    err = int(ws[8:12].view(torch.int32).cpu()[0])
    out = dict(nmat=nmat, active=int(m.sum()), launches=launches, n=n,
               us_per_launch=1e6 * wall / launches, error_word=err)
    del A, piv, perm, b, x, ws
    torch.cuda.empty_cache()
    return out


def plain_full_leg(graph_rounds):
    """One run (amp 3e-2) alone on the FP64 device stepper, unbatched, from
    t = 0 to tstop (after a two-step warm-up of the same model)."""
    import torch
    from specmc_amd import ConvectionConfig, api
    m = api.build_model(ConvectionConfig(**WORKLOAD))
    m.eng.graph_rounds = int(graph_rounds)
    th0 = _theta0(3e-2)
    api.march(m, th0, 4e-5, 2)
    m = api.build_model(ConvectionConfig(**WORKLOAD))
    m.eng.graph_rounds = int(graph_rounds)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s, info = api.march_adaptive(m, th0, TSTOP)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return dict(steps=info["steps"], wall_s=wall, steps_per_s=info["steps"] / wall,
                graph_rounds=graph_rounds,
                vs_fp64_oracle=_vs_oracle(s, m.diagnostics(s), _oracle_full()))


class _WallLimit(Exception):
    pass


def run_heated(precisions=("fp64", "mixed"), max_wall=None, routes=("legacy",)):
    """The heated 2:1 case marched to t = 0.03 with the single-run model on
    each precision and Stokes route, diagnostics at the first steps with
    t >= 0.01, 0.02, 0.03 (stopped early, and reported as such, past
    ``max_wall`` seconds per run)."""
    from specmc_amd import api
    from specmc_amd.workloads import heated_2to1_requests, HEATED_LOG_TIMES
    import threading
    import torch
    reqs = {route: heated_2to1_requests(route)[0] for route in routes}
    out, errors = {}, []
    combos = [(p, rt) for rt in routes for p in precisions]

    def one(prec, route):
        try:
            with torch.cuda.stream(torch.cuda.Stream()):
                march_one(prec, route)
                torch.cuda.current_stream().synchronize()
        except BaseException as exc:
            errors.append(exc)

    def march_one(prec, route):
        r = reqs[route]
        label = "%s/%s" % (prec, route)
        m = api.build_model(r.cfg, precision=prec)
        logs, marks = [], list(HEATED_LOG_TIMES)
        t0 = time.perf_counter()
        last = {}

        def cb(n, t, s):
            last.update(n=n, t_prev=last.get("t", 0.0), t=t, s=s)
            if max_wall is not None and time.perf_counter() - t0 > max_wall:
                raise _WallLimit()
            if n % 25 == 0:
                dsc = m.describe()
                th = m.theta
                print("heated-2to1 %s: step %d t=%.5f dt=%.3e wall %.1fs | theta its %d info %d "
                      "nonconv %d | stokes %s bfbt its %d res %.1e rescue res %.1e" % (
                          label, n, t, t - last.get("t_prev", 0.0), time.perf_counter() - t0,
                          th.last_iters, th.last_info, th.n_nonconv,
                          dsc.get("last_stokes_route"), dsc.get("last_bfbt_iters", -1),
                          dsc.get("last_bfbt_true_res", float("nan")),
                          dsc.get("last_defect_res", float("nan"))),
                      file=sys.stderr, flush=True)
            if marks and t >= marks[0] - 1e-15:
                d = m.diagnostics(s)
                logs.append(dict(t_mark=marks.pop(0), step=n, t=t,
                                 Nu_top=d["Nu_top"], Nu_energy=d["Nu_energy"], urms=d["urms"],
                                 wall_s=time.perf_counter() - t0))
                print("heated-2to1 %s: %s" % (label, json.dumps(logs[-1])), file=sys.stderr,
                      flush=True)

        try:
            s, info = api.march_adaptive(m, r.theta0, r.tstop, on_step=cb)
            complete = True
        except _WallLimit:
            s, info, complete = last["s"], dict(steps=last["n"], t=last["t"]), False
        wall = time.perf_counter() - t0
        dsc = m.describe()
        out[label] = dict(steps=info["steps"], t=info["t"], complete=complete, wall_s=wall,
                         steps_per_s=info["steps"] / wall,
                         log=logs, rescues=dsc.get("n_bfbt_rescue"),
                         rescues_fp32=dsc.get("n_rescue_fp32"),
                         rescues_fp32_redone=dsc.get("n_rescue_fp32_redo"),
                         rescues_carried=dsc.get("n_rescue_carry"),
                         bfbt_probes=dsc.get("n_bfbt_probe"),
                         theta_nonconv=m.theta.n_nonconv,
                         last_rescue_res=dsc.get("last_defect_res"),
                         dt_last=(info["t"] - last.get("t_prev", 0.0)) if last else None,
                         theta=np.real(s))
    # the runs march concurrently, each on its own stream
    threads = [threading.Thread(target=one, args=c) for c in combos]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    res = dict(metric="heated 2:1 run (64x96, Ra=1.25e4, H=1.5, eta_T 1e3) to t=0.03",
               unit="steps/s", config=dict(workload="heated-2to1", Nx=64, Nz=96, Lx=4.0))
    for key in out:
        o = dict(out[key])
        o.pop("theta")
        res[key] = o
    for route in routes:
        a, b = out.get("fp64/" + route), out.get("mixed/" + route)
        if a and b and a["log"] and b["log"]:
            e = {k: max(abs(lb[k] - la[k]) / abs(la[k]) for la, lb in zip(a["log"], b["log"]))
                 for k in ("Nu_top", "Nu_energy", "urms")}
            if a["complete"] and b["complete"]:
                e["theta_l2"] = float(np.linalg.norm(b["theta"] - a["theta"])
                                      / np.linalg.norm(a["theta"]))
            res["mixed_vs_fp64/" + route] = e
    return res


def run_served(reqs, slots, precision, graph_rounds, max_wall, label, concurrent=True):
    from specmc_amd.serve import Server
    srv = Server(slots=slots, precision=precision, graph_rounds=graph_rounds,
                 concurrent=concurrent)
    srv.submit(reqs)
    srv.run(max_wall=max_wall,
            progress=lambda p: print("%s: %s" % (label, json.dumps(p)), file=sys.stderr,
                                     flush=True))
    st = srv.stats()
    st.update(workload=label, precision=precision, concurrent=concurrent, submitted=len(reqs),
              unfinished=len(reqs) - st["requests"])
    return st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--phases", default="early,hard,late")
    ap.add_argument("--legs", default="full,plainfull,plain,batched,mixed,dense",
                    help="full: B served full runs in mixed precision (the reported value); "
                         "plainfull: one FP64 run to tstop; plain / batched / mixed: the "
                         "phase legs (one FP64 run, B FP64 runs, B mixed runs); dense: the "
                         "FP32-factor dense solve alone on B factors at the full runs' "
                         "mean active count")
    ap.add_argument("--pc", default=None, choices=("lu", "inv"),
                    help="preconditioner block application (default: the engine's)")
    ap.add_argument("--ortho", default=None, choices=("mgs", "bgs2"),
                    help="FGMRES orthogonalisation (default: the engine's)")
    ap.add_argument("--graph-rounds", type=int, default=4,
                    help="FGMRES rounds per HIP graph replay in every leg but the clocked "
                         "batched FP64 one (0: eager launches)")
    ap.add_argument("--arrive-s", type=float, default=4.0,
                    help="full runs: the Poisson arrivals span about this many seconds")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the GEMM autotuner pass before the timed legs")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="copy MIB MiB device-to-device once first (a known-size "
                         "reference for hardware byte counters)")
    ap.add_argument("--workload", default="etaT-64x96",
                    choices=("etaT-64x96", "heated-2to1", "ensemble-256", "stream"),
                    help="etaT-64x96: the batched 64x96 eta_T headline (default); heated-2to1: "
                         "the heated 2:1 run to t=0.03; ensemble-256: 256-run uneven-grid "
                         "ensemble; stream: Poisson stream of requests up to 256x192")
    ap.add_argument("--precision", default="mixed", choices=("fp64", "mixed"),
                    help="precision of the ensemble / stream server")
    ap.add_argument("--slots", type=int, default=64, help="slots per bucket (ensemble, stream)")
    ap.add_argument("--requests", type=int, default=0,
                    help="ensemble-256: runs (default 256); stream: requests (default 48)")
    ap.add_argument("--rate", type=float, default=0.5, help="stream arrivals per second")
    ap.add_argument("--tstop", type=float, default=1e-3, help="ensemble-256 run length")
    ap.add_argument("--max-wall", type=float, default=600.0, help="ensemble / stream wall limit (s)")
    ap.add_argument("--sequential", action="store_true",
                    help="ensemble / stream: step the buckets one after another on one stream")
    ap.add_argument("--heated-precision", default="fp64,mixed")
    ap.add_argument("--heated-route", default="legacy",
                    help="heated-2to1: Stokes routes, comma-separated (legacy, bfbt)")
    ap.add_argument("--make-states", action="store_true")
    ap.add_argument("--states-out", default=STATES)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the runs of this benchmark are independent "
              "and shard by run across GPUs without communication; run it with "
              "--gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    from specmc_amd import runtime
    runtime.require_device()
    if a.make_states:
        make_states(a.states_out)
        return
    if a.workload != "etaT-64x96":
        from specmc_amd import workloads as W
        if a.workload == "heated-2to1":
            out = run_heated(tuple(a.heated_precision.split(",")), a.max_wall,
                             tuple(a.heated_route.split(",")))
        elif a.workload == "ensemble-256":
            out = run_served(W.ensemble_requests(n=a.requests or 256, tstop=a.tstop),
                             a.slots, a.precision, a.graph_rounds, a.max_wall,
                             "ensemble-256: 256-run uneven-grid ensemble, tstop=%g" % a.tstop,
                             not a.sequential)
        else:
            out = run_served(W.stream_requests(n=a.requests or 48, rate=a.rate),
                             a.slots, a.precision, a.graph_rounds, a.max_wall,
                             "stream: Poisson stream at %g/s" % a.rate, not a.sequential)
        print(json.dumps(out), flush=True)
        return
    if not os.path.exists(STATES):
        raise SystemExit("missing %s: run `python bench.py --make-states` first" % STATES)
    if a.copy_probe > 0:
        x = torch.ones(a.copy_probe << 17, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y
    z = np.load(STATES, allow_pickle=False)
    log = z["log"]
    w = phase_weights(log)
    names = ("early", "hard", "late")
    legs = legs_of(a)
    tuning = None if a.no_tune else _tune(z, a, legs)
    from specmc_amd.kclock import KernelClock
    rng = np.random.default_rng(1)
    clock = KernelClock(CLOCKED, capacity=40000)
    clock32 = KernelClock("dense_solve", capacity=4000)
    ub = {}
    res = {}
    for ph, step in zip(names, CHECKPOINTS):
        if ph not in a.phases.split(","):
            continue
        s0 = z["s%d" % step]
        sts = [s0 * (1.0 + 1e-3 * rng.normal()) for _ in range(a.batch)]
        # the rescue phase times at least one lag period, so that every run's
        # lagged-factor refresh falls inside the window at its real rate
        ns = max(a.steps, HARD_MIN_STEPS) if ph == "hard" else a.steps
        r = {"steps": ns}
        if "plain" in legs:
            plain = _ensemble([s0], pc=a.pc, ortho=a.ortho, graph_rounds=a.graph_rounds)
            n1, t1 = _timed(plain, ns, a.warmup)
            r["plain"] = n1 / t1
            del plain
            torch.cuda.empty_cache()
        ref = None
        if "batched" in legs:
            ens = _ensemble(sts, pc=a.pc, ortho=a.ortho)
            n2, t2 = _timed(ens, ns, a.warmup, clock)
            for k in CLOCKED:
                ub[k] = unit_bytes(ens.eng, k)
            r.update(batched=n2 / t2, ms_batched_step=1e3 * t2 / ns,
                     bfbt_iters=int(ens.eng.last_bfbt_iters[0]),
                     rescues=int(ens.eng.n_rescue.sum()), pc_apply=ens.eng.pc_apply,
                     ortho=ens.eng.ortho)
            if "mixed" in legs:
                ref = (ens.states(), ens.diagnostics())
            del ens
            torch.cuda.empty_cache()
        if "mixed" in legs:
            ens = _ensemble(sts, pc=a.pc, ortho=a.ortho, precision="mixed",
                            graph_rounds=a.graph_rounds)
            n3, t3 = _timed(ens, ns, a.warmup)
            r.update(mixed=n3 / t3, ms_mixed_step=1e3 * t3 / ns,
                     lag_refreshes=int(ens.eng.n_refresh.sum()),
                     rescues_fp32=int(ens.eng.n_rescue_fp32.sum()),
                     rescues_fp32_redone=int(ens.eng.n_rescue_fp32_redo.sum()),
                     rescue_fgmres_iters=int(ens.eng.rescue_iters.sum()),
                     rescues_carried=int(ens.eng.n_rescue_carry.sum()),
                     bfbt_probes=int(ens.eng.n_bfbt_probe.sum()),
                     bfbt_rejections=int(ens.eng.n_rescue.sum()))
            if ref is not None:
                r["mixed_vs_fp64"] = _diag_errors(ens, *ref)
            if ph == "hard":
                # two more steps with the clock on the FP32-factor solves of
                # the carried rescues (eager launches; not part of the rate)
                _timed(ens, 2, 0, clock32)
                ub["dense_solve_fp32"] = unit_bytes(ens.eng, "dense_solve", factor_bytes=4)
            del ens
            torch.cuda.empty_cache()
        res[ph] = r
        print("phase %s: %s" % (ph, json.dumps(r)), file=sys.stderr, flush=True)
    ws = np.array([w[names.index(p)] for p in res])
    ws = ws / ws.sum() if len(ws) else ws
    # whole-run rates: time per step is the weighted harmonic combination
    rate = lambda leg: (1.0 / sum(wi / res[p][leg] for wi, p in zip(ws, res))
                        if res and all(leg in res[p] for p in res) else None)
    plain_rate, fp64_rate, mixed_rate = rate("plain"), rate("batched"), rate("mixed")
    full = plainfull = None
    clock_full = KernelClock("dense_solve", capacity=40000)
    if "full" in legs:
        full = full_leg(a.batch, "mixed", a.graph_rounds, a.arrive_s, clock=clock_full)
        print("full runs: %s" % json.dumps(full), file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    dense = None
    clock_dense = KernelClock("dense_solve", capacity=1000)
    if "dense" in legs:
        dense = dense_leg(a.batch, clock=clock_dense)
        print("dense solve: %s" % json.dumps(dense), file=sys.stderr, flush=True)
    if "plainfull" in legs:
        plainfull = plain_full_leg(a.graph_rounds)
        print("plain full run: %s" % json.dumps(plainfull), file=sys.stderr, flush=True)
    base = plainfull["steps_per_s"] if plainfull else plain_rate
    value = next(v for v in ((full or {}).get("steps_per_s"), mixed_rate, fp64_rate,
                             plain_rate, 0.0) if v is not None)
    dtype = "f64+f32" if (full or mixed_rate is not None) else "f64"
    out = dict(metric="batched steps/s at 64x96 eta_T deta=1e3 vs plain FP64 GPU; "
                      "p50/p99 run latency",
               value=value, unit="run-steps/s", n_gpus=1, steps=a.steps,
               warmup=a.warmup, ms_per_step=(1e3 * a.batch / value if value else None),
               higher_is_better=True, scaling="weak", vs_baseline=None, dtype=dtype,
               data="synthetic",
               config=dict(workload="64x96 eta_T deta_T=1e3 Ra=1e4 unit box (Lx=2 carrier), "
                                    "BFBT route, theta0 conductive + amp cos(pi x) sin(pi z), "
                                    "amp ~ U[2.5e-2, 3.5e-2], t=0 to tstop=5e-3",
                           value=("served full runs: sum of run-steps / wall over B "
                                  "Poisson-arriving requests, t=0 to tstop, mixed "
                                  "precision; --steps/--warmup set the phase legs"
                                  if full else "phase-weighted rate of the phase legs"),
                           batch=a.batch, Nx=64, Nz=96,
                           precision=("mixed: FP64 Krylov/operators/blocks; FP32 dense "
                                      "rescue LU under FP64 FGMRES refinement, reused "
                                      "across steps; rescued runs carried past BFBT "
                                      "with a 100-iteration probe every 17th solve"
                                      if dtype != "f64" else "fp64"),
                           phase_weights=dict(zip(names, [float(x) for x in w]))),
               plain_fp64_steps_per_s=base,
               speedup_vs_plain=(value / base if base else None),
               full_runs=full, plain_full=plainfull, dense_solve=dense,
               phase_legs=dict(plain=plain_rate, batched_fp64=fp64_rate,
                               batched_mixed=mixed_rate,
                               mixed_vs_plain=(mixed_rate / plain_rate
                                               if mixed_rate and plain_rate else None)),
               phases=res, autotune=tuning)
    if full:
        out["p50_s"], out["p99_s"] = full["p50_s"], full["p99_s"]
    errs = [res[p]["mixed_vs_fp64"] for p in res if "mixed_vs_fp64" in res[p]]
    if errs:
        out["mixed_error_vs_fp64"] = {k: max(e[k] for e in errs) for k in errs[0]}
    roofs = {}
    if full or dense:
        import types
        shape = types.SimpleNamespace(n=(WORKLOAD["Nx"] // 2 + 1) * WORKLOAD["Nz"],
                                      Nz=WORKLOAD["Nz"], esize=8)
        ub["dense_solve_full"] = ub["dense_solve_fixed"] = unit_bytes(shape, "dense_solve",
                                                                      factor_bytes=4)
    for k, clk, key in (("dense_solve_full", clock_full, "dense_solve"),
                        ("dense_solve_fixed", clock_dense, "dense_solve"),
                        ("pc_apply", clock, "pc_apply"), ("fg_post", clock, "fg_post"),
                        ("dense_solve_fp64", clock, "dense_solve"),
                        ("dense_solve_fp32", clock32, "dense_solve")):
        nl, ms, units = clk.summary(key)
        if not nl or k not in ub and key not in ub:
            continue
        b_unit = ub.get(k, ub.get(key))
        achieved = b_unit * units / nl / (ms * 1e-3)
        traffic = None
        tf = TRAFFIC_FILES.get(k)
        if tf and os.path.exists(tf):
            with open(tf) as f:
                tj = json.load(f)
            # a traffic file measured per unit scales to this leg's units per launch
            traffic = (tj["hbm_bytes_per_unit"] * units / nl if "hbm_bytes_per_unit" in tj
                       else tj.get("hbm_bytes_per_launch"))
        roofs[k] = dict(bound="hbm", achieved=achieved / 1e9, peak=PEAK_HBM / 1e9,
                        unit="GB/s", frac=achieved / PEAK_HBM, traffic=traffic,
                        kernel=KERNEL_SYMBOLS[key], launches=nl, mean_launch_us=1e3 * ms,
                        bytes_per_unit=b_unit, units_per_launch=units / nl)
    # the roofline is the dominant kernel of the value leg: the FP32-factor
    # dense solve of the rescue route in the full runs (the preconditioner
    # apply when the full runs are not measured)
    order = ("dense_solve_full", ROOF_KERNEL, "dense_solve_fixed")
    top = next((k for k in order if k in roofs), next(iter(roofs), None))
    if top is not None:
        out["roofline"] = roofs.pop(top)
        out["roofline"]["leg"] = top
        out["rooflines_other"] = roofs
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(z, w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
