"""Host-side logic of serving and of the benchmark workloads (no device)."""
import numpy as np

from specmc_amd import workloads as W
from specmc_amd.engine import KIND_CONST, KIND_COUPLED, KIND_DEPTH
from specmc_amd.serve import Request, Server, lag_slots, run_bytes


def test_run_bytes_grows_with_shape_and_solver():
    for kind in (KIND_CONST, KIND_DEPTH, KIND_COUPLED):
        sizes = [run_bytes(kind, Nx, Nz) for Nx, Nz in ((16, 24), (64, 96), (256, 192))]
        assert sizes == sorted(sizes) and sizes[0] > 0
    for Nx, Nz in ((32, 24), (128, 96)):
        assert run_bytes(KIND_CONST, Nx, Nz) < run_bytes(KIND_DEPTH, Nx, Nz) \
            < run_bytes(KIND_COUPLED, Nx, Nz)
    # the Stokes Krylov bases dominate: (2 * 80 + 6) vectors of n doubles
    n = 33 * 96
    assert run_bytes(KIND_COUPLED, 64, 96) > 166 * n * 8


def test_bucket_keys():
    reqs = W.ensemble_requests(n=40)
    keys = {Server.key_of(r.cfg) for r in reqs}
    for k, Nx, Nz, Lx in keys:
        assert k in (KIND_DEPTH, KIND_COUPLED)
        assert (Nx, Nz) in ((32, 24), (64, 48), (64, 96), (128, 96)) and Lx in (2.0, 4.0)
    # eta_T and arrhenius share the coupled engine kind
    kinds = {(r.cfg.viscosity_model, Server.key_of(r.cfg)[0]) for r in reqs}
    assert ("arrhenius", KIND_COUPLED) in kinds and ("eta_T", KIND_COUPLED) in kinds


def test_workloads_are_seeded_and_well_formed():
    a, b = W.ensemble_requests(n=16), W.ensemble_requests(n=16)
    for p, q in zip(a, b):
        assert p.cfg == q.cfg and np.array_equal(p.theta0, q.theta0)
        assert p.theta0.shape == (p.cfg.Nx // 2 + 1, p.cfg.Nz)
        assert np.isrealobj(p.theta0) and p.arrival == 0.0
        if p.cfg.viscosity_model == "arrhenius":
            assert p.cfg.arrhenius_Ts == 1.0
            assert abs(p.cfg.arrhenius_E - 2.0 * np.log(p.cfg.deta_T)) < 1e-12
    s = W.stream_requests(n=32, rate=2.0)
    arr = np.array([r.arrival for r in s])
    assert np.all(np.diff(arr) > 0) and arr[0] > 0
    # mean inter-arrival near 1/rate
    assert 0.25 < np.mean(np.diff(np.r_[0.0, arr])) < 1.0
    for r in s:
        assert 1e-4 <= r.tstop <= 1e-2
        assert r.cfg.eta_min == 1e-5 and r.cfg.eta_max == 1.0
        assert 0.0 <= r.cfg.dissipation_number <= 0.5
    heated = W.heated_2to1_requests()[0]
    assert heated.cfg.Lx == 4.0 and heated.cfg.internal_heating == 1.5 and heated.tstop == 0.03
    # the ConvectionConfig default route unless another is asked for
    assert heated.cfg.precond == "legacy"
    assert W.heated_2to1_requests("bfbt")[0].cfg.precond == "bfbt"


def test_theta0_vanishes_at_the_walls():
    """The seeded perturbations are sin(n pi z) in z and the heated profile is
    z(1-z)/2: theta0 is zero on both walls."""
    r = W.ensemble_requests(n=3)[0]
    g = W._Grid(r.cfg.Nx, r.cfg.Nz, r.cfg.Lx)
    th = np.fft.irfft((g.V @ r.theta0.T).T * r.cfg.Nx, n=r.cfg.Nx, axis=0)
    assert np.max(np.abs(th[:, [0, -1]])) < 1e-12


def test_request_latency():
    r = Request("x", W.bench_64x96_config(), W.bench_64x96_theta0().real, 1e-3, arrival=2.0)
    assert r.latency is None
    r.finished = 3.5
    assert r.latency == 1.5


def test_lag_slots_cap_only_the_largest_coupled_grids():
    from specmc_amd.engine import LAG_STORE_BYTES
    # 256 x 192: one FP32 factor of n = 24768 is 2.45 GB, so 64 slots do not
    # fit the store; the cap keeps every slot's factor inside it
    n = 129 * 192
    S = lag_slots(256, 192, "mixed")
    assert S < 64 and S * n * n * 4 <= LAG_STORE_BYTES < (S + 1) * n * n * 4
    # n > 8192 takes the FP32 factor in either precision
    assert lag_slots(256, 192, "fp64") == S
    # every grid of the ensemble and at 64 x 96 keeps its 64 slots
    for Nx, Nz in ((32, 24), (64, 48), (64, 96), (128, 96), (128, 192)):
        for prec in ("fp64", "mixed"):
            assert lag_slots(Nx, Nz, prec) >= 64, (Nx, Nz, prec)


def test_rescue_gate_admits_transient_bytes_up_to_its_budget():
    """The transient dense matrices of rescues and refreshes are admitted by
    bytes: concurrent holders share the budget (small requests go ahead
    beside a large one), a request that would pass it waits for a release, a
    request above the whole budget goes ahead alone, and a thread's nested
    request (a rescue inside a refresh) never waits."""
    import threading
    import time
    from specmc_amd import engine
    gate = engine._ByteGate(100)

    def take(nbytes, out, hold_s=0.0):
        with gate.hold(nbytes):
            out.append((nbytes, time.perf_counter()))
            time.sleep(hold_s)

    got = []
    with gate.hold(60):
        t = threading.Thread(target=take, args=(30, got))     # fits beside 60
        t.start()
        t.join(timeout=5)
        assert not t.is_alive() and got and got[0][0] == 30
        got.clear()
        t = threading.Thread(target=take, args=(50, got))     # 60 + 50 > 100
        t.start()
        time.sleep(0.2)
        assert got == [] and t.is_alive()
        with gate.hold(500):                                 # nested: never waits
            assert gate.used == 560
        t0 = time.perf_counter()
    t.join(timeout=5)
    assert not t.is_alive() and got and got[0][1] >= t0
    assert gate.used == 0 and gate.held == {}
    got.clear()
    t = threading.Thread(target=take, args=(1000, got))       # above the budget, alone
    t.start()
    t.join(timeout=5)
    assert not t.is_alive() and got[0][0] == 1000 and gate.used == 0
    # many threads of mixed sizes finish, and the budget held at once stays within
    peak = [0]
    lock = threading.Lock()

    def work(nbytes):
        with gate.hold(nbytes):
            with lock:
                peak[0] = max(peak[0], gate.used)
            time.sleep(0.01)

    ts = [threading.Thread(target=work, args=(b,)) for b in (10, 90, 40, 70, 20, 60, 30, 80)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=10)
    assert all(not t.is_alive() for t in ts)
    assert gate.used == 0 and peak[0] <= 100
    assert isinstance(engine._RESCUE_GATE, engine._ByteGate)
    assert engine._RESCUE_GATE.budget == engine.RESCUE_BYTES


def test_rescue_gate_admits_waiting_requests_in_order():
    """A request above the budget that waits for the gate to empty is not
    overtaken by later small requests: they wait behind it, it runs alone,
    and they follow."""
    import threading
    import time
    from specmc_amd import engine
    gate = engine._ByteGate(100)
    order = []
    lock = threading.Lock()

    def take(name, nbytes, hold_s):
        with gate.hold(nbytes):
            with lock:
                order.append((name, gate.used))
            time.sleep(hold_s)

    small0 = threading.Thread(target=take, args=("small0", 10, 0.3))
    small0.start()
    time.sleep(0.05)
    big = threading.Thread(target=take, args=("big", 500, 0.1))
    big.start()
    time.sleep(0.05)
    later = [threading.Thread(target=take, args=("small%d" % i, 10, 0.01)) for i in (1, 2, 3)]
    for t in later:
        t.start()
        time.sleep(0.02)
    for t in [small0, big] + later:
        t.join(timeout=10)
    assert all(not t.is_alive() for t in [small0, big] + later)
    names = [n for n, _ in order]
    assert names[:2] == ["small0", "big"]
    assert dict(order)["big"] == 500            # alone
    assert sorted(names[2:]) == ["small1", "small2", "small3"]
    assert gate.used == 0 and not gate.queue
