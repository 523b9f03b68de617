"""Continuous batching of independent runs on one GPU.

A ``Server`` takes requests (a ``ConvectionConfig``, an initial state, a
``tstop`` and an arrival time), groups them by shape into buckets and steps
every bucket's active runs together.  Runs join a bucket in a free slot while
its other runs keep stepping, and leave it when their own ``t >= tstop``.  A
bucket is one ``Engine`` whose batch dimension is its slot count.  Admission
is by HBM footprint: a bucket is created only when its resident buffers fit
the memory budget, and idle buckets are released when a waiting request of
another shape needs the room.

Each run follows the march of api.py:169-190 with the production CFL policy
(cfl_policy.py:64-74), exactly as it would alone: every kernel is masked per
run and no arithmetic couples two runs, so a result does not depend on what
else was in the bucket (``tests/test_gpu_serve.py`` checks this against solo
marches).

Time-to-solution of a request is its finish time minus its arrival time
(queueing included); ``Server.stats()`` reports its p50 / p99 and the
aggregate run-steps/s.
"""
import queue
import threading
import time
from collections import OrderedDict, deque

import numpy as np
import torch

from . import _numerics as NUM
from .api import model_spec
from .batch import run_params, _KIND
from .core.cfl_policy import clamp_safety
from .engine import (ANTISYM_TOL, DENSE_FP64_MAXN, Engine, KIND_CONST, KIND_COUPLED,
                     LAG_STORE_BYTES, RESCUE_RESTART, antisym_part)


def run_bytes(kind, Nx, Nz, precision="fp64", stokes_restart=80, theta_restart=40):
    """Resident HBM bytes of one slot of an FP64 ``Engine`` of this shape
    (Krylov bases, per-mode blocks, product-grid fields, state, and for the
    coupled kind the slot's lagged dense factor: FP32 for ``precision``
    'mixed' or above the FP64 factor's size, else FP64), not counting the
    transient dense rescue matrices (bounded by ``engine.RESCUE_BYTES``)."""
    Nk, N = Nx // 2 + 1, Nz
    n, nrp, nru = Nk * N, Nx + 1, Nx // 2 + 1
    Q = (3 * N - 2 + 1) // 2 + 1
    w = 8
    b = (12 * n + 2 * nrp * N * 6 + 3 * nru * N + 8 * n) * w      # state, grids, scratch
    if kind == KIND_CONST:
        return b + (Nk * N * N + n + nrp * N) * w
    b += (2 * theta_restart + 6) * n * w + Nk * N * N * w          # theta FGMRES, Helmholtz
    b += (6 * N * N + Nk * N * N) * w                               # reduced-operator scratch
    if kind == KIND_COUPLED:
        b += (2 * stokes_restart + 6) * n * w                       # Stokes FGMRES
        b += Nk * N * N * w                                         # BFBT blocks
        b += (2 * Nk * 3 * Q + Nk * 7 * N + nrp * Q + 2 * nrp * N + 4 * n) * w
        f32 = precision == "mixed" or n > DENSE_FP64_MAXN
        b += n * n * (4 if f32 else 8) + n * 4                      # lagged factor
        if f32:
            b += (2 * RESCUE_RESTART + 6) * n * w                   # its FGMRES
    else:
        b += Nk * N * N * w                                         # eta(z) operator LU
    return int(b)


def lag_slots(Nx, Nz, precision="fp64"):
    """Most slots of a coupled bucket whose lagged dense factors (one n x n
    factor per slot, FP32 for ``precision`` 'mixed' or above the FP64
    factor's size) fit the engine's lagged-factor store: the legacy route
    needs them, and the mixed-precision rescue reuses and carries with them."""
    n = (Nx // 2 + 1) * Nz
    es = 4 if (precision == "mixed" or n > DENSE_FP64_MAXN) else 8
    return max(1, int(LAG_STORE_BYTES // (n * n * es)))


class Request:
    """One run to serve.  ``priority``: a request that finds its bucket full
    preempts a running request of lower priority there (the run is saved as
    a device snapshot, re-queued, and resumed from the snapshot in the next
    free slot)."""

    __slots__ = ("rid", "cfg", "theta0", "tstop", "arrival", "admitted", "finished",
                 "steps", "t", "theta", "diag", "bucket", "slot", "priority", "snapshot",
                 "preemptions")

    def __init__(self, rid, cfg, theta0, tstop, arrival=0.0, priority=0):
        self.rid, self.cfg = rid, cfg
        self.theta0 = np.asarray(theta0)
        self.tstop, self.arrival = float(tstop), float(arrival)
        self.priority = priority
        self.admitted = self.finished = None
        self.steps, self.t = 0, 0.0
        self.theta = self.diag = None
        self.bucket = self.slot = None
        self.snapshot = None
        self.preemptions = 0

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class Bucket:
    """Slots of one (kind, Nx, Nz, Lx) sharing an Engine."""

    def __init__(self, key, cfg, slots, precision="fp64", **engine_kw):
        kind, Nx, Nz, Lx = key
        self.key = key
        self.S = int(slots)
        params = [run_params(cfg)] * self.S
        self.eng = Engine(kind, Nx, Nz, Lx, params, precision=precision, **engine_kw)
        self.req = [None] * self.S
        self.t = np.zeros(self.S)
        self.tstop = np.zeros(self.S)
        self.live = np.zeros(self.S, dtype=bool)
        self.steps = 0
        self.last_used = 0.0
        # concurrent serving: the bucket's own stream, worker thread and inbox
        self.stream = None
        self.thread = None
        self.inbox = deque()
        self.stop = False
        self.error = None
        self.assigned = 0           # routed here and not yet collected (main thread only)

    def free_slots(self):
        return [i for i in range(self.S) if self.req[i] is None]

    def admit(self, slot, r, now):
        """Start ``r`` in ``slot``, or resume it from its snapshot."""
        if r.snapshot is not None:
            self.eng.restore(slot, r.snapshot)
            self.t[slot] = r.t
            r.snapshot = None
        else:
            self.eng.admit(slot, run_params(r.cfg), r.theta0)
            self.t[slot] = 0.0
            r.admitted = now
        self.req[slot] = r
        self.tstop[slot] = r.tstop
        self.live[slot] = self.t[slot] < r.tstop
        r.bucket, r.slot = self.key, slot

    def victim(self, r):
        """The slot of the lowest-priority running request below ``r``'s
        priority (None if there is none)."""
        best = None
        for i, q in enumerate(self.req):
            if q is not None and self.live[i] and q.priority < r.priority:
                if best is None or q.priority < self.req[best].priority:
                    best = i
        return best

    def preempt(self, slot):
        """Save the run in ``slot`` as a device snapshot and free the slot;
        -> its request (to be queued again)."""
        r = self.req[slot]
        r.snapshot = self.eng.snapshot(slot)
        r.t = float(self.t[slot])
        r.preemptions += 1
        r.bucket = r.slot = None
        self.req[slot] = None
        self.live[slot] = False
        return r

    def step(self, safety, dtmax):
        """One step of every live slot; -> run-steps taken."""
        live = self.live
        if not np.any(live):
            return 0
        act = None if np.all(live) else torch.as_tensor(live.astype(np.int32),
                                                        device=self.eng.dev)
        dt = self.eng.cfl(safety, dtmax, NUM.DTMIN, act)
        dt = np.minimum(dtmax, dt)
        self.eng.step(dt, act)
        self.t[live] += dt[live]
        for i in np.where(live)[0]:
            self.req[i].steps += 1
        self.steps += 1
        self.live = live & (self.t < self.tstop)
        return int(live.sum())

    def retire(self, now, diagnostics=True):
        """Release the slots whose run reached its tstop; -> their requests."""
        done = [i for i in range(self.S) if self.req[i] is not None and not self.live[i]]
        if not done:
            return []
        if diagnostics:
            mask = np.zeros(self.S, dtype=np.int32)
            mask[done] = 1
            d = self.eng.diagnostics(torch.as_tensor(mask, device=self.eng.dev))
        th = self.eng.get_theta()
        out = []
        names = ("Nu_top", "Nu_bot", "Nu_energy", "urms", "umax", "amp")
        for i in done:
            r = self.req[i]
            r.t = float(self.t[i])
            r.theta = th[i].copy()
            if diagnostics:
                r.diag = dict(zip(names, (float(v) for v in d[i])))
            r.finished = now
            self.req[i] = None
            out.append(r)
        return out

    def busy(self):
        return any(r is not None for r in self.req)

    def idle(self):
        """Nothing routed here is still unfinished."""
        return self.assigned == 0 and not self.busy()


class Server:
    """Admission, stepping and retirement of requests on one GPU.

    ``slots``     slots per bucket (at most; fewer when memory is short);
    ``mem_budget`` HBM bytes the buckets may hold (default: 80% of the free
                  device memory at construction, less the rescue headroom);
    ``precision`` "fp64" or "mixed" (see ``engine.PRECISIONS``);
    ``concurrent`` every bucket steps in a worker thread of its own, on its
                  own HIP stream, with no barrier between buckets: a bucket
                  of small runs is not held back by one of large runs.  The
                  worker admits requests from its inbox into free slots
                  between its steps; the main thread routes arrivals,
                  creates and releases buckets and collects retirements."""

    def __init__(self, slots=64, mem_budget=None, precision="fp64", safety=None, dtmax=None,
                 diagnostics=True, concurrent=False, **engine_kw):
        from .engine import RESCUE_BYTES
        self.slots = int(slots)
        if mem_budget is None:
            free, _ = torch.cuda.mem_get_info()
            mem_budget = max(0, int(0.8 * free) - RESCUE_BYTES)
        self.mem_budget = int(mem_budget)
        self.precision = precision
        self.engine_kw = dict(engine_kw)
        self.safety = clamp_safety(NUM.CFL_SAFETY if safety is None else safety)
        self.dtmax = NUM.DTMAX if dtmax is None else float(dtmax)
        self.diagnostics = bool(diagnostics)
        self.buckets = OrderedDict()
        self.bucket_bytes = {}
        self.pending = deque()
        self.done = []
        self.run_steps = 0
        self.t0 = None
        self.t_end = None
        self.concurrent = bool(concurrent)
        self._finished = queue.Queue()
        self._requeue = queue.Queue()
        self._steps_lock = threading.Lock()
        self.n_preempted = 0

    # ---------------------------------------------------------------- input
    def submit(self, requests):
        """Queue requests.  A variable-viscosity request must start from a
        mirror-symmetric state: its march takes the CFL dt of the state as
        given before the step projects it (convection_modes.py:198), and
        the device carries those models on the mirror subspace only."""
        for r in requests:
            if (self.key_of(r.cfg)[0] != KIND_CONST
                    and antisym_part(r.theta0) > ANTISYM_TOL):
                raise ValueError("request %r: a %s model needs a mirror-symmetric initial "
                                 "state (relative antisymmetric part %.2e)"
                                 % (r.rid, self.key_of(r.cfg)[0], antisym_part(r.theta0)))
        for r in sorted(requests, key=lambda q: q.arrival):
            self.pending.append(r)

    @staticmethod
    def key_of(cfg):
        spec = model_spec(cfg)
        return (_KIND[spec["model"]], int(cfg.Nx), int(cfg.Nz), float(cfg.Lx))

    def used_bytes(self):
        return sum(self.bucket_bytes.values())

    # ------------------------------------------------------------ admission
    def _bucket_for(self, r, now):
        key = self.key_of(r.cfg)
        b = self.buckets.get(key)
        if b is not None:
            if self.concurrent:
                return b            # its worker admits from the inbox as slots free
            return b if (b.free_slots() or b.victim(r) is not None) else None
        per = run_bytes(key[0], key[1], key[2], self.precision)
        room = self.mem_budget - self.used_bytes()
        if room < per:
            # release idle buckets of other shapes, oldest first
            for k in [k for k, bb in self.buckets.items() if bb.idle()]:
                self._release(k)
                room = self.mem_budget - self.used_bytes()
                if room >= self.slots * per:
                    break
        S = min(self.slots, room // per)
        if key[0] == KIND_COUPLED:
            # slots whose lagged factors fit the store (large grids only)
            S = min(S, lag_slots(key[1], key[2], self.precision))
        if S < 1:
            return None
        b = Bucket(key, r.cfg, S, precision=self.precision, **self.engine_kw)
        self.buckets[key] = b
        self.bucket_bytes[key] = S * per
        if self.concurrent:
            b.stream = torch.cuda.Stream()
            b.stream.wait_stream(torch.cuda.default_stream())
            b.thread = threading.Thread(target=self._worker, args=(b,), daemon=True)
            b.thread.start()
        return b

    def _release(self, key):
        b = self.buckets.pop(key)
        self.bucket_bytes.pop(key)
        if b.thread is not None:
            b.stop = True
            b.thread.join()
        del b
        torch.cuda.empty_cache()

    def _worker(self, b):
        """Bucket loop of concurrent serving: admit from the inbox, step the
        live slots, retire, until told to stop."""
        try:
            with torch.cuda.stream(b.stream):
                while not b.stop:
                    while b.inbox and b.free_slots():
                        b.admit(b.free_slots()[0], b.inbox.popleft(), self.now())
                    while b.inbox and not b.free_slots():
                        v = b.victim(b.inbox[0])
                        if v is None:
                            break
                        self._requeue.put((b, b.preempt(v)))
                        b.admit(v, b.inbox.popleft(), self.now())
                    if not np.any(b.live) and not b.busy():
                        time.sleep(0.001)
                        continue
                    n = b.step(self.safety, self.dtmax)
                    fin = b.retire(self.now(), self.diagnostics)
                    with self._steps_lock:
                        self.run_steps += n
                    for r in fin:
                        self._finished.put((b, r))
        except BaseException as exc:          # surfaced by the main loop
            b.error = exc

    def _admit(self, now):
        """FIFO over the arrived requests; a request that cannot be placed
        does not block later arrivals of other shapes."""
        waiting = deque()
        while self.pending and self.pending[0].arrival <= now:
            r = self.pending.popleft()
            b = self._bucket_for(r, now)
            if b is None:
                waiting.append(r)
                continue
            b.assigned += 1
            if self.concurrent:
                b.inbox.append(r)
                continue
            free = b.free_slots()
            if free:
                b.admit(free[0], r, now)
            else:
                # a full bucket and a running request of lower priority
                v = b.victim(r)
                q = b.preempt(v)
                b.assigned -= 1
                self.n_preempted += 1
                b.admit(v, r, now)
                waiting.append(q)
        while waiting:
            self.pending.appendleft(waiting.pop())

    # ----------------------------------------------------------------- loop
    def now(self):
        return time.perf_counter() - self.t0

    def run(self, max_wall=None, on_retire=None, progress=None, every=15.0):
        """Serve until every submitted request is finished (or ``max_wall``
        seconds have passed).  ``progress(stats)`` is called about every
        ``every`` seconds of serving."""
        if self.t0 is None:
            self.t0 = time.perf_counter()
        next_report = self.now() + every
        while self.pending or any(b.assigned > 0 for b in self.buckets.values()):
            if progress is not None and self.now() >= next_report:
                next_report = self.now() + every
                progress(dict(t=self.now(), done=len(self.done), pending=len(self.pending),
                              live=sum(int(b.live.sum()) for b in self.buckets.values()),
                              run_steps=self.run_steps, buckets=len(self.buckets)))
            now = self.now()
            if max_wall is not None and now > max_wall:
                break
            self._admit(now)
            if self.concurrent:
                self._collect(on_retire, now)
                continue
            stepped = 0
            busy = [b for b in self.buckets.values() if b.busy()]
            results = [(b, b.step(self.safety, self.dtmax),
                        b.retire(self.now(), self.diagnostics)) for b in busy]
            for b, n, fin in results:
                stepped += n
                b.assigned -= len(fin)
                if fin:
                    b.last_used = self.now()
                    self.done.extend(fin)
                    if on_retire is not None:
                        for r in fin:
                            on_retire(r)
            self.run_steps += stepped
            if stepped == 0 and self.pending:
                arrived = [r for r in self.pending if r.arrival <= now]
                if not arrived:       # idle until the next arrival
                    time.sleep(max(0.0, min(0.05, self.pending[0].arrival - self.now())))
                elif not any(b.busy() for b in self.buckets.values()):
                    raise RuntimeError("request %r does not fit the memory budget (%d bytes)"
                                       % (arrived[0].rid, self.mem_budget))
        if self.concurrent:
            for b in self.buckets.values():
                b.stop = True
            for b in self.buckets.values():
                if b.thread is not None:
                    b.thread.join()
            self._collect(on_retire, None, wait=False)
        torch.cuda.synchronize()
        self.t_end = self.now()
        return self.done

    def _collect(self, on_retire, now, wait=True):
        """Concurrent serving: gather retirements, surface worker errors,
        and wait briefly when nothing happened.  ``now`` is the time of the
        admission pass just made (a request arrived later is not late)."""
        got = False
        while True:
            try:
                b, r = self._requeue.get_nowait()
            except queue.Empty:
                break
            # preempted by the bucket's worker: queued again ahead of later
            # arrivals, resumed from its snapshot
            b.assigned -= 1
            self.n_preempted += 1
            self.pending.appendleft(r)
            got = True
        while True:
            try:
                b, r = self._finished.get(timeout=0.005 if (wait and not got) else 0.0)
            except queue.Empty:
                break
            got = True
            b.assigned -= 1
            b.last_used = self.now()
            self.done.append(r)
            if on_retire is not None:
                on_retire(r)
        for b in self.buckets.values():
            if b.error is not None:
                raise RuntimeError("bucket %r failed" % (b.key,)) from b.error
        if now is not None and not got and self.pending and not any(
                b.assigned > 0 for b in self.buckets.values()):
            arrived = [r for r in self.pending if r.arrival <= now]
            if arrived:
                raise RuntimeError("request %r does not fit the memory budget (%d bytes)"
                                   % (arrived[0].rid, self.mem_budget))

    def stats(self):
        lat = np.array([r.latency for r in self.done if r.latency is not None])
        wall = self.t_end if self.t_end is not None else self.now()
        return dict(requests=len(self.done), run_steps=int(self.run_steps), wall_s=wall,
                    preempted=self.n_preempted,
                    steps_per_s=self.run_steps / wall if wall > 0 else None,
                    p50_s=float(np.percentile(lat, 50)) if lat.size else None,
                    p99_s=float(np.percentile(lat, 99)) if lat.size else None,
                    buckets=[dict(key=list(k), slots=b.S, steps=b.steps)
                             for k, b in self.buckets.items()])
