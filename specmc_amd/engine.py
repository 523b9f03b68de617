"""Batched device engine: B runs of one ``(Nx, Nz, Lx)`` shape advance together.

Every numerical operation of a step is a kernel of ``libspecmc_hip.so``; this
module only sequences launches, owns the HBM buffers (torch tensors used as
plain device memory) and reads back the few per-run scalars that decide
control flow (FGMRES completion, CFL dt, memo hits).

Real arithmetic on the mirror-symmetric subspace
    With the box symmetry the theta coefficients are real and the stream
    function, vorticity and Stokes unknowns purely imaginary
    (stokes_coupled.py:78-112): ``b = i k Ra theta`` with real operators keeps
    the whole Stokes Krylov process in ``i R^n``.  Stream-function-type
    vectors are therefore stored by their imaginary part (the "s-part") and
    every FGMRES inner product equals the reference's ``vdot`` exactly.

States with an x-antisymmetric part (isoviscous model)
    RBC.step (rbc_galerkin.py:426-436) evolves any complex state.  The engine
    then carries a second real channel ``theta_o`` = Im theta (x-odd fields)
    beside ``theta`` = Re theta: the per-mode solves run on both channels
    with the same factors and the quadratic terms couple them on the half
    grid (even part f_e g_e + f_o g_o, odd part f_e g_o + f_o g_e).  The
    stream function of channel b is stored as ``psi_o`` with
    psi = i psi + (-psi_o) in the reference's convention.  Coefficient
    arrays whose imaginary part stays within ANTISYM_TOL of their largest
    entry are mirror states and keep the one-channel path.

Layouts (row-major, run-major)
    spectral vector      [B][Nk][Nz]          (n = Nk*Nz per run)
    padded half grid     [B][Nx+1][Nz]        (x = j*Lx/(2Nx), j = 0..Nx)
    unpadded half grid   [B][Nx/2+1][Nz]
    Stokes product grid  [B][Nx+1][Q]         (Q z-padded CGL nodes)
    Stokes z fields      [B][Nk][4][Q]
    Krylov bases         [B][m+1][n], [B][m][n]
    LU factors           [B][Nk][Nz*Nz] column-major, pivots [B][Nk][Nz]
"""
import json
import collections
import contextlib
import threading

import numpy as np
import torch

from . import _numerics as NUM
from . import kclock, runtime
from .basis import shape_ops

KIND_CONST, KIND_DEPTH, KIND_COUPLED = "const", "eta_z", "coupled"
# an imaginary part of at most this fraction of the largest coefficient is
# roundoff of a mirror-symmetric state (the rfft of an x-even grid field)
ANTISYM_TOL = 1e-10


def antisym_part(theta):
    """-> the relative size of the imaginary part of a coefficient array:
    max |Im theta| / max |theta| (0 for real arrays)."""
    a = np.asarray(theta)
    if not np.iscomplexobj(a) or a.size == 0:
        return 0.0
    return float(np.abs(a.imag).max()) / max(float(np.abs(a).max()), 1e-300)
PROFILES = {"harm": 0, "diag": 1, "geom": 2}

MEMO_RTOL = 1e-12          # coupled_model.py:327, 749-765
PC_APPLY = "inv"           # default application of the per-mode preconditioner blocks
# "fp64": every kernel in FP64.  "mixed": the dense rescue factorisation in
# FP32 (FP32 LU, FP32 triangular solves) under the FP64 defect correction
# against the FP64 matrix-free operator, which refines its result to the
# reference's acceptance test; runs whose FP32 attempt misses it are redone
# with the FP64 factor.  The Krylov operators, preconditioner blocks,
# Gram-Schmidt, least squares and Helmholtz solves stay FP64.
PRECISIONS = ("fp64", "mixed")
RESCUE_BYTES = 24 << 30    # HBM for dense rescue matrices factored together
# FGMRES of the mixed-precision rescue (FP32 dense LU as the preconditioner)
RESCUE_RESTART, RESCUE_MAXITER, RESCUE_RTOL = 20, 60, 1e-12
# after an FP32-factored rescue of a run misses the acceptance test, that
# run's next rescues use the FP64 factor directly this many times
RESCUE_FP32_BACKOFF = 8
# HBM for the lagged dense factors of one engine (one n x n factor per slot):
# the legacy route's frozen factorisation and the mixed-precision rescue's
# reused FP32 factor
LAG_STORE_BYTES = 96 << 30
# the mixed-precision rescue refactors a run's reused FP32 factor once the
# FGMRES on it took more than this many iterations, or after RESCUE_REUSE
# uses (the BFBT route has no lagged factor of the reference's, so its
# ``lag`` field does not apply; the legacy route keeps ``lag``)
RESCUE_REFRESH_ITERS = 32
RESCUE_REUSE = 24
# when a run of the mixed rescue refreshes, the runs past this fraction of
# their factor's life (uses since the refresh against ``lag``, FGMRES
# iterations against RESCUE_REFRESH_ITERS) are refreshed in the same batch
CO_REFRESH_FRACTION = 0.5
# mixed precision: a run whose last BFBT solve was rejected and rescued
# skips the BFBT attempt (800 iterations that end in the rescue anyway) and
# solves on its reused FP32 factor directly, for at most this many solves in
# a row; then it tries BFBT again with the full budget.  The rescue starts
# from zero, so a carried solve equals the rescue after a failed attempt
# bit for bit.  0 turns the carry off (the FP64 precision never carries).
RESCUE_CARRY = 16
# iteration budget of such a renewed BFBT attempt (a probe): the rescue route
# costs about as much as ~110 BFBT iterations per solve, so a BFBT that needs
# more is not worth returning to; a probe that misses keeps the run carried
RESCUE_PROBE_ITERS = 100
# mixed precision: a run whose BFBT solve converged but needed more than this
# many iterations is carried onto the factor route as well (with the same
# probes back): with the reused FP32 factor a solve there costs about as much
# as ~100 BFBT iterations at 64x96, and BFBT needs 100-800 from step ~60 of
# the 64x96 run on
RESCUE_SWITCH_ITERS = 100
# why a run is carried: its last BFBT attempt was rejected (it then holds a
# reused factor), or its last BFBT solve was accepted but cost more than
# RESCUE_SWITCH_ITERS (its first carried solve factors)
CARRY_REJECTED, CARRY_COSTLY = 1, 2
DENSE_FP64_MAXN = 8192     # largest n of the FP64 dense factor (FP32: 32768)
ORTHO = {"mgs": 0, "bgs2": 1}
ORTHO_DEFAULT = "mgs"      # FGMRES orthogonalisation (block variant: "bgs2")


LAW_ROW = 8      # doubles per run in the viscosity-law parameter rows


def _law_row(p):
    """Per-run viscosity-law parameters as the eta kernels read them:
    (ln deta_T, Ts, eta_min, eta_max, eta_smooth, law, E, V) with law 0 the
    exponential law, 1 the T-only Arrhenius law, 2 the pressure-dependent pV
    law calibrated by ``reference_state.calibrate_pv`` (Ts, E, V its own;
    coupled_model.py:186-229, 369-375)."""
    Ts = float(p.get("Ts", 0.0))
    V = float(p.get("arrhenius_V", 0.0))
    law, E = (1.0 if Ts > 0.0 else 0.0), 0.0
    if V > 0.0:
        from .reference_state import REFERENCE_DEFAULTS, calibrate_pv
        a = str(p.get("reference_anchor", "D"))
        rec = calibrate_pv(REFERENCE_DEFAULTS, anchor=a, V=V,
                           deta_T=float(p.get("deta_T", 1000.0)) if a == "D" else None)
        law, Ts, E, V = 2.0, rec["Ts"], rec["E"], rec["V"]
    return np.array([np.log(float(p.get("deta_T", 1000.0))), Ts,
                     float(p.get("eta_min", 1e-3)), float(p.get("eta_max", 1e3)),
                     float(p.get("eta_smooth", 0.0)), law, E, V])


def _is_legacy(p):
    """Whether a run's parameters select the lagged-LU route
    (precond='legacy', the ConvectionConfig default)."""
    return str(p.get("precond", "bfbt")).strip().lower() == "legacy"


LDS_BYTES = 160 * 1024      # per workgroup on gfx950


def inverse_fits_lds(N, esize):
    """Whether the Gauss-Jordan inversion of an N x N block fits in LDS."""
    return N * N * esize + 2 * N * esize + 4 * N <= LDS_BYTES


# HIP-graph captures are taken one at a time in the process: torch registers
# generator state when a capture begins, which it refuses while another
# thread's capture is open.  Launches and replays from other threads go on.
_CAPTURE_LOCK = threading.Lock()
class _ByteGate:
    """Admission of the transient dense matrices of rescues and lagged-factor
    refreshes across the engines of a process (the server's bucket workers):
    a thread asks for the bytes it will hold and waits while the bytes other
    threads hold plus its own would exceed ``budget``, so the server's one
    RESCUE_BYTES of headroom covers them all, and small grids' rescues are not
    queued behind large ones.  A request larger than the budget goes ahead
    alone.  Waiting requests are admitted in arrival order (a later one does
    not pass an earlier one that still waits), so a large request is not
    starved by a stream of small ones.  A thread's nested requests (a rescue
    inside a refresh) never wait, so a waiting thread holds nothing and no two
    threads wait on each other."""

    def __init__(self, budget):
        self.budget = int(budget)
        self.used = 0
        self.cv = threading.Condition()
        self.held = {}       # thread ident -> [depth, bytes]
        self.queue = collections.deque()   # tickets of the waiting requests
        self.tickets = 0

    @contextlib.contextmanager
    def hold(self, nbytes):
        nbytes = max(0, int(nbytes))
        me = threading.get_ident()
        with self.cv:
            h = self.held.get(me)
            if h is None:
                ticket = self.tickets
                self.tickets += 1
                self.queue.append(ticket)
                while (self.queue[0] != ticket
                       or (self.used > 0 and self.used + nbytes > self.budget)):
                    self.cv.wait()
                self.queue.popleft()
                self.cv.notify_all()          # the next waiting request may fit too
                h = self.held[me] = [0, 0]
            h[0] += 1
            h[1] += nbytes
            self.used += nbytes
        try:
            yield
        finally:
            with self.cv:
                h = self.held[me]
                h[0] -= 1
                h[1] -= nbytes
                self.used -= nbytes
                if h[0] == 0:
                    del self.held[me]
                self.cv.notify_all()


_RESCUE_GATE = _ByteGate(RESCUE_BYTES)
_TRANSIENT_STREAMS = {}


def _transient(shape, dtype, dev):
    """A transient dense rescue buffer.  The caching allocator keeps freed
    blocks per allocating stream, so buffers of engines on different streams
    would each keep their own cache; these come from one side stream per
    device (one pool), marked as used on the caller's stream so a block is
    not handed out again before the caller's kernels on it are done."""
    side = _TRANSIENT_STREAMS.get(dev)
    if side is None:
        side = _TRANSIENT_STREAMS.setdefault(dev, torch.cuda.Stream(device=dev))
    cur = torch.cuda.current_stream(dev)
    with torch.cuda.stream(side):
        t = torch.empty(*shape, dtype=dtype, device=dev)
    t.record_stream(cur)
    return t


class Solver:
    """Device FGMRES workspace for B independent systems of size n
    (krylov_fgmres.py:60-179 per run, driven in lockstep rounds)."""

    def __init__(self, eng, n, m, maxiter, rtol):
        self.e = eng
        B, dt, dev = eng.B, eng.dtype, eng.dev
        self.n, self.m, self.maxiter, self.rtol = int(n), int(m), int(maxiter), float(rtol)
        z = lambda *s: torch.zeros(*s, dtype=dt, device=dev)
        self.x, self.b, self.w = z(B, n), z(B, n), z(B, n)
        self.vin, self.zbuf = z(B, n), z(B, n)
        self.V, self.Z = z(B, m + 1, n), z(B, m, n)
        self.H, self.cs, self.sn = z(B, (m + 1) * m), z(B, m), z(B, m)
        self.g, self.y = z(B, m + 1), z(B, m)
        self.ctl = torch.zeros(10, B, dtype=torch.int32, device=dev)
        self.sc = torch.zeros(8, B, dtype=torch.float64, device=dev)
        # unfinished-run counts of the last two rounds: device slots, and
        # pinned host copies read one round late
        self.ortho = ORTHO[eng.ortho]
        self.cnt = torch.zeros(2, dtype=torch.int32, device=dev)
        self.cnt_host = torch.zeros(2, dtype=torch.int32, pin_memory=True)
        self.cnt_ev = [torch.cuda.Event(), torch.cuda.Event()]
        self.rounds = 0
        self._graphs = {}

    def act_pre(self):
        return self.ctl[5]

    def act_mv(self):
        return self.ctl[6]

    def _round(self, precond, op):
        """Enqueue one FGMRES round: cycle end, M^{-1}, A, Arnoldi step."""
        e, L = self.e, self.e.L
        B, n, m, P = e.B, self.n, self.m, e.prec
        s = e.stream()
        L.fg_cycle_end(P, B, n, m, e.p(self.x), e.p(self.Z), e.p(self.V),
                       e.p(self.H), e.p(self.g), e.p(self.y), e.p(self.vin),
                       e.p(self.zbuf), e.p(self.ctl), s)
        precond(self.vin, self.zbuf, self.act_pre())
        op(self.zbuf, self.w, self.act_mv())
        L.fg_post(P, B, n, m, e.p(self.b), e.p(self.w), e.p(self.zbuf),
                  e.p(self.Z), e.p(self.V), e.p(self.H), e.p(self.cs),
                  e.p(self.sn), e.p(self.g), e.p(self.ctl), e.p(self.sc), self.ortho, s)

    def _poll(self, slot):
        """Count the unfinished runs into device slot ``slot`` and start its
        copy to the pinned host slot."""
        e = self.e
        e.L.fg_count(e.B, e.p(self.ctl), e.p(self.cnt, slot), e.stream())
        self.cnt_host[slot:slot + 1].copy_(self.cnt[slot:slot + 1], non_blocking=True)
        self.cnt_ev[slot].record()

    def solve(self, precond, op, active=None, maxiter=None, key=None, budget=None):
        """Run FGMRES on every active run; ``precond(vin, zout, mask)`` and
        ``op(zin, wout, mask)`` apply M^{-1} and A to the masked runs.
        ``budget`` (host int array [B], optional) replaces the iteration
        budget per run.

        With ``key`` set and the engine's ``graph_rounds`` G > 0, the rounds
        run as replays of a HIP graph of G rounds captured once per ``key``
        (the launches of a round have fixed arguments: buffers, sizes and the
        device-side masks), and the host checks completion once per replay.
        Rounds after a run's completion skip every kernel for it, so the
        result is the eager one."""
        e, L = self.e, self.e.L
        B, n, m, P = e.B, self.n, self.m, e.prec
        mi = self.maxiter if maxiter is None else int(maxiter)
        L.fg_init(P, B, n, m, e.p(self.b), e.p(self.x), e.p(self.V), e.p(self.g),
                  e.p(self.ctl), e.p(self.sc), self.rtol, 0.0, mi, e.p(active), e.stream())
        if budget is not None:
            self.ctl[7].copy_(torch.as_tensor(np.asarray(budget, dtype=np.int32),
                                              device=e.dev))
        G = int(getattr(e, "graph_rounds", 0)) if key is not None else 0
        if G > 0 and kclock.armed():
            G = 0             # clocked launches stay eager
        graph = None
        rounds = 0
        if G > 0:
            graph = self._graphs.get(key)
            if graph is None:
                # one eager round first: kernel attributes and lazily queried
                # device limits are set outside the capture
                self._round(precond, op)
                rounds += 1
                graph = torch.cuda.CUDAGraph()
                with _CAPTURE_LOCK:
                    with torch.cuda.graph(graph, capture_error_mode="thread_local"):
                        for _ in range(G):
                            self._round(precond, op)
                self._graphs[key] = graph
        # Round (or replay) r is enqueued before the host waits for the
        # count of r-1, so the device never idles on the host's check; a run
        # that is done skips every kernel of a round (the masks), so the work
        # enqueued after the last run finished is a no-op.
        polls = 0
        while True:
            slot = polls & 1
            if graph is not None:
                graph.replay()
                rounds += G
            else:
                self._round(precond, op)
                rounds += 1
            self._poll(slot)
            polls += 1
            if polls >= 2:
                prev = slot ^ 1
                self.cnt_ev[prev].synchronize()
                if int(self.cnt_host[prev]) == 0:
                    break
        self.cnt_ev[(polls - 1) & 1].synchronize()
        self.rounds += rounds
        return rounds

    def results(self):
        """-> (iters[B], info[B], true_rel_res[B]) on the host."""
        c = self.ctl.cpu().numpy()
        r = self.sc.cpu().numpy()
        return c[3].copy(), c[4].copy(), r[3].copy()


class Engine:
    """B runs of one kind and one shape.  ``params`` is a list of per-run
    dicts with keys Ra, deta_T, Ts, eta_min, eta_max, eta_smooth,
    internal_heating, dissipation_number (missing keys take the defaults)."""

    def __init__(self, kind, Nx, Nz, Lx, params, prec=0, device=None,
                 profile="harm", stokes_rtol=1e-8, stokes_restart=80,
                 stokes_maxiter=800, theta_rtol=1e-10, theta_restart=40,
                 theta_maxiter=200, defect_tol=1e-8, max_K=8, K=3, pc_apply=None,
                 ortho=None, precision="fp64", graph_rounds=0):
        self.dev = runtime.require_device() if device is None else device
        # > 0: FGMRES rounds replayed from HIP graphs of this many rounds
        self.graph_rounds = int(graph_rounds)
        self.L = runtime.load()
        if precision not in PRECISIONS:
            raise ValueError("precision must be one of %s, got %r" % (PRECISIONS, precision))
        if precision == "mixed" and int(prec) != 0:
            raise ValueError("precision='mixed' runs on the FP64 engine (prec=0)")
        self.precision = precision
        self.rescue_prec = 1 if precision == "mixed" else 0
        self.rescue_carry = RESCUE_CARRY if self.rescue_prec == 1 else 0
        pc_apply = PC_APPLY if pc_apply is None else pc_apply
        if pc_apply not in ("lu", "inv"):
            raise ValueError("pc_apply must be 'lu' or 'inv', got %r" % (pc_apply,))
        # per-mode preconditioner blocks: LU + two triangular sweeps per
        # apply ("lu"), or an explicit inverse applied as one matrix-vector
        # product ("inv")
        self.pc_apply = pc_apply
        self.ortho = ORTHO_DEFAULT if ortho is None else ortho
        if self.ortho not in ORTHO:
            raise ValueError("ortho must be one of %s, got %r" % (sorted(ORTHO), ortho))
        self.kind = kind
        self.prec = int(prec)
        self.dtype = torch.float64 if self.prec == 0 else torch.float32
        self.esize = 8 if self.prec == 0 else 4
        self.ops = shape_ops(Nx, Nz, Lx)
        o = self.ops
        self.Nx, self.Nz, self.Nk, self.Q = o.Nx, o.Nz, o.Nk, o.Q
        if self.pc_apply == "inv" and not inverse_fits_lds(self.Nz, self.esize):
            self.pc_apply = "lu"     # the in-LDS inversion does not take this Nz
        self.nrp, self.nru = o.nrp, o.nru
        self.n = self.Nk * self.Nz
        self.B = B = len(params)
        self.t = o.device_tensors(self.dtype, self.dev, schur=(kind == KIND_COUPLED))
        self.profile = PROFILES[profile]
        dt, dev, N, Nk, n = self.dtype, self.dev, self.Nz, self.Nk, self.n
        z = lambda *s: torch.zeros(*s, dtype=dt, device=dev)
        # per-run physics
        self.params = [dict(p) for p in params]
        P = self.params
        self.Ra = torch.tensor([float(p.get("Ra", 1e4)) for p in P], dtype=dt, device=dev)
        self.Hh = np.array([float(p.get("internal_heating", 0.0)) for p in P])
        self.Dih = np.array([float(p.get("dissipation_number", 0.0)) for p in P])
        self.H = torch.tensor(self.Hh, dtype=dt, device=dev)
        self.diss_scale = torch.tensor(self.Dih / np.array([float(p.get("Ra", 1e4)) for p in P]),
                                       dtype=dt, device=dev)
        law = np.stack([_law_row(p) for p in P])
        self.lawh = law
        self.law = torch.tensor(law, dtype=torch.float64, device=dev)
        self.counters = torch.zeros(B, 8, dtype=torch.float64, device=dev)
        # state
        self.theta, self.theta_prev = z(B, n), z(B, n)
        self.general = False           # x-odd channel carried (isoviscous only)
        self.theta_o = self.psi_o = None
        self.psi, self.psi_prev, self.vsol = z(B, n), z(B, n), z(B, n)
        self.omega = z(B, n)
        self.has_hist = np.zeros(B, dtype=bool)
        self.dt_prev = np.zeros(B)
        # scratch
        self.zv = z(B, n * 4)
        self.zv2 = z(B, n * 4)
        self.gp1, self.gp2 = z(B, self.nrp * N), z(B, self.nrp * N)
        self.gp3, self.gp4 = z(B, self.nrp * N), z(B, self.nrp * N)
        self.ux_p, self.uz_p = z(B, self.nrp * N), z(B, self.nrp * N)
        self.ux_u, self.uz_u = z(B, self.nru * N), z(B, self.nru * N)
        self.th_u = z(B, self.nru * N)
        self.s1, self.s2, self.s3 = z(B, n), z(B, n), z(B, n)
        self.umax = torch.zeros(B, dtype=torch.float64, device=dev)
        self.umax2 = torch.zeros(B, dtype=torch.float64, device=dev)
        self.dtv = torch.zeros(B, dtype=torch.float64, device=dev)
        self.cflp = torch.zeros(B, 4, dtype=torch.float64, device=dev)
        self.diag_out = torch.zeros(B, 6, dtype=torch.float64, device=dev)
        self.a0 = z(B)
        self.b0 = z(B)
        self.b1 = z(B)
        self.hlam = z(B)
        self._a0_fact = np.full(B, np.nan)
        self.all_on = torch.ones(B, dtype=torch.int32, device=dev)
        self.k2 = (self.t["kx"] * self.t["kx"]).contiguous()
        self.n_steps = np.zeros(B, dtype=np.int64)
        self.n_stokes = np.zeros(B, dtype=np.int64)
        # theta machinery (implicit BDF2) for eta_z / coupled; the explicit
        # isoviscous step for const
        if kind == KIND_CONST:
            self._init_const()
        else:
            self.helm = z(B, Nk, N * N)
            self.helm_piv = torch.zeros(B, Nk, N, dtype=torch.int32, device=dev)
            self.tsol = Solver(self, n, theta_restart, theta_maxiter, theta_rtol)
            self.d2v, self.adv = z(B, n), z(B, n)
            self.src = z(B, n)
            if kind == KIND_DEPTH:
                self._init_depth()
            elif kind == KIND_COUPLED:
                self._init_coupled(stokes_rtol, stokes_restart, stokes_maxiter,
                                   defect_tol, max_K, K)
            else:
                raise ValueError("unknown engine kind %r" % (kind,))
        self._init_budget()

    # ------------------------------------------------------------- plumbing
    def p(self, t, off=0):
        if t is None:
            return None
        return t.data_ptr() + off * t.element_size()

    def stream(self):
        return torch.cuda.current_stream(self.dev).cuda_stream

    # GEMM roles (they name the kernel symbols in traces)
    Z_ROLE, X_ROLE, MODE_ROLE, DENSE_ROLE = 0, 1, 2, 3

    def gemm(self, M, N, K, A, ars, acs, Bm, brs, bcs, C, crs, ccs,
             alpha=1.0, beta=0.0, nz1=1, nz2=1, sa=(0, 0), sb=(0, 0), sc=(0, 0),
             aoff=0, boff=0, coff=0, rscale=None, rperiod=1, active=None,
             mask_mode=0, rpr=1, role=0):
        self.L.gemm_ex(role, self.prec, M, N, K, alpha, beta,
                    self.p(A, aoff), ars, acs, sa[0], sa[1],
                    self.p(Bm, boff), brs, bcs, sb[0], sb[1],
                    self.p(C, coff), crs, ccs, sc[0], sc[1],
                    nz1, nz2, self.p(rscale), rperiod, self.p(active),
                    mask_mode if active is not None else 0, rpr, self.stream())

    def zrows(self, src, Mat, dst, K, O, src_ld, dst_ld, alpha=1.0, beta=0.0,
              rscale=None, active=None, soff=0, doff=0, nb=None):
        """dst[(b,k)][o] = alpha rs(k) sum_i src[(b,k)][i] Mat[o][i] over all
        B*Nk mode rows (a z-transform of every x-mode row)."""
        nb = self.B if nb is None else nb
        self.gemm(nb * self.Nk, O, K, src, src_ld, 1, Mat, 1, K, dst, dst_ld, 1,
                  alpha=alpha, beta=beta, aoff=soff, coff=doff, rscale=rscale,
                  rperiod=self.Nk, active=active, mask_mode=1, rpr=self.Nk)

    def xleft(self, S, R, Kin, src, src_ld, src_bs, dst, dst_ld, dst_bs, C,
              soff=0, doff=0, active=None, beta=0.0):
        """per run b: dst[b][r][c] = sum_m S[r][m] src[b][m][c] (x-DFT)."""
        self.gemm(R, C, Kin, S, Kin, 1, src, src_ld, 1, dst, dst_ld, 1, beta=beta,
                  nz1=self.B, sb=(src_bs, 0), sc=(dst_bs, 0), boff=soff, coff=doff,
                  active=active, mask_mode=2, role=self.X_ROLE)

    def green(self, src, dst, active=None, nb=None):
        """dst[b][k] = G_k src[b][k] (stokes_coupled.py:318-325)."""
        N, n = self.Nz, self.n
        nb = self.B if nb is None else nb
        self.gemm(nb, N, N, src, n, 1, self.t["G"], 1, N, dst, n, 1,
                  nz1=1, nz2=self.Nk, sa=(0, N), sb=(0, N * N), sc=(0, N),
                  active=active, mask_mode=1, rpr=1, role=self.MODE_ROLE)

    # ------------------------------------------------------------ state I/O
    def set_theta(self, th, idx=None):
        """Upload coefficient arrays ``th`` (B', Nk, Nz) (runs ``idx``).  The
        imaginary part of an isoviscous state is kept (``theta_o``) when it
        exceeds ANTISYM_TOL; other imaginary parts are dropped."""
        dst = (slice(None) if idx is None else torch.as_tensor(idx, device=self.dev))
        th = np.asarray(th).reshape(-1, self.n)
        self.theta[dst] = self._dev_real(np.real(th))
        odd = np.array([self.kind == KIND_CONST and antisym_part(r) > ANTISYM_TOL for r in th])
        if odd.any():
            self._ensure_general()
        if self.general:
            self.theta_o[dst] = self._dev_real(np.where(odd[:, None], np.imag(th), 0.0))

    def _dev_real(self, a):
        a = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64))
        return a.reshape(-1, self.n).to(device=self.dev, dtype=self.dtype)

    def get_theta(self):
        """-> the (B, Nk, Nz) states: real, or complex once the x-odd channel
        is carried."""
        th = self.get(self.theta)
        return th + 1j * self.get(self.theta_o) if self.general else th

    def _ensure_general(self):
        """Allocate the x-odd channel of the isoviscous state and its work
        buffers; every later step of this engine carries both channels (runs
        without an antisymmetric part carry zeros in it)."""
        if self.general:
            return
        if self.kind != KIND_CONST:
            raise ValueError("only the isoviscous model carries states with an "
                             "x-antisymmetric part")
        B, n, N = self.B, self.n, self.Nz
        z = lambda *sh: torch.zeros(*sh, dtype=self.dtype, device=self.dev)
        self.theta_o, self.psi_o, self.omega_o = z(B, n), z(B, n), z(B, n)
        self.nc_o, self.s1o = z(B, n), z(B, n)
        self.zvo, self.zvo2 = z(B, 2 * n), z(B, 2 * n)
        gp = lambda: z(B, self.nrp * N)
        gu = lambda: z(B, self.nru * N)
        self.ux_pe, self.uz_po, self.gpo1, self.gpo2, self.eg1o = gp(), gp(), gp(), gp(), gp()
        self.ux_ue, self.uz_uo, self.th_uo = gu(), gu(), gu()
        self.bud_o = torch.zeros(B, dtype=torch.float64, device=self.dev)
        self.general = True

    def get(self, t):
        return t.detach().to("cpu", torch.float64).numpy().reshape(self.B, self.Nk, self.Nz)

    def reset_history(self, idx=None):
        if idx is None:
            self.has_hist[:] = False
        else:
            self.has_hist[idx] = False

    def admit(self, slot, params, theta):
        """Load a new run into ``slot``: its physics, its initial real
        coefficients ``theta`` (Nk, Nz), cleared history, memo, cached
        factors and counters.  The other slots are untouched, so runs can
        join a bucket while the rest keep stepping (continuous batching)."""
        i = int(slot)
        if not 0 <= i < self.B:
            raise IndexError("slot %d outside 0..%d" % (i, self.B - 1))
        p = dict(params)
        self.params[i] = p
        Ra = float(p.get("Ra", 1e4))
        self.Hh[i] = float(p.get("internal_heating", 0.0))
        self.Dih[i] = float(p.get("dissipation_number", 0.0))
        self.lawh[i] = _law_row(p)
        vals = torch.tensor([Ra, self.Hh[i], self.Dih[i] / Ra], dtype=self.dtype)
        self.Ra[i], self.H[i], self.diss_scale[i] = vals.to(self.dev)
        self.law[i] = torch.from_numpy(self.lawh[i]).to(self.dev)
        self.counters[i].zero_()
        self.set_theta(np.asarray(theta)[None], [i])
        for t in (self.theta_prev, self.psi, self.psi_prev, self.vsol, self.omega):
            t[i].zero_()
        if self.general:
            self.psi_o[i].zero_()
            self.omega_o[i].zero_()
        self.has_hist[i] = False
        self.dt_prev[i] = 0.0
        self._a0_fact[i] = np.nan
        self.n_steps[i] = 0
        self.n_stokes[i] = 0
        if self.kind == KIND_CONST:
            self._lam_fact[i] = np.nan
        elif self.kind == KIND_DEPTH:
            eg = self._depth_profile(i)
            self.eta_grid_h[i] = eg
            prof = self.ops.Vi @ eg
            self.eta_prof_c[i] = torch.from_numpy(prof).to(self.dev, self.dtype)
            self.eta_c_depth[i].zero_()
            self.eta_c_depth[i, :self.Nz] = self.eta_prof_c[i]
            mask = torch.zeros(self.B, dtype=torch.int32, device=self.dev)
            mask[i] = 1
            self._reduced_op_factor(self.eta_prof_c, self.vv_lu, self.vv_piv, mask)
        else:
            self.memo_valid[i] = False
            for a in (self.memo_hits, self.bfbt_iters, self.last_bfbt_iters,
                      self.last_bfbt_info, self.n_rescue, self.n_rescue_direct,
                      self.n_rescue_weak, self.n_defect_failed, self.n_defect_double_fail,
                      self.n_rescue_fp32, self.n_rescue_fp32_redo, self.rescue_iters,
                      self.fp32_backoff, self.n_since_refresh, self.n_refresh,
                      self.n_defect, self.n_lag_krylov, self.last_K, self.lag_iters,
                      self.rescue_streak, self.carry_run, self.n_rescue_carry,
                      self.n_bfbt_probe, self.n_singular_factor):
                a[i] = 0
            self.last_bfbt_res[i] = 0.0
            self.last_defect_res[i] = 0.0
            self.last_corr[i] = 0.0
            self.lag_have[i] = False
            self.force_refresh[i] = False
            self.legacy[i] = _is_legacy(p)
            self.lag[i] = max(1, int(p.get("lag", 16)))
            if self.legacy[i] and self._lag_bytes() > LAG_STORE_BYTES:
                raise ValueError("precond='legacy' does not fit the lagged-factor store "
                                 "at this shape and slot count")
            self.last_route[i] = "legacy" if self.legacy[i] else "bfbt"

    # coupled-model counters carried by a snapshot (per-run host arrays)
    _SNAP_COUNTERS = ("memo_hits", "bfbt_iters", "last_bfbt_iters", "last_bfbt_info",
                      "last_bfbt_res", "n_rescue", "n_rescue_direct", "n_rescue_weak",
                      "n_defect_failed", "n_defect_double_fail", "n_rescue_fp32",
                      "n_rescue_fp32_redo", "rescue_iters", "fp32_backoff", "last_defect_res",
                      "n_since_refresh", "n_refresh", "n_defect", "n_lag_krylov", "last_K",
                      "last_corr", "lag_iters", "force_refresh", "rescue_streak",
                      "carry_run", "n_rescue_carry", "n_bfbt_probe", "n_singular_factor")

    def snapshot(self, slot):
        """-> the state of the run in ``slot`` as a dict of numpy arrays:
        physics, theta, the BDF2 history, counters, and the lagged factor
        when the run holds one.  ``restore`` of it into a slot of an engine of
        the same kind, shape and precision continues the run exactly."""
        i = int(slot)
        if not 0 <= i < self.B:
            raise IndexError("slot %d outside 0..%d" % (i, self.B - 1))
        h = lambda t: t[i].detach().cpu().numpy().copy()
        snap = dict(kind=np.array(self.kind), shape=np.array([self.Nx, self.Nz]),
                    Lx=np.array(float(self.ops.Lx)),
                    precision=np.array(self.precision), prec=np.array(self.prec),
                    params=np.array(json.dumps(self.params[i], sort_keys=True)),
                    theta=h(self.theta), n_steps=np.array(self.n_steps[i]),
                    n_stokes=np.array(self.n_stokes[i]), counters=h(self.counters))
        if self.general:
            snap["theta_odd"] = h(self.theta_o)
        if self.kind != KIND_CONST:
            snap.update(theta_prev=h(self.theta_prev), psi_prev=h(self.psi_prev),
                        dt_prev=np.array(self.dt_prev[i]), has_hist=np.array(self.has_hist[i]))
        if self.kind == KIND_COUPLED:
            for k in self._SNAP_COUNTERS:
                snap["c_" + k] = np.array(getattr(self, k)[i])
            snap["last_route"] = np.array(self.last_route[i])
            if self.lag_have[i]:
                snap.update(lag_lu=h(self.lag_lu), lag_piv=h(self.lag_piv),
                            eta_lag=h(self.eta_lag), lag_prec=np.array(self.lag_prec))
        return snap

    def restore(self, slot, snap):
        """Load a ``snapshot`` into ``slot`` (any slot; the other slots are
        untouched)."""
        i = int(slot)
        kind, shape = str(snap["kind"]), [int(v) for v in snap["shape"]]
        if (kind != self.kind or shape != [self.Nx, self.Nz]
                or float(snap["Lx"]) != float(self.ops.Lx)):
            raise ValueError("snapshot of a %s %dx%d run does not fit this %s %dx%d engine"
                             % (kind, shape[0], shape[1], self.kind, self.Nx, self.Nz))
        if str(snap["precision"]) != self.precision or int(snap["prec"]) != self.prec:
            raise ValueError("snapshot precision %s does not match the engine's %s"
                             % (snap["precision"], self.precision))
        params = json.loads(str(snap["params"]))
        th = np.asarray(snap["theta"]).reshape(self.Nk, self.Nz)
        if "theta_odd" in snap:
            th = th + 1j * np.asarray(snap["theta_odd"]).reshape(self.Nk, self.Nz)
        self.admit(i, params, th)
        d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(self.dev)
        self.n_steps[i] = int(snap["n_steps"])
        self.n_stokes[i] = int(snap["n_stokes"])
        self.counters[i] = d(snap["counters"]).to(self.counters.dtype)
        if self.kind != KIND_CONST:
            self.theta_prev[i] = d(snap["theta_prev"]).to(self.dtype)
            self.psi_prev[i] = d(snap["psi_prev"]).to(self.dtype)
            self.dt_prev[i] = float(snap["dt_prev"])
            self.has_hist[i] = bool(snap["has_hist"])
        if self.kind == KIND_COUPLED:
            for k in self._SNAP_COUNTERS:
                if "c_" + k in snap:      # snapshots written before a counter existed
                    getattr(self, k)[i] = snap["c_" + k]
            self.last_route[i] = str(snap["last_route"])
            if "lag_lu" in snap:
                if int(snap["lag_prec"]) != self.lag_prec or self._lag_store() is None:
                    raise ValueError("the snapshot's lagged factor does not fit this engine")
                self.lag_lu[i] = d(snap["lag_lu"])
                self.lag_piv[i] = d(snap["lag_piv"])
                self.L.dense_lu_perm(1, self.n, self.p(self.lag_piv[i]), self.p(self.lag_perm[i]),
                                     None, self.stream())
                self.eta_lag[i] = d(snap["eta_lag"]).to(self.dtype)
                self.lag_have[i] = True

    # --------------------------------------------------------- energy budget
    def _init_budget(self):
        """Quadrature weights of the energy budget (heating.py:269-306,
        dissipation.py:380-444, 540-571): <f> = vol_mean(grid(f_hat)) =
        (wq V) . f_hat[k=0]; on the padded half grid the x-mean weights are
        the k = 0 row of the analysis matrix."""
        o, N, dev = self.ops, self.Nz, self.dev
        c0 = o.wq @ o.V                       # <f> from the k = 0 coefficients
        ct = c0.copy()
        ct[-2:] = 0.0                         # the source's tau rows are zeroed
        wx = o.Ae_p[0]                        # padded-grid x-mean
        tt = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=self.dtype, device=dev)
        self.bud_w = dict(c0=tt(c0), phi=tt(np.outer(wx, o.Vi.T @ ct).ravel()),
                          pi=tt(np.outer(wx, o.wq).ravel()))
        # <theta^{n-1}>, <theta^n>, <theta^{n+1}>, Ra <u_z theta>, (Di/Ra) <Phi>
        self.bud = torch.zeros(self.B, 5, dtype=torch.float64, device=dev)
        self.bud_abc = np.zeros((self.B, 3))

    def _budget_means(self, prev, cur, abc, active):
        """Record <theta^{n-1}> (``prev``, or None), <theta^n> and the step's
        (a0, b0, b1) before the step overwrites the history."""
        L, s, N, n = self.L, self.stream(), self.Nz, self.n
        c0 = self.p(self.bud_w["c0"])
        if prev is not None:
            L.wdot(self.prec, self.B, N, self.p(prev), n, None, 0, c0, self.p(self.bud[:, 0]),
                   5, self.p(active), s)
        L.wdot(self.prec, self.B, N, self.p(cur), n, None, 0, c0, self.p(self.bud[:, 1]), 5,
               self.p(active), s)
        act = np.ones(self.B, bool) if active is None else active.cpu().numpy().astype(bool)
        self.bud_abc[act] = abc[act]

    def energy_budget(self, active=None):
        """The discrete energy budget of the step just taken, for every
        active run at once (heating.balance, heating.py:279-306, with the
        dissipation term of dissipation.balance, dissipation.py:540-571):

            storage  = a0 <theta^{n+1}> - b0 <theta^n> + b1 <theta^{n-1}>
            residual = storage - (nu_bot - nu_top) - H - (Di/Ra) <Phi>

        with <Phi> of the tau-zeroed source field for the stream function of
        theta^{n+1} and the eta it consumed (dissipation.py:380-392), and
        Ra <u_z theta> through the same padded pipeline (dissipation.py:423-444).
        The advection term is 0 (the exact operator).  -> dict of (B,) arrays."""
        L, t, s, N, n, nrp = self.L, self.t, self.stream(), self.Nz, self.n, self.nrp
        B = self.B
        d = self.diagnostics(active)             # Stokes of theta^{n+1}; Nu pair
        c0 = self.p(self.bud_w["c0"])
        L.wdot(self.prec, B, N, self.p(self.theta), n, None, 0, c0, self.p(self.bud[:, 2]), 5,
               self.p(active), s)
        # Ra <u_z theta> on the padded grid
        self._padded_velocity(self.psi, active)
        self.zrows(self.theta, t["V"], self.zv, N, N, N, N, active=active)
        self.xleft(t["Ce_p"], nrp, self.Nk, self.zv, N, n, self.gp1, N, nrp * N, N,
                   active=active)
        L.wdot(self.prec, B, nrp * N, self.p(self.uz_p), nrp * N, self.p(self.gp1), nrp * N,
               self.p(self.bud_w["pi"]), self.p(self.bud[:, 3]), 5, self.p(active), s)
        if self.general:
            # the x-odd parts add <uz_o theta_o>; the odd part of Phi has zero mean
            self.zrows(self.theta_o, t["V"], self.zvo, N, N, N, N, active=active)
            self.xleft(t["So_p"], nrp, self.Nk, self.zvo, N, n, self.gpo1, N, nrp * N, N,
                       active=active)
            L.wdot(self.prec, B, nrp * N, self.p(self.uz_po), nrp * N, self.p(self.gpo1),
                   nrp * N, self.p(self.bud_w["pi"]), self.p(self.bud_o), 1, self.p(active), s)
            self.bud[:, 3] += self.bud_o
        self.bud[:, 4].zero_()
        if np.any(self.Dih != 0.0):
            eta_c = (self.eta_c if self.kind == KIND_COUPLED else
                     self.eta_c_depth if self.kind == KIND_DEPTH else None)
            self._phi_grid(self.psi, eta_c, self.gp4, active,
                           out_o=self.eg1o if self.general else None)
            L.wdot(self.prec, B, nrp * N, self.p(self.gp4), nrp * N, None, 0,
                   self.p(self.bud_w["phi"]), self.p(self.bud[:, 4]), 5, self.p(active), s)
        m = self.bud.cpu().numpy()
        a0, b0, b1 = self.bud_abc.T
        storage = a0 * m[:, 2] - b0 * m[:, 1] + b1 * m[:, 0]
        nu_top, nu_bot = d[:, 0], d[:, 1]
        H = self.Hh.astype(float)
        diss = m[:, 4]
        Ra = np.array([float(p.get("Ra", 1e4)) for p in self.params])
        out = dict(storage=storage, nu_bot=nu_bot, nu_top=nu_top, heating=H,
                   advection=np.zeros(B), dissipation=diss,
                   residual=storage - (nu_bot - nu_top) - H - diss,
                   theta_mean=m[:, 2], T_mean=0.5 + m[:, 2], pi_theta=Ra * m[:, 3],
                   phi_integral=np.where(self.Dih != 0.0, diss * Ra / np.where(
                       self.Dih != 0.0, self.Dih, 1.0), 0.0))
        if active is not None:
            act = active.cpu().numpy().astype(bool)
            for k in out:
                out[k] = np.where(act, out[k], np.nan)
        return out

    # ------------------------------------------------------- shared pieces
    def project(self, src, dst, active=None):
        """Vi (Re V src): project_mirror on real fields (convection_modes.py:134-153)."""
        N = self.Nz
        self.zrows(src, self.t["V"], self.s3, N, N, N, N, active=active)
        self.zrows(self.s3, self.t["Vi"], dst, N, N, N, N, active=active)

    def velocity_unpadded(self, psi, active=None, psi_o=None):
        """ux = grid(D psi), uz = grid(-i k psi) on the unpadded half grid
        (rbc_galerkin.py:355-372).  With ``psi_o`` (the x-odd channel's
        stream function) also the other parity parts: ux_ue = Ce(-D psi_o),
        uz_uo = So(k psi_o)."""
        N, n, nru = self.Nz, self.n, self.nru
        t = self.t
        self.zrows(psi, t["VD"], self.zv, N, N, N, N, active=active)
        self.zrows(psi, t["V"], self.zv2, N, N, N, N, rscale=t["kx"], active=active)
        self.xleft(t["So_u"], nru, self.Nk, self.zv, N, n, self.ux_u, N, nru * N, N,
                   active=active)
        self.xleft(t["Ce_u"], nru, self.Nk, self.zv2, N, n, self.uz_u, N, nru * N, N,
                   active=active)
        if psi_o is not None:
            self.zrows(psi_o, t["VD"], self.zvo, N, N, N, N, alpha=-1.0, active=active)
            self.zrows(psi_o, t["V"], self.zvo2, N, N, N, N, rscale=t["kx"], active=active)
            self.xleft(t["Ce_u"], nru, self.Nk, self.zvo, N, n, self.ux_ue, N, nru * N, N,
                       active=active)
            self.xleft(t["So_u"], nru, self.Nk, self.zvo2, N, n, self.uz_uo, N, nru * N, N,
                       active=active)

    def cfl(self, safety, dtmax, dtmin, active=None):
        """Per-run CFL dt of the current theta (rbc_galerkin.py:465-471,
        cfl_policy.py:64-74).  Solves Stokes (memoised for the coupled model)."""
        self.stokes(self.theta, active)
        g = self.general
        self.velocity_unpadded(self.psi, active, self.psi_o if g else None)
        L, B = self.L, self.B
        nh = self.nru * self.Nz
        if g:
            # full-grid max |u| = max |u_e| + |u_o| over the half grid
            L.maxabs_pair(self.prec, B, nh, self.p(self.ux_u), self.p(self.ux_ue),
                          self.p(self.umax), self.p(active), self.stream())
            L.maxabs_pair(self.prec, B, nh, self.p(self.uz_u), self.p(self.uz_uo),
                          self.p(self.umax2), self.p(active), self.stream())
        else:
            L.maxabs(self.prec, B, nh, self.p(self.ux_u), self.p(self.umax), self.p(active),
                     self.stream())
            L.maxabs(self.prec, B, nh, self.p(self.uz_u), self.p(self.umax2), self.p(active),
                     self.stream())
        torch.maximum(self.umax, self.umax2, out=self.umax)
        hc = self.ops.hcfl
        s = np.broadcast_to(np.asarray(safety, dtype=np.float64), (B,))
        pr = np.stack([s, np.full(B, hc), np.broadcast_to(dtmin, (B,)),
                       np.broadcast_to(dtmax, (B,))], 1).astype(np.float64)
        self.cflp.copy_(torch.from_numpy(pr))
        L.cfl_dt(B, self.p(self.umax), self.p(self.cflp), self.p(self.dtv), self.p(active),
                 self.stream())
        return self.dtv.cpu().numpy().copy()

    def diagnostics(self, active=None):
        """(B, 6): Nu_top, Nu_bot, Nu_energy, urms, umax, amp of the current
        theta (rbc_galerkin.py:439-453)."""
        t, N, n = self.t, self.Nz, self.n
        g = self.general
        self.stokes(self.theta, active)
        self.velocity_unpadded(self.psi, active, self.psi_o if g else None)
        self.zrows(self.theta, t["V"], self.zv, N, N, N, N, active=active)
        self.xleft(t["Ce_u"], self.nru, self.Nk, self.zv, N, n, self.th_u, N,
                   self.nru * N, N, active=active)
        self.zrows(self.theta, t["D"], self.s1, N, N, N, N, active=active)
        if g:
            self.zrows(self.theta_o, t["V"], self.zvo, N, N, N, N, active=active)
            self.xleft(t["So_u"], self.nru, self.Nk, self.zvo, N, n, self.th_uo, N,
                       self.nru * N, N, active=active)
            self.L.diag_general(self.prec, self.B, self.Nx, N, self.p(self.ux_u),
                                self.p(self.uz_u), self.p(self.th_u), self.p(self.ux_ue),
                                self.p(self.uz_uo), self.p(self.th_uo), self.p(self.s1), n,
                                self.p(t["V0"]), self.p(t["V1"]), self.p(t["wq"]),
                                self.p(self.diag_out), self.p(active), self.stream())
            return self.diag_out.cpu().numpy().copy()
        self.L.diag(self.prec, self.B, self.Nx, N, self.p(self.ux_u), self.p(self.uz_u),
                    self.p(self.th_u), self.p(self.s1), n, self.p(t["V0"]), self.p(t["V1"]),
                    self.p(t["wq"]), self.p(self.diag_out), self.p(active), self.stream())
        return self.diag_out.cpu().numpy().copy()

    # ------------------------------------------------- isoviscous (const)
    def _init_const(self):
        B, Nk, N, dev = self.B, self.Nk, self.Nz, self.dev
        z = lambda *s: torch.zeros(*s, dtype=self.dtype, device=dev)
        # lam = 0 factors: one set per shape, shared by every run
        self.z0 = z(1, Nk, N * N)
        self.z0_piv = torch.zeros(1, Nk, N, dtype=torch.int32, device=dev)
        zero = z(1)
        self.L.tau_assemble(self.prec, 0, 1, Nk, N, self.p(self.t["D2"]), self.p(self.t["kx"]),
                            self.p(zero), None, None, None, None, None, self.p(self.z0),
                            None, self.stream())
        self.L.lu_factor(self.prec, Nk, N, self.p(self.z0), self.p(self.z0_piv), Nk, None,
                         self.stream())
        self.zl = z(B, Nk, N * N)
        self.zl_piv = torch.zeros(B, Nk, N, dtype=torch.int32, device=dev)
        self.lam = z(B)
        self._lam_fact = np.full(B, np.nan)
        self.nc = z(B, self.n)
        self.eg1 = z(B, self.nrp * N)

    def _stokes_const(self, theta, active, omega=None, psi=None):
        """omega = Z^{-1}(-i k Ra theta), psi = Z^{-1}(-omega) (rbc_galerkin.py:342-353).
        The solves are real-linear: the x-odd channel b gives omega = -omega_o,
        psi = -psi_o with omega_o, psi_o from the same two solves of b."""
        L, B, Nk, N = self.L, self.B, self.Nk, self.Nz
        s = self.stream()
        omega = self.omega if omega is None else omega
        psi = self.psi if psi is None else psi
        L.lu_solve(self.prec, B * Nk, N, self.p(self.z0), self.p(self.z0_piv), self.p(theta),
                   self.p(omega), 1, -1.0, self.p(self.t["kx"]), self.p(self.Ra), Nk, 1,
                   self.p(active), s)
        L.lu_solve(self.prec, B * Nk, N, self.p(self.z0), self.p(self.z0_piv),
                   self.p(omega), self.p(psi), 1, -1.0, None, None, Nk, 1,
                   self.p(active), s)

    def _padded_velocity(self, psi, active):
        """ux = synth_pad(D psi) (x-odd), uz = synth_pad(-i k psi) (x-even)."""
        t, N, n, nrp = self.t, self.Nz, self.n, self.nrp
        self.zrows(psi, t["VD"], self.zv, N, N, N, N, active=active)
        self.zrows(psi, t["V"], self.zv2, N, N, N, N, rscale=t["kx"], active=active)
        self.xleft(t["So_p"], nrp, self.Nk, self.zv, N, n, self.ux_p, N, nrp * N, N,
                   active=active)
        self.xleft(t["Ce_p"], nrp, self.Nk, self.zv2, N, n, self.uz_p, N, nrp * N, N,
                   active=active)
        if self.general and psi is self.psi:
            # x-odd channel: ux_e = synth(-D psi_o), uz_o = synth(k psi_o)
            self.zrows(self.psi_o, t["VD"], self.zvo, N, N, N, N, alpha=-1.0, active=active)
            self.zrows(self.psi_o, t["V"], self.zvo2, N, N, N, N, rscale=t["kx"],
                       active=active)
            self.xleft(t["Ce_p"], nrp, self.Nk, self.zvo, N, n, self.ux_pe, N, nrp * N, N,
                       active=active)
            self.xleft(t["So_p"], nrp, self.Nk, self.zvo2, N, n, self.uz_po, N, nrp * N, N,
                       active=active)

    def _phi_grid(self, psi, eta_c, out, active, out_o=None):
        """(Di/Ra) Phi on the padded half grid (dissipation.py:313-362) for the
        (extrapolated) psi and eta coefficients ``eta_c`` (None: eta = 1)."""
        t, N, n, nrp = self.t, self.Nz, self.n, self.nrp
        # exx = -k D psi (x-even), exz = 0.5 (D D psi + k^2 psi) (x-odd)
        self.zrows(psi, t["D"], self.s1, N, N, N, N, active=active)
        self.zrows(self.s1, t["D"], self.s2, N, N, N, N, active=active)
        self.L.rowscale(self.prec, self.B, self.Nk, N, self.p(self.s1), self.p(t["kx"]), -1.0,
                        None, 0, self.p(self.s1), self.p(active), self.stream())
        self.L.rowscale(self.prec, self.B, self.Nk, N, self.p(psi), self.p(self.k2), 1.0,
                        None, 0, self.p(self.s3), self.p(active), self.stream())
        self.L.axpby(self.prec, n, self.B, 0.5, None, self.p(self.s2), 0.5, None,
                     self.p(self.s3), self.p(self.s2), self.p(active), self.stream())
        self.zrows(self.s1, t["V"], self.zv, N, N, N, N, active=active)
        self.zrows(self.s2, t["V"], self.zv2, N, N, N, N, active=active)
        self.xleft(t["Ce_p"], nrp, self.Nk, self.zv, N, n, self.gp1, N, nrp * N, N,
                   active=active)
        self.xleft(t["So_p"], nrp, self.Nk, self.zv2, N, n, self.gp2, N, nrp * N, N,
                   active=active)
        if out_o is not None:
            # x-odd channel (psi = -psi_o): exx_o = synth(-k D psi_o) (odd),
            # exz_e = synth(-(D D psi_o + k^2 psi_o) / 2) (even); eta = 1
            if eta_c is not None:
                raise ValueError("the two-channel dissipation is isoviscous only")
            L, B, s = self.L, self.B, self.stream()
            self.zrows(self.psi_o, t["D"], self.s1, N, N, N, N, active=active)
            self.zrows(self.s1, t["D"], self.s2, N, N, N, N, active=active)
            L.rowscale(self.prec, B, self.Nk, N, self.p(self.s1), self.p(t["kx"]), -1.0,
                       None, 0, self.p(self.s1), self.p(active), s)
            L.rowscale(self.prec, B, self.Nk, N, self.p(self.psi_o), self.p(self.k2), 1.0,
                       None, 0, self.p(self.s3), self.p(active), s)
            L.axpby(self.prec, n, B, 0.5, None, self.p(self.s2), 0.5, None,
                    self.p(self.s3), self.p(self.s2), self.p(active), s)
            self.zrows(self.s1, t["V"], self.zvo, N, N, N, N, active=active)
            self.zrows(self.s2, t["V"], self.zvo2, N, N, N, N, alpha=-1.0, active=active)
            self.xleft(t["So_p"], nrp, self.Nk, self.zvo, N, n, self.gpo1, N, nrp * N, N,
                       active=active)
            self.xleft(t["Ce_p"], nrp, self.Nk, self.zvo2, N, n, self.gpo2, N, nrp * N, N,
                       active=active)
            L.phi_grid_general(self.prec, B, nrp * N, self.p(self.gp1), self.p(self.gp2),
                               self.p(self.gpo1), self.p(self.gpo2), self.p(self.diss_scale),
                               self.p(out), self.p(out_o), self.p(active), s)
            return
        eta_g = None
        if eta_c is not None:
            self.zrows(eta_c, t["V"], self.zv, N, N, N, N, active=active)
            self.xleft(t["Ce_p"], nrp, self.Nk, self.zv, N, n, self.gp3, N, nrp * N, N,
                       active=active)
            eta_g = self.gp3
        self.L.phi_grid(self.prec, self.B, nrp * N, self.p(self.gp1), self.p(self.gp2),
                        self.p(eta_g), self.p(self.diss_scale), self.p(out), self.p(active),
                        self.stream())

    def _step_const(self, dt, active):
        """Explicit advection + BE diffusion step (rbc_galerkin.py:396-436)."""
        L, t, B, Nk, N, n, nrp = self.L, self.t, self.B, self.Nk, self.Nz, self.n, self.nrp
        s = self.stream()
        g = self.general
        self._stokes_const(self.theta, active)
        if g:
            self._stokes_const(self.theta_o, active, self.omega_o, self.psi_o)
        self._padded_velocity(self.psi, active)
        e2 = e2o = None
        if np.any(self.Dih != 0.0):
            self._phi_grid(self.psi, None, self.eg1, active, out_o=self.eg1o if g else None)
            e2, e2o = self.eg1, (self.eg1o if g else None)
        # theta derivatives: thx = synth(i k theta) (odd), thz = synth(D theta)
        self.zrows(self.theta, t["V_VD"], self.zv, N, 2 * N, N, 2 * N, active=active)
        hh = self.H if np.any(self.Hh != 0.0) else None
        if g:
            self.zrows(self.theta_o, t["V_VD"], self.zvo, N, 2 * N, N, 2 * N, active=active)
            L.advect_prod_general(self.prec, B, Nk, nrp, N, self.p(self.zv), self.p(self.zvo),
                                  self.p(t["kx"]), self.p(self.ux_pe), self.p(self.ux_p),
                                  self.p(self.uz_p), self.p(self.uz_po), self.p(hh),
                                  self.p(e2), self.p(e2o), self.p(self.s1), self.p(self.s1o),
                                  self.p(t["So_p"]), self.p(t["Ce_p"]), self.p(t["Ae_p"]),
                                  self.p(t["Ao_p"]), self.p(active), s)
            self.zrows(self.s1o, t["Vi"], self.nc_o, N, N, N, N, active=active)
        else:
            L.advect_prod(self.prec, B, Nk, nrp, N, self.p(self.zv), self.p(t["kx"]),
                          self.p(self.ux_p), self.p(self.uz_p), self.p(self.uz_p), self.p(hh),
                          self.p(e2), self.p(self.s1), self.p(t["So_p"]), self.p(t["Ce_p"]),
                          self.p(t["Ae_p"]), self.p(active), s)
        self.zrows(self.s1, t["Vi"], self.nc, N, N, N, N, active=active)
        # rhs = theta / dt + Nc ; out = Z_lam^{-1}(-rhs), lam = 1/dt
        dt = np.asarray(dt, dtype=np.float64)
        inv_dt = 1.0 / np.where(dt > 0.0, dt, 1.0)     # runs outside the mask carry dt = 0
        self.lam.copy_(torch.from_numpy(inv_dt).to(self.dtype))
        refac = np.where(self._lam_fact != inv_dt)[0]
        if active is not None:
            act = active.cpu().numpy().astype(bool)
            refac = refac[act[refac]]
        if len(refac):
            mask = torch.zeros(B, dtype=torch.int32, device=self.dev)
            mask[torch.as_tensor(refac, device=self.dev)] = 1
            L.tau_assemble(self.prec, 0, B, Nk, N, self.p(t["D2"]), self.p(t["kx"]),
                           self.p(self.lam), None, None, None, None, None, self.p(self.zl),
                           self.p(mask), s)
            L.lu_factor(self.prec, B * Nk, N, self.p(self.zl), self.p(self.zl_piv), Nk,
                        self.p(mask), s)
            self._lam_fact[refac] = inv_dt[refac]
        self._budget_means(None, self.theta, np.stack([inv_dt, inv_dt, np.zeros(B)], axis=1),
                           active)
        L.axpby(self.prec, n, B, 1.0, self.p(self.lam), self.p(self.theta), 1.0, None,
                self.p(self.nc), self.p(self.s1), self.p(active), s)
        L.lu_solve(self.prec, B * Nk, N, self.p(self.zl), self.p(self.zl_piv), self.p(self.s1),
                   self.p(self.theta), 1, -1.0, None, None, Nk, 0, self.p(active), s)
        if g:
            L.axpby(self.prec, n, B, 1.0, self.p(self.lam), self.p(self.theta_o), 1.0, None,
                    self.p(self.nc_o), self.p(self.s1o), self.p(active), s)
            L.lu_solve(self.prec, B * Nk, N, self.p(self.zl), self.p(self.zl_piv),
                       self.p(self.s1o), self.p(self.theta_o), 1, -1.0, None, None, Nk, 0,
                       self.p(active), s)

    def set_pc_apply(self, mode):
        """Switch the per-mode preconditioner blocks between 'lu' and 'inv';
        cached Helmholtz factors are rebuilt at the next step."""
        if mode not in ("lu", "inv"):
            raise ValueError("pc_apply must be 'lu' or 'inv', got %r" % (mode,))
        if mode == "inv" and not inverse_fits_lds(self.Nz, self.esize):
            raise ValueError("pc_apply='inv' needs the %d x %d block in LDS" % (self.Nz, self.Nz))
        if mode != self.pc_apply:
            self.pc_apply = mode
            self._a0_fact[:] = np.nan

    def _pc_factor(self, mats, piv, mask):
        """Factor the assembled per-mode blocks ``mats`` [B][Nk][N*N] in place
        (LU with pivots, or explicit inverse)."""
        B, Nk, N = self.B, self.Nk, self.Nz
        if self.pc_apply == "inv":
            self.L.invert(self.prec, B * Nk, N, self.p(mats), Nk, self.p(mask), self.stream())
        else:
            self.L.lu_factor(self.prec, B * Nk, N, self.p(mats), self.p(piv), Nk, self.p(mask),
                             self.stream())

    def _pc_apply(self, mats, piv, vin, zout, scale, mask):
        """zout_k = block_k^{-1} (scale * vin_k) with the tau entries zeroed."""
        B, Nk, N = self.B, self.Nk, self.Nz
        if self.pc_apply == "inv":
            self.L.inv_apply(self.prec, B * Nk, N, self.p(mats), self.p(vin), self.p(zout), 1,
                             scale, None, None, Nk, 0, self.p(mask), self.stream())
        else:
            self.L.lu_solve(self.prec, B * Nk, N, self.p(mats), self.p(piv), self.p(vin),
                            self.p(zout), 1, scale, None, None, Nk, 0, self.p(mask),
                            self.stream())

    # ---------------------------------------------------- depth-only eta(z)
    def _reduced_op_factor(self, prof_c, lu, piv, active, as_pc=False):
        """LU of Lm_k M + 2k^2 [D,[D,M]] G_k with tau rows v(0), v(1) for the
        Chebyshev profile ``prof_c`` [B][N] of every run (vv_stokes.py:266-307,
        bfbt.py:157-181)."""
        L, t, B, Nk, N = self.L, self.t, self.B, self.Nk, self.Nz
        s = self.stream()
        NN = N * N
        M0, T1, T2, T3, CM = self.rb_M0, self.rb_T1, self.rb_T2, self.rb_T3, self.rb_C
        L.cheb_mult(self.prec, B, N, self.p(prof_c), self.p(M0), self.p(active), s)
        mm = lambda A, sa, Bm, sb, C: self.gemm(N, N, N, A, N, 1, Bm, N, 1, C, N, 1, nz1=B,
                                                sa=(sa, 0), sb=(sb, 0), sc=(NN, 0),
                                                active=active, mask_mode=2,
                                                role=self.MODE_ROLE)
        mm(t["D2"], 0, M0, NN, T1)           # D2 M
        mm(t["D"], 0, M0, NN, T2)            # D M
        mm(T2, NN, t["D"], 0, T3)            # (D M) D
        mm(M0, NN, t["D2"], 0, CM)           # M D2
        L.axpby(self.prec, NN, B, 1.0, None, self.p(T1), -2.0, None, self.p(T3), self.p(T3),
                self.p(active), s)
        L.axpby(self.prec, NN, B, 1.0, None, self.p(T3), 1.0, None, self.p(CM), self.p(CM),
                self.p(active), s)           # Comm = (D2 M - 2 D M D) + M D2
        self.gemm(N, N, N, CM, N, 1, t["G"], N, 1, self.rb_CG, N, 1, nz1=B, nz2=Nk,
                  sa=(NN, 0), sb=(0, NN), sc=(Nk * NN, NN), active=active, mask_mode=2,
                  role=self.MODE_ROLE)
        L.tau_assemble(self.prec, 1, B, Nk, N, None, self.p(t["kx"]), None, self.p(M0),
                       self.p(T1), self.p(self.rb_CG), self.p(t["V0"]), self.p(t["V1"]),
                       self.p(lu), self.p(active), s)
        if as_pc:          # a preconditioner: factored the way _pc_apply applies it
            self._pc_factor(lu, piv, active)
        else:              # the depth-only Stokes operator itself: LU
            L.lu_factor(self.prec, B * Nk, N, self.p(lu), self.p(piv), Nk, self.p(active), s)

    def _alloc_reduced(self):
        B, Nk, N, dev = self.B, self.Nk, self.Nz, self.dev
        z = lambda *s: torch.zeros(*s, dtype=self.dtype, device=dev)
        self.rb_M0, self.rb_T1, self.rb_T2 = z(B, N * N), z(B, N * N), z(B, N * N)
        self.rb_T3, self.rb_C = z(B, N * N), z(B, N * N)
        self.rb_CG = z(B, Nk, N * N)

    def _depth_profile(self, i):
        """Window-clipped eta(z) profile of run ``i`` on the CGL nodes
        (convection_modes.py:86-110)."""
        o = self.ops
        lo, hi, w = self.lawh[i, 2], self.lawh[i, 3], self.lawh[i, 4]
        eg = np.exp(-self.lawh[i, 0] * (1.0 + (0.0 - 1.0) * o.z))
        elo, ehi = lo * (1.0 - 1e-12), hi * (1.0 + 1e-12)
        if w > 0.0:
            ylo, yhi = np.log(lo), np.log(hi)
            y = np.log(eg)
            sm = lambda q: (lambda c: c * c * c * (c * (6.0 * c - 15.0) + 10.0))(np.clip(q, 0.0, 1.0))
            Y = ylo + (y - ylo) * sm((y - ylo) / w)
            eg = np.exp(yhi - (yhi - Y) * sm((yhi - Y) / w))
        elif np.any((eg < elo) | (eg > ehi)):
            eg = np.clip(eg, lo, hi)
        return eg

    def _init_depth(self):
        B, Nk, N, dev = self.B, self.Nk, self.Nz, self.dev
        z = lambda *s: torch.zeros(*s, dtype=self.dtype, device=dev)
        self._alloc_reduced()
        o = self.ops
        prof = np.zeros((B, N))
        self.eta_grid_h = np.zeros((B, N))
        for i in range(B):
            self.eta_grid_h[i] = self._depth_profile(i)
            prof[i] = o.Vi @ self.eta_grid_h[i]
        self.eta_prof_c = torch.tensor(prof, dtype=self.dtype, device=dev)
        ec = np.zeros((B, Nk, N))
        ec[:, 0, :] = prof
        self.eta_c_depth = torch.tensor(ec.reshape(B, -1), dtype=self.dtype, device=dev)
        self.vv_lu = z(B, Nk, N * N)
        self.vv_piv = torch.zeros(B, Nk, N, dtype=torch.int32, device=dev)
        self._reduced_op_factor(self.eta_prof_c, self.vv_lu, self.vv_piv, None)

    def _stokes_depth(self, theta, active):
        """u = A_k^{-1}(i k Ra theta), psi = G_k u, omega = -u (vv_stokes.py:309-321)."""
        B, Nk, N = self.B, self.Nk, self.Nz
        self.L.lu_solve(self.prec, B * Nk, N, self.p(self.vv_lu), self.p(self.vv_piv),
                        self.p(theta), self.p(self.vsol), 1, 1.0, self.p(self.t["kx"]),
                        self.p(self.Ra), Nk, 0, self.p(active), self.stream())
        self.green(self.vsol, self.psi, active)

    # ------------------------------------------------------ coupled eta(T)
    def _init_coupled(self, rtol, restart, maxiter, defect_tol=1e-8, max_K=8, K=3):
        B, Nk, N, Q, dev = self.B, self.Nk, self.Nz, self.Q, self.dev
        z = lambda *s: torch.zeros(*s, dtype=self.dtype, device=dev)
        self._alloc_reduced()
        self.eta_p = z(B, self.nrp * N)        # padded law field
        self.eta_u = z(B, self.nru * N)        # eta on the Nx grid (last_eta)
        self.eta_c = z(B, self.n)              # its coefficients
        self.eta_S = z(B, self.nrp * Q)        # eta on the Stokes product grid
        self.prof = z(B, N)
        self.prof_c = z(B, N)
        self.bf_lu = z(B, Nk, N * N)
        self.bf_piv = torch.zeros(B, Nk, N, dtype=torch.int32, device=dev)
        self.yz = z(B, Nk * 4 * Q)
        # the intermediates of the three-launch Schur operator apply
        self.sw_elems = int(self.L.schur_apply_work(B, Nk, Q))
        self.sw = z(self.sw_elems)
        self.ssol = Solver(self, self.n, restart, maxiter, rtol)
        self.memo_X = z(B, self.n)
        self.memo_new = z(B, self.n)
        self.memo_valid = np.zeros(B, dtype=bool)
        self.memo_hits = np.zeros(B, dtype=np.int64)
        self.bfbt_iters = np.zeros(B, dtype=np.int64)
        self.last_bfbt_iters = np.zeros(B, dtype=np.int64)
        self.last_bfbt_info = np.zeros(B, dtype=np.int64)
        self.last_bfbt_res = np.zeros(B)
        self.n_rescue = np.zeros(B, dtype=np.int64)
        self.n_rescue_direct = np.zeros(B, dtype=np.int64)
        self.n_rescue_weak = np.zeros(B, dtype=np.int64)
        self.n_defect_failed = np.zeros(B, dtype=np.int64)
        self.n_defect_double_fail = np.zeros(B, dtype=np.int64)
        self.n_rescue_fp32 = np.zeros(B, dtype=np.int64)
        self.n_rescue_fp32_redo = np.zeros(B, dtype=np.int64)
        self.rescue_iters = np.zeros(B, dtype=np.int64)
        self.fp32_backoff = np.zeros(B, dtype=np.int64)
        # rescue carry (mixed precision): 1 while a run's last BFBT attempt
        # was rejected; carried solves since that attempt; carried in total
        self.rescue_streak = np.zeros(B, dtype=np.int64)
        self.carry_run = np.zeros(B, dtype=np.int64)
        self.n_rescue_carry = np.zeros(B, dtype=np.int64)
        self.n_bfbt_probe = np.zeros(B, dtype=np.int64)
        self.last_defect_res = np.zeros(B)
        self.last_route = ["bfbt"] * B
        self.defect_tol = defect_tol
        self.max_K, self.K = int(max_K), int(K)
        # legacy route (lagged factor + Picard defect correction) and the
        # lagged FP32 factor of the mixed-precision rescue
        # (coupled_model.py:509-581, 693-716; lagged_stokes.py:274-343)
        self.legacy = np.array([_is_legacy(p) for p in self.params])
        self.lag = np.array([max(1, int(p.get("lag", 16))) for p in self.params])
        self.eta_drift_tol = float(NUM.ETA_DRIFT_TOL)
        self.lag_prec = 1 if (self.rescue_prec == 1 or self.n > DENSE_FP64_MAXN) else 0
        self.lag_lu = self.lag_piv = self.lag_perm = self.lag_info = self.eta_lag = None
        self._ludm_ws = None
        self.lag_have = np.zeros(B, dtype=bool)
        self.force_refresh = np.zeros(B, dtype=bool)
        self.n_since_refresh = np.zeros(B, dtype=np.int64)
        self.n_refresh = np.zeros(B, dtype=np.int64)
        self.n_defect = np.zeros(B, dtype=np.int64)
        self.n_lag_krylov = np.zeros(B, dtype=np.int64)
        self.last_K = np.zeros(B, dtype=np.int64)
        self.last_corr = np.zeros(B)
        self.lag_iters = np.zeros(B, dtype=np.int64)
        self.lag_red = torch.zeros(4, B, dtype=torch.float64, device=dev)
        self.n_refresh_batches = 0
        self.n_singular_factor = np.zeros(B, dtype=np.int64)
        if self.legacy.any() and self._lag_bytes() > LAG_STORE_BYTES:
            raise ValueError("precond='legacy' keeps one %d x %d factor per slot: %d slots "
                             "need %.1f GB, above the %.0f GB lagged-factor store"
                             % (self.n, self.n, B, self._lag_bytes() / 1e9,
                                LAG_STORE_BYTES / 1e9))

    def _eta_of(self, theta, active):
        """eta from theta: padded synthesis, law, window, truncation to Nk
        modes, Nx-grid synthesis, positivity guard (coupled_model.py:401-496),
        the BFBT profile, the coefficients and the Stokes-grid values."""
        L, t, B, Nk, N, n, Q = self.L, self.t, self.B, self.Nk, self.Nz, self.n, self.Q
        nrp, nru = self.nrp, self.nru
        s = self.stream()
        self.zrows(theta, t["V"], self.zv, N, N, N, N, active=active)
        self.xleft(t["Ce_p"], nrp, Nk, self.zv, N, n, self.gp1, N, nrp * N, N, active=active)
        L.eta_law(self.prec, B, nrp, N, self.p(self.gp1), self.p(t["z"]), self.p(self.law),
                  self.p(self.eta_p), self.p(self.counters), self.p(active), s)
        self.xleft(t["Ae_p"], Nk, nrp, self.eta_p, N, nrp * N, self.s1, N, n, N, active=active)
        self.xleft(t["Ce_u"], nru, Nk, self.s1, N, n, self.eta_u, N, nru * N, N, active=active)
        L.eta_guard_profile(self.prec, B, self.Nx, N, self.p(self.eta_u), self.p(self.law),
                            self.p(self.prof), self.profile, self.p(self.counters),
                            self.p(active), s)
        self.xleft(t["Ae_u"], Nk, nru, self.eta_u, N, nru * N, self.s2, N, n, N, active=active)
        self.zrows(self.s2, t["Vi"], self.eta_c, N, N, N, N, active=active)
        # eta on the z-padded Stokes product grid
        self.zrows(self.eta_c, t["Vp"], self.yz, N, Q, N, Q, active=active)
        self.xleft(t["Ce_p"], nrp, Nk, self.yz, Q, Nk * Q, self.eta_S, Q, nrp * Q, Q,
                   active=active)

    def schur_matvec(self, v, out, active, eta=None, eta_bs=None):
        """Coupled Schur operator (FreeSchurOperator.matvec, lagged_stokes.py:202-220)
        with the eta products evaluated pseudo-spectrally, in three launches
        (``smx_schur_apply``): the per-mode z-synthesis of v and of its Green
        field, the four x-products with eta, and the truncation/derivative
        combination with the tau rows.  ``eta`` / ``eta_bs``: the eta field on
        the product grid and its batch stride (default: this engine's)."""
        t = self.t
        eta = self.eta_S if eta is None else eta
        eta_bs = self.nrp * self.Q if eta_bs is None else eta_bs
        self.L.schur_apply(self.prec, self.B, self.Nk, self.Nz, self.nrp, self.Q, self.p(v),
                           self.p(eta), eta_bs, self.p(out), self.p(t["Wz"]), self.p(t["Mc"]),
                           self.p(t["kx"]), self.p(t["So_p"]), self.p(t["Ce_p"]),
                           self.p(t["Ao_p"]), self.p(t["Ae_p"]), self.p(t["V0"]),
                           self.p(t["V1"]), self.p(self.sw), self.sw_elems, self.p(active),
                           self.stream())

    def assemble_schur(self, r, out=None):
        """Dense real Schur matrix of run ``r`` at its current eta (the
        imaginary-to-imaginary block of CoupledStokes.assemble_schur,
        stokes_coupled.py:242-289): per-shift Chebyshev multiplication matrices,
        the block combinations, one batched GEMM against the Green operators
        and the tau rows."""
        L, t, N, Nk, n = self.L, self.t, self.Nz, self.Nk, self.n
        NN = N * N
        st = self.stream()
        if getattr(self, "_asm", None) is None:
            z = lambda *sh: torch.zeros(*sh, dtype=self.dtype, device=self.dev)
            self._asm = dict(M=z(Nk, NN), P=z(Nk, NN), Q=z(Nk, NN), DM=z(Nk, NN),
                             R=z(Nk, NN), U=z(Nk * Nk, NN))
        w = self._asm
        A = torch.empty(n, n, dtype=self.dtype, device=self.dev) if out is None else out
        L.cheb_mult(self.prec, Nk, N, self.p(self.eta_c[r]), self.p(w["M"]), None, st)
        mm = lambda X, sx, Y, sy, C: self.gemm(N, N, N, X, N, 1, Y, N, 1, C, N, 1, nz1=Nk,
                                               sa=(sx, 0), sb=(sy, 0), sc=(NN, 0),
                                               role=self.DENSE_ROLE)
        mm(t["D2"], 0, w["M"], NN, w["P"])        # D2 M_d
        mm(w["M"], NN, t["D2"], 0, w["Q"])        # M_d D2
        mm(t["D"], 0, w["M"], NN, w["DM"])        # D M_d
        mm(w["DM"], NN, t["D"], 0, w["R"])        # (D M_d) D
        L.schur_blocks(self.prec, Nk, N, self.p(w["M"]), self.p(w["P"]), self.p(w["Q"]),
                       self.p(w["R"]), self.p(t["kx"]), -1.0, self.p(A), self.p(w["U"]), st)
        # A_ij += U_ij G_j
        self.gemm(N, N, N, w["U"], N, 1, t["G"], N, 1, A, n, 1, beta=1.0, nz1=Nk, nz2=Nk,
                  sa=(Nk * NN, NN), sb=(0, NN), sc=(N * n, N), role=self.DENSE_ROLE)
        L.schur_tau(self.prec, Nk, N, self.p(t["V0"]), self.p(t["V1"]), self.p(A), st)
        return A

    def _mask(self, sel):
        """Device int32 run mask of the host bool array ``sel``."""
        return torch.as_tensor(np.asarray(sel).astype(np.int32), device=self.dev)

    def _stokes_coupled(self, theta, active):
        """eta_of, the Schur right-hand side, then per run the BFBT route
        (preconditioner + FGMRES + rescue) or the legacy route (lagged factor +
        Picard defect correction), and the unpack (coupled_model.py:656-723)."""
        L, t, B, Nk, N, n = self.L, self.t, self.B, self.Nk, self.Nz, self.n
        s = self.stream()
        self._eta_of(theta, active)
        sv = self.ssol
        L.rowscale(self.prec, B, Nk, N, self.p(theta), self.p(t["kx"]), 1.0, self.p(self.Ra), 1,
                   self.p(sv.b), self.p(active), s)
        act = np.ones(B, bool) if active is None else active.cpu().numpy().astype(bool)
        leg, bf = act & self.legacy, act & ~self.legacy
        if leg.any():
            self._stokes_legacy(leg)
            m = torch.as_tensor(leg, device=self.dev)
            self.vsol[m] = sv.x[m]
        if bf.any():
            self._stokes_bfbt(theta, self._mask(bf) if leg.any() else active)
        self.green(self.vsol, self.psi, active)

    def _stokes_bfbt(self, theta, active):
        """BFBT preconditioner + FGMRES on the Schur system, with the rescue
        of rejected solves (coupled_model.py:588-654, precond='bfbt'); the
        solution goes to self.vsol.  sv.b holds the right-hand side."""
        t, B, N = self.t, self.B, self.Nz
        sv = self.ssol
        act = np.ones(B, bool) if active is None else active.cpu().numpy().astype(bool)
        carry = self._carry_select(act)
        bf = act & ~carry
        if bf.any():
            amask = self._mask(bf) if carry.any() else active
            self.gemm(B, N, N, self.prof, N, 1, t["Qprof"], 1, N, self.prof_c, N, 1,
                      active=amask, mask_mode=1, rpr=1)
            self._reduced_op_factor(self.prof_c, self.bf_lu, self.bf_piv, amask, as_pc=True)

            def precond(vin, zout, mask):
                self._pc_apply(self.bf_lu, self.bf_piv, vin, zout, 1.0, mask)

            def op(zin, wout, mask):
                self.schur_matvec(zin, wout, mask)

            # a run back from carried solves attempts BFBT as a probe
            probe = bf & self._carry_ready() & self._carries()
            budget = None
            if probe.any():
                self.n_bfbt_probe[probe] += 1
                budget = np.where(probe, min(RESCUE_PROBE_ITERS, sv.maxiter), sv.maxiter)
            sv.solve(precond, op, amask, key="stokes", budget=budget)
            it, info, res = sv.results()
            self.bfbt_iters[bf] += it[bf]
            self.last_bfbt_iters[bf] = it[bf]
            self.last_bfbt_info[bf] = info[bf]
            self.last_bfbt_res[bf] = res[bf]
            # coupled_model.py:607-610: a BFBT solve resets the defect fields
            self.last_K[bf] = 0
            self.last_defect_res[bf] = res[bf]
            for r in np.where(bf)[0]:
                self.last_route[r] = "bfbt"
            bad = bf & (info != 0) & ~(res <= sv.rtol)
            self.carry_run[bf] = 0
            costly = bf & ~bad & (it > RESCUE_SWITCH_ITERS) & self._carries()
            self.rescue_streak[bf] = np.where(bad, CARRY_REJECTED,
                                              np.where(costly, CARRY_COSTLY, 0))[bf]
            if np.any(bad):
                self.n_rescue[bad] += 1
                self._rescue(theta, np.where(bad)[0])
        if carry.any():
            self.last_bfbt_iters[carry] = 0
            self.carry_run[carry] += 1
            self.n_rescue_carry[carry] += 1
            self._rescue(theta, np.where(carry)[0])
            for r in np.where(carry)[0]:
                self.last_route[r] = "carry:" + self.last_route[r]
        if active is None:
            self.vsol.copy_(sv.x)
        else:
            m = active.bool()
            self.vsol[m] = sv.x[m]

    def _carry_select(self, act):
        """Runs of ``act`` whose Stokes solve skips the BFBT attempt and goes
        straight to the mixed-precision rescue on the run's reused FP32
        factor: the run's last BFBT attempt was rejected and it holds a
        factor, or its last BFBT solve was accepted after more than
        RESCUE_SWITCH_ITERS iterations; and fewer than ``rescue_carry`` solves
        were carried since that attempt.  The attempt after them is a probe
        with a budget of RESCUE_PROBE_ITERS iterations.  Nothing is carried in
        FP64, or when the factors do not fit the lagged-factor store (a fresh
        FP32 factorisation per step would then replace the attempt)."""
        if not self._carries():
            return np.zeros(self.B, dtype=bool)
        return act & self._carry_ready() & (self.carry_run < self.rescue_carry)

    def _carry_ready(self):
        """Runs in a carry: rejected ones holding a reused factor, and costly
        ones (which factor on their first carried solve)."""
        return (((self.rescue_streak == CARRY_REJECTED) & self.lag_have)
                | (self.rescue_streak == CARRY_COSTLY))

    def _carries(self):
        """Whether this engine carries rescued runs (mixed precision, the
        reused factors fit the lagged-factor store)."""
        return self.rescue_carry > 0 and self._lag_bytes() <= LAG_STORE_BYTES

    # dense LU solves: "mw" spreads each matrix's sweeps over workgroups of
    # 64 rows (smx_dense_lu_solve_mw); "batched" runs one workgroup per
    # matrix (smx_dense_lu_solve)
    dense_solve_kind = "mw"

    def _dense_solve(self, prec, nmat, LU, perm, b, x, mask, piv=None):
        """x[q] = A_q^-1 b[q] for the dense factors LU [nmat][n][n] (row
        order ``perm`` from smx_dense_lu_perm, pivots ``piv``) of the
        matrices whose ``mask`` entry is set (all when None), on this
        engine's stream (the multi-workgroup form with its own workspace)."""
        n = self.n
        if self.dense_solve_kind == "batched":
            self.L.dense_lu_solve(prec, nmat, n, self.p(LU), n * n, self.p(piv), self.p(b),
                                  self.p(x), None if mask is None else self.p(mask),
                                  self.stream())
            return
        need = self.L.dense_lu_solve_mw_bytes(prec, nmat, n)
        if self._ludm_ws is None or self._ludm_ws.numel() < need:
            self._ludm_ws = torch.empty(need, dtype=torch.uint8, device=self.dev)
        ws = self._ludm_ws
        self.L.dense_lu_solve_mw(prec, nmat, n, self.p(LU), n * n, self.p(perm), self.p(b),
                                 self.p(x), None if mask is None else self.p(mask),
                                 self.p(ws), ws.numel(), self.stream())

    def _rescue(self, theta, idx, rprec=None):
        """Recompute rejected BFBT solves with the robust route
        (coupled_model.py:613-654) for every listed run at once: dense Schur LU
        at this eta + Picard defect correction against the matrix-free
        operator, accepted when the true relative residual meets the
        tolerance, else the direct dense solve (Nk*Nz <= 4096) or a counted
        weak acceptance.  ``rprec`` 1 factors in FP32 (mixed precision, see
        ``_rescue_mixed``)."""
        sv, n, B = self.ssol, self.n, self.B
        bad = [int(r) for r in idx]
        if rprec is None and self.rescue_prec == 1 and n <= DENSE_FP64_MAXN:
            # runs whose last FP32 attempt failed take the FP64 factor for a while
            back = [r for r in bad if self.fp32_backoff[r] > 0]
            if back:
                self.fp32_backoff[back] -= 1
                self._rescue(theta, back, rprec=0)
                bad = [r for r in bad if r not in set(back)]
                if not bad:
                    return
        lagged = rprec is None and self.rescue_prec == 1
        rprec = self.rescue_prec if rprec is None else int(rprec)
        if n > DENSE_FP64_MAXN:
            rprec = 1          # the FP64 factor does not go this large
        if lagged and self._lag_store() is not None:
            self._rescue_lagged(theta, bad)
            return
        # dense matrices of at most RESCUE_BYTES at a time
        cap = max(1, int(RESCUE_BYTES // (n * n * (4 if rprec else 8))))
        if len(bad) > cap:
            for c0 in range(0, len(bad), cap):
                self._rescue(theta, bad[c0:c0 + cap], rprec)
            return
        # the rescue's transient matrices: nb factors, and the FP64 assembly
        # matrix of an FP32 rescue
        es = 4 if rprec else 8
        with _RESCUE_GATE.hold(len(bad) * n * n * es + (n * n * 8 if rprec else 0)):
            redo = self._rescue_dense(theta, bad, rprec)
        if redo:
            # runs the FP32 factor did not bring to the acceptance test, with
            # the FP64 factor (after the FP32 matrices are released)
            self.n_rescue_fp32_redo[redo] += 1
            self.fp32_backoff[redo] = RESCUE_FP32_BACKOFF
            self._rescue(theta, redo, rprec=0)

    def _rescue_dense(self, theta, bad, rprec):
        """The rescue of ``bad`` (at most RESCUE_BYTES of dense matrices) in
        factor precision ``rprec``.  -> the runs an FP32 factor left above
        the acceptance test, for the caller to redo in FP64."""
        sv, n, B = self.ssol, self.n, self.B
        nb = len(bad)
        tol = self.defect_tol
        bt = torch.as_tensor(bad, device=self.dev)
        L, st = self.L, self.stream()
        if rprec == 0:
            A = _transient((nb, n, n), self.dtype, self.dev)
            for q, r in enumerate(bad):
                self.assemble_schur(r, out=A[q])
        else:
            # assemble each matrix in FP64, keep its FP32 rounding
            A = _transient((nb, n, n), torch.float32, self.dev)
            A64 = _transient((n, n), self.dtype, self.dev)
            for q, r in enumerate(bad):
                self.assemble_schur(r, out=A64)
                A[q].copy_(A64)
            del A64
        piv = torch.empty(nb, n, dtype=torch.int32, device=self.dev)
        info = torch.empty(nb, dtype=torch.int32, device=self.dev)
        L.dense_lu(rprec, nb, n, self.p(A), n * n, self.p(piv), self.p(info), st)
        perm = torch.empty(nb, n, dtype=torch.int32, device=self.dev)
        L.dense_lu_perm(nb, n, self.p(piv), self.p(perm), None, st)
        zero_pivot = info.cpu().numpy() != 0
        if zero_pivot.any():
            self.n_singular_factor[np.asarray(bad)[zero_pivot]] += 1
        mask = torch.zeros(B, dtype=torch.int32, device=self.dev)
        mask[bt] = 1
        if rprec != 0:
            return self._rescue_mixed(theta, bad, bt, A, perm, mask, piv)

        def solve(R):
            R = R.contiguous()
            X = torch.empty_like(R)
            self._dense_solve(0, nb, A, perm, R, X, None, piv)
            return X

        def afree(X):
            self.s1[bt] = X
            self.schur_matvec(self.s1, self.s2, mask)
            return self.s2[bt].clone()

        prec = self.prec
        nrm = torch.empty(nb, dtype=torch.float64, device=self.dev)

        def norms(X):
            L.rownorm(prec, nb, n, n, self.p(X), self.p(nrm), None, st)
            return nrm.cpu().numpy().copy()

        def residual(X, out):
            # out = b - A_free X
            L.axpby(prec, n, nb, 1.0, None, self.p(b), -1.0, None, self.p(afree(X)),
                    self.p(out), None, st)
            return out

        b = sv.b[bt].clone()
        x0 = solve(b)
        x = x0.clone()
        rv = torch.empty_like(b)
        bn = np.maximum(norms(b), 1e-300)
        xn = np.maximum(norms(x0), 1e-300)
        kmax = self.max_K if tol is not None else self.K
        live = np.ones(nb, dtype=bool)
        live_dev = torch.ones(nb, dtype=torch.int32, device=self.dev)
        lastc = np.zeros(nb)
        kused = np.zeros(nb, dtype=np.int64)
        for _ in range(kmax):
            dx = solve(residual(x, rv))
            c = norms(dx) / xn
            # x += dx on the runs still iterating
            L.axpby(prec, n, nb, 1.0, None, self.p(x), 1.0, None, self.p(dx), self.p(x),
                    self.p(live_dev), st)
            lastc = np.where(live, c, lastc)
            kused += live
            if tol is not None:
                live = live & ~(c <= tol)
                if not live.any():
                    break
                live_dev.copy_(torch.from_numpy(live.astype(np.int32)))
        res = norms(residual(x, rv)) / bn
        res_direct = None
        for q, r in enumerate(bad):
            # the rescue refreshes the lagged factor at this eta and runs the
            # defect solve; a failed solve is refreshed and run once more
            # (coupled_model.py:634-641), which repeats the same arithmetic at
            # the same eta, so its counters count both passes
            failed = tol is not None and (not np.isfinite(lastc[q]) or lastc[q] > tol)
            passes = 2 if failed else 1
            self.n_refresh[r] += passes
            self.n_defect[r] += passes
            self.last_K[r] = int(kused[q])
            if failed:
                self.n_defect_failed[r] += 2
                self.n_defect_double_fail[r] += 1
            self.last_defect_res[r] = res[q]
            if res[q] <= sv.rtol or not np.isfinite(res[q]):
                self.last_route[r] = "bfbt->legacy"
            elif self.Nk * self.Nz > 4096:
                self.n_rescue_weak[r] += 1
                self.last_route[r] = "bfbt->legacy(weak)"
            else:
                if res_direct is None:
                    res_direct = norms(residual(x0, rv)) / bn
                x[q] = x0[q]
                self.n_rescue_direct[r] += 1
                self.last_defect_res[r] = res_direct[q]
                self.last_route[r] = "bfbt->direct"
        sv.x[bt] = x
        return []

    def _rescue_mixed(self, theta, bad, bt, A32, perm, mask, piv):
        """Mixed-precision rescue: FGMRES in FP64 on the matrix-free Schur
        operator, right-preconditioned by the FP32 dense LU of the same
        operator (its solves take and return FP64 vectors).  The outer FP64
        residual refines the FP32 factor's answer (GMRES-based iterative
        refinement), and the Krylov acceleration removes the few directions
        in which the FP32 factor is poor, which a plain defect iteration
        would leave at the FP32 level.  -> the runs whose true residual
        misses the acceptance test (for the FP64 factor)."""
        sv, n, nb, L, st = self.ssol, self.n, len(bad), self.L, self.stream()
        rs = getattr(self, "rsol", None)
        if rs is None:
            rs = self.rsol = Solver(self, n, RESCUE_RESTART, RESCUE_MAXITER, RESCUE_RTOL)
        rs.b.copy_(sv.b)
        R = torch.empty(nb, n, dtype=self.dtype, device=self.dev)
        X = torch.empty_like(R)

        def precond(vin, zout, m):
            # rows whose mask is off this round keep zout: in a residual
            # round it holds the iterate the operator is applied to
            R.copy_(vin[bt])
            self._dense_solve(2, nb, A32, perm, R, X, None, piv)
            on = m[bt].bool()[:, None]
            zout[bt] = torch.where(on, X, zout[bt])

        def op(zin, wout, m):
            self.schur_matvec(zin, wout, m)

        rs.solve(precond, op, mask)
        it, info, res = rs.results()
        good, redo = [], []
        for r in bad:
            if np.isfinite(res[r]) and res[r] <= sv.rtol:
                good.append(r)
                self.n_rescue_fp32[r] += 1
                self.rescue_iters[r] += int(it[r])
                self.last_defect_res[r] = res[r]
                self.last_route[r] = "bfbt->legacy"
            else:
                redo.append(r)
        if good:
            g = torch.as_tensor(good, device=self.dev)
            sv.x[g] = rs.x[g]
        if redo and n > DENSE_FP64_MAXN:
            # no FP64 factor at this size: the refined solution is kept as a
            # counted weak acceptance (the reference's route for Nk*Nz > 4096)
            for r in redo:
                self.n_rescue_weak[r] += 1
                self.last_defect_res[r] = res[r]
                self.last_route[r] = "bfbt->legacy(weak)"
            g = torch.as_tensor(redo, device=self.dev)
            sv.x[g] = rs.x[g]
            return []
        return redo

    def _rescue_lagged(self, theta, bad):
        """Mixed-precision rescue on the reused FP32 factor.  FGMRES on the
        matrix-free operator at the current eta is preconditioned by the
        factor of an earlier rescue (``_lag_krylov``); the true residual
        decides acceptance, so the factor's age moves only the iteration
        count.  A run refactors at this eta when it has no factor, after
        RESCUE_REUSE uses, or when its last solve took more than
        RESCUE_REFRESH_ITERS iterations; a miss with an older factor
        refactors and solves again, and a miss with a fresh one is redone
        with the FP64 factor (the weak acceptance above the FP64 factor's
        size)."""
        B, n = self.B, self.n
        sel = np.zeros(B, dtype=bool)
        sel[bad] = True
        life = np.where(self.legacy, self.lag, RESCUE_REUSE)
        ref = sel & (~self.lag_have | self.force_refresh | (self.n_since_refresh >= life)
                     | (self.lag_iters > RESCUE_REFRESH_ITERS))
        if ref.any():
            # a factorisation batch costs about the same for one matrix as for
            # dozens (its panel chain is latency-bound), so runs past half of
            # their factor's life are refreshed with it rather than alone later
            f = CO_REFRESH_FRACTION
            ref |= sel & self.lag_have & ((self.n_since_refresh >= f * life)
                                         | (self.lag_iters > f * RESCUE_REFRESH_ITERS))
        self._lag_refresh(ref)
        self.n_since_refresh[sel] += 1
        ok, it, res = self._lag_krylov(sel)
        self.lag_iters[sel] = it[sel]
        retry = sel & ~ok & ~ref
        if retry.any():
            self._lag_refresh(retry)
            self.n_since_refresh[retry] += 1
            ok2, it2, res2 = self._lag_krylov(retry)
            self.lag_iters[retry] = it2[retry]
            ok = ok | (retry & ok2)
            res = np.where(retry, res2, res)
        good = sel & ok
        self.n_rescue_fp32[good] += 1
        for r in np.where(good)[0]:
            self.last_route[r] = "bfbt->legacy"
        miss = sel & ~ok
        if not miss.any():
            return
        if n > DENSE_FP64_MAXN:
            self.n_rescue_weak[miss] += 1
            for r in np.where(miss)[0]:
                self.last_route[r] = "bfbt->legacy(weak)"
            return
        self.n_rescue_fp32_redo[miss] += 1
        self.fp32_backoff[miss] = RESCUE_FP32_BACKOFF
        self._rescue(theta, np.where(miss)[0], rprec=0)

    # ------------------------------------------ lagged factors, legacy route
    def _lag_bytes(self):
        return self.B * self.n * self.n * (4 if self.lag_prec else 8)

    def _lag_store(self):
        """The per-slot lagged factors (allocated on first use), or None when
        they do not fit LAG_STORE_BYTES."""
        if self.lag_lu is None:
            if self._lag_bytes() > LAG_STORE_BYTES:
                return None
            B, n = self.B, self.n
            dt = torch.float32 if self.lag_prec else torch.float64
            self.lag_lu = torch.empty(B, n, n, dtype=dt, device=self.dev)
            self.lag_piv = torch.zeros(B, n, dtype=torch.int32, device=self.dev)
            self.lag_perm = torch.zeros(B, n, dtype=torch.int32, device=self.dev)
            self.lag_info = torch.zeros(B, dtype=torch.int32, device=self.dev)
            self.eta_lag = torch.zeros_like(self.eta_u)
        return self.lag_lu

    def _lag_refresh(self, sel):
        """Assemble and factorise the Schur matrix at the current eta for the
        selected runs and freeze that eta (LaggedStokes.refresh,
        lagged_stokes.py:274-300; RBCCoupled._refresh, coupled_model.py:575-581).
        Runs of consecutive slots factor as one batch in place; scattered
        runs (slots admitted at different times) are gathered into one batch
        and their factors copied into their slots, so the LU keeps its batch
        width."""
        runs = np.where(sel)[0]
        if runs.size == 0:
            return
        # transient matrices of a scattered refresh (at most RESCUE_BYTES) and
        # the FP64 assembly matrix of an FP32 store; a refresh of consecutive
        # slots factors in the store itself
        n = self.n
        es = 4 if self.lag_prec else 8
        scattered = runs[-1] - runs[0] + 1 != runs.size
        nb = min(runs.size, max(1, int(RESCUE_BYTES // (n * n * es)))) if scattered else 0
        with _RESCUE_GATE.hold(nb * n * n * es + (n * n * 8 if self.lag_prec else 0)):
            self._lag_refresh_runs(sel, runs)

    def _lag_refresh_runs(self, sel, runs):
        self._lag_store()
        L, st, n = self.L, self.stream(), self.n
        self._A64 = None
        self.n_refresh_batches += 1
        if runs[-1] - runs[0] + 1 == runs.size:
            g0, g1 = int(runs[0]), int(runs[-1]) + 1
            self._assemble_into(range(g0, g1), self.lag_lu[g0:g1])
            L.dense_lu(self.lag_prec, g1 - g0, n, self.p(self.lag_lu[g0]), n * n,
                       self.p(self.lag_piv[g0]), self.p(self.lag_info[g0:g1]), st)
        else:
            cap = max(1, int(RESCUE_BYTES // (n * n * self.lag_lu.element_size())))
            for c0 in range(0, runs.size, cap):
                chunk = runs[c0:c0 + cap]
                A = _transient((chunk.size, n, n), self.lag_lu.dtype, self.dev)
                piv = torch.empty(chunk.size, n, dtype=torch.int32, device=self.dev)
                info = torch.empty(chunk.size, dtype=torch.int32, device=self.dev)
                self._assemble_into(chunk, A)
                L.dense_lu(self.lag_prec, chunk.size, n, self.p(A), n * n, self.p(piv),
                           self.p(info), st)
                idx = torch.as_tensor(chunk, device=self.dev)
                self.lag_lu.index_copy_(0, idx, A)
                self.lag_piv.index_copy_(0, idx, piv)
                self.lag_info.index_copy_(0, idx, info)
                del A, piv, info
        self._A64 = None
        L.dense_lu_perm(self.B, n, self.p(self.lag_piv), self.p(self.lag_perm),
                        self.p(self._mask(sel)), st)
        rt = torch.as_tensor(runs, device=self.dev)
        self.eta_lag[rt] = self.eta_u[rt]
        self.lag_have[runs] = True
        self.n_since_refresh[runs] = 0
        self.n_refresh[runs] += 1
        self.force_refresh[runs] = False
        # a zero pivot: the factor of that run is not used; its solves go to
        # the FP64 factor of the rescue (mixed precision) and it is counted
        sing = runs[self.lag_info[rt].cpu().numpy() != 0]
        if sing.size:
            self.n_singular_factor[sing] += 1
            self.lag_have[sing] = False
            if self.rescue_prec == 1 and self.n <= DENSE_FP64_MAXN:
                self.fp32_backoff[sing] = RESCUE_FP32_BACKOFF

    def _assemble_into(self, runs, A):
        """Dense Schur matrices of ``runs`` at their current eta into the
        consecutive matrices of ``A`` (FP64 assembly, rounded once when A is
        FP32)."""
        for q, r in enumerate(runs):
            if A.dtype == self.dtype:
                self.assemble_schur(int(r), out=A[q])
            else:
                if self._A64 is None:
                    self._A64 = _transient((self.n, self.n), self.dtype, self.dev)
                self.assemble_schur(int(r), out=self._A64)
                A[q].copy_(self._A64)

    def _lag_drift(self, sel):
        """Pointwise relative eta drift since each selected run's frozen
        factorisation; inf where a run has none (coupled_model.py:509-522)."""
        out = np.full(self.B, np.inf)
        have = sel & self.lag_have
        if have.any():
            d = self.lag_red[0]
            self.L.eta_drift(self.prec, self.B, self.nru * self.Nz, self.p(self.eta_u),
                             self.p(self.eta_lag), self.p(d), self.p(self._mask(have)),
                             self.stream())
            out[have] = d.cpu().numpy()[have]
        return out

    def _defect(self, sel):
        """Picard defect correction with the lagged FP64 factor against the
        matrix-free operator at the current eta, for the selected runs at
        once: x0 = A_lag^-1 b, x += A_lag^-1 (b - A x) until the relative
        correction ||dx|| / ||x0|| <= defect_tol (max_K corrections), or K
        fixed corrections without a tolerance (RBCCoupled._defect_solve,
        coupled_model.py:524-573).  b is in ssol.b, x goes to ssol.x.
        -> the runs whose correction missed the tolerance."""
        L, st, B, n = self.L, self.stream(), self.B, self.n
        sv = self.ssol
        m = self._mask(sel)
        bn, xn, cn = self.lag_red[1], self.lag_red[2], self.lag_red[3]
        self._dense_solve(0, B, self.lag_lu, self.lag_perm, sv.b, sv.x, m, self.lag_piv)
        L.rownorm(self.prec, B, n, n, self.p(sv.b), self.p(bn), self.p(m), st)
        L.rownorm(self.prec, B, n, n, self.p(sv.x), self.p(xn), self.p(m), st)
        bnh = np.maximum(bn.cpu().numpy(), 1e-300)
        xnh = np.maximum(xn.cpu().numpy(), 1e-300)
        tol = self.defect_tol
        kmax = self.max_K if tol is not None else self.K
        live = sel.copy()
        K = np.zeros(B, dtype=np.int64)
        for _ in range(kmax):
            lm = self._mask(live)
            self.schur_matvec(sv.x, self.s2, lm)
            L.axpby(self.prec, n, B, 1.0, None, self.p(sv.b), -1.0, None, self.p(self.s2),
                    self.p(self.s1), self.p(lm), st)
            self._dense_solve(0, B, self.lag_lu, self.lag_perm, self.s1, self.s3, lm,
                              self.lag_piv)
            L.rownorm(self.prec, B, n, n, self.p(self.s3), self.p(cn), self.p(lm), st)
            L.axpby(self.prec, n, B, 1.0, None, self.p(sv.x), 1.0, None, self.p(self.s3),
                    self.p(sv.x), self.p(lm), st)
            c = cn.cpu().numpy() / xnh
            self.last_corr[live] = c[live]
            K[live] += 1
            if tol is not None:
                live &= ~(c <= tol)
                if not live.any():
                    break
        self.schur_matvec(sv.x, self.s2, m)
        L.axpby(self.prec, n, B, 1.0, None, self.p(sv.b), -1.0, None, self.p(self.s2),
                self.p(self.s1), self.p(m), st)
        L.rownorm(self.prec, B, n, n, self.p(self.s1), self.p(cn), self.p(m), st)
        res = cn.cpu().numpy() / bnh
        self.last_K[sel] = K[sel]
        self.last_defect_res[sel] = res[sel]
        self.n_defect[sel] += 1
        if tol is None:
            return np.zeros(B, dtype=bool)
        c = self.last_corr
        failed = sel & (~np.isfinite(c) | (c > tol))
        self.n_defect_failed[failed] += 1
        return failed

    def _lag_krylov(self, sel):
        """FGMRES in FP64 on the matrix-free Schur operator at the current
        eta, right-preconditioned by each selected run's lagged FP32 factor
        (the frozen-preconditioner FGMRES of LaggedStokes.solve_fgmres,
        lagged_stokes.py:345-360, with the dense factor as the preconditioner).
        b is in ssol.b, x goes to ssol.x.  -> (accepted, iterations, true
        relative residual) per run."""
        L, st, B, n = self.L, self.stream(), self.B, self.n
        sv = self.ssol
        rs = getattr(self, "rsol", None)
        if rs is None:
            rs = self.rsol = Solver(self, n, RESCUE_RESTART, RESCUE_MAXITER, RESCUE_RTOL)
        rs.b.copy_(sv.b)
        def precond(vin, zout, mm):
            # masked rows keep zout (a residual round holds the iterate there)
            self._dense_solve(2, B, self.lag_lu, self.lag_perm, vin, zout, mm, self.lag_piv)

        def op(zin, wout, mm):
            self.schur_matvec(zin, wout, mm)

        m = self._mask(sel)
        rs.solve(precond, op, m)
        it, info, res = rs.results()
        mb = m.bool()
        sv.x[mb] = rs.x[mb]
        ok = sel & np.isfinite(res) & (res <= sv.rtol)
        self.n_lag_krylov[sel] += 1
        self.rescue_iters[sel] += it[sel]
        self.last_defect_res[sel] = res[sel]
        return ok, it, res

    def _stokes_legacy(self, sel):
        """The legacy route for the selected runs (coupled_model.py:693-716):
        refresh the frozen factorisation when a run has none, when ``lag``
        solves have used it or when eta drifted more than the tolerance since;
        then the defect correction, and on failure a refresh at this eta and
        a second correction (a second failure is counted and accepted).  With
        the FP32 factor (precision 'mixed', or n above the FP64 factor's
        limit) FGMRES on the lagged factor takes the place of the Picard
        loop.  ssol.b holds the right-hand side; the solution goes to ssol.x."""
        drift = self._lag_drift(sel)
        ref = sel & (~self.lag_have | self.force_refresh | (self.n_since_refresh >= self.lag)
                     | (drift > self.eta_drift_tol))
        self._lag_refresh(ref)
        self.n_since_refresh[sel] += 1
        if self.lag_prec == 0:
            failed = self._defect(sel)
        else:
            ok, _, _ = self._lag_krylov(sel)
            failed = sel & ~ok
            self.n_defect_failed[failed] += 1
        if failed.any():
            self._lag_refresh(failed)
            self.n_since_refresh[failed] += 1
            if self.lag_prec == 0:
                failed2 = self._defect(failed)
            else:
                ok, _, _ = self._lag_krylov(failed)
                failed2 = failed & ~ok
                self.n_defect_failed[failed2] += 1
            self.n_defect_double_fail[failed2] += 1
        for r in np.where(sel)[0]:
            self.last_route[r] = "legacy"

    def stokes(self, theta, active=None):
        """(psi, v = -omega) of ``theta`` for the active runs into self.psi /
        self.vsol (and self.omega for const)."""
        if self.kind == KIND_CONST:
            self._stokes_const(theta, active)
            if self.general and theta is self.theta:
                self._stokes_const(self.theta_o, active, self.omega_o, self.psi_o)
        elif self.kind == KIND_DEPTH:
            self._stokes_depth(theta, active)
        else:
            # the coupled model counts fresh solves only (coupled_model.py:758-761)
            self._stokes_coupled_memo(theta, active)
            return
        act = np.ones(self.B, bool) if active is None else active.cpu().numpy().astype(bool)
        self.n_stokes[act] += 1

    def _stokes_coupled_memo(self, theta, active):
        """Reuse the last solve when the rfft signature V theta matches to 1e-12
        relative (coupled_model.py:725-765)."""
        B, N = self.B, self.Nz
        act = np.ones(B, bool) if active is None else active.cpu().numpy().astype(bool)
        self.zrows(theta, self.t["V"], self.memo_new, N, N, N, N)
        need = act.copy()
        if np.any(self.memo_valid & act):
            self.L.axpby(self.prec, self.n, B, 1.0, None, self.p(self.memo_new), -1.0, None,
                         self.p(self.memo_X), self.p(self.s2), None, self.stream())
            self.L.maxabs(self.prec, B, self.n, self.p(self.s2), self.p(self.umax), None,
                          self.stream())
            self.L.maxabs(self.prec, B, self.n, self.p(self.memo_X), self.p(self.umax2), None,
                          self.stream())
            d = self.umax.cpu().numpy()
            c = np.maximum(self.umax2.cpu().numpy(), 1e-300)
            hit = act & self.memo_valid & (d <= MEMO_RTOL * c)
            self.memo_hits[hit] += 1
            need = act & ~hit
        if np.any(need):
            mask = None if np.all(need) else torch.as_tensor(need.astype(np.int32), device=self.dev)
            self._stokes_coupled(theta, mask)
            m = torch.as_tensor(need, device=self.dev)
            self.memo_X[m] = self.memo_new[m]
            self.memo_valid |= need
            self.n_stokes[need] += 1

    # ------------------------------------------------- BDF2 implicit theta
    def theta_matvec(self, v, out, active):
        """a0 v - (D2 v - k^2 v) - advect(v) with tau rows (implicit_theta.py:296-383)."""
        L, t, B, Nk, N = self.L, self.t, self.B, self.Nk, self.Nz
        s = self.stream()
        self.zrows(v, t["V_VD"], self.zv, N, 2 * N, N, 2 * N, active=active)
        L.advect_prod(self.prec, B, Nk, self.nrp, N, self.p(self.zv), self.p(t["kx"]),
                      self.p(self.ux_p), self.p(self.uz_p), None, None, None, self.p(self.s1),
                      self.p(t["So_p"]), self.p(t["Ce_p"]), self.p(t["Ae_p"]), self.p(active), s)
        self.zrows(self.s1, t["Vi"], self.adv, N, N, N, N, active=active)
        self.zrows(v, t["D2"], self.d2v, N, N, N, N, active=active)
        L.theta_combine(self.prec, B, Nk, N, self.p(v), self.p(self.d2v), self.p(self.adv),
                        self.p(t["kx"]), self.p(self.a0), self.p(t["V0"]), self.p(t["V1"]),
                        self.p(out), self.p(active), s)

    def _set_step(self, dt):
        """BDF2 / BE coefficients per run (implicit_theta.py:215-238)."""
        dt = np.asarray(dt, dtype=np.float64)
        dt = np.where(dt > 0.0, dt, 1.0)      # runs outside the mask carry dt = 0
        a0 = np.empty(self.B)
        b0 = np.empty(self.B)
        b1 = np.zeros(self.B)
        for i in range(self.B):
            if not self.has_hist[i]:
                a0[i] = b0[i] = 1.0 / dt[i]
            else:
                r = dt[i] / self.dt_prev[i]
                a0[i] = (1.0 + 2.0 * r) / ((1.0 + r) * dt[i])
                b0[i] = (1.0 + r) / dt[i]
                b1[i] = r * r / ((1.0 + r) * dt[i])
        return a0, b0, b1

    def _step_implicit(self, dt, active):
        """RBCImplicitTheta.step (implicit_theta.py:523-535, 608-620)."""
        L, t, B, Nk, N, n = self.L, self.t, self.B, self.Nk, self.Nz, self.n
        s = self.stream()
        act = np.ones(B, bool) if active is None else active.cpu().numpy().astype(bool)
        dt = np.broadcast_to(np.asarray(dt, dtype=np.float64), (B,)).copy()
        th = torch.empty_like(self.theta)
        self.project(self.theta, th, active)             # th = project(that)
        t2 = torch.empty_like(self.theta)
        self.project(th, t2, active)                     # projected again
        self.stokes(t2, active)                          # psi_n (memo hit)
        psi_n = self.psi.clone()
        # extrapolated velocity psi* = 2 psi^n - psi^{n-1}
        ca = torch.tensor(np.where(self.has_hist, 2.0, 1.0), dtype=self.dtype, device=self.dev)
        cb = torch.tensor(np.where(self.has_hist, -1.0, 0.0), dtype=self.dtype, device=self.dev)
        psx = self.s2
        L.axpby(self.prec, n, B, 1.0, self.p(ca), self.p(psi_n), 1.0, self.p(cb),
                self.p(self.psi_prev), self.p(psx), self.p(active), s)
        psx = psx.clone()
        self._padded_velocity(psx, active)
        src = None
        if np.any(self.Dih != 0.0):
            eta_c = self.eta_c if self.kind == KIND_COUPLED else self.eta_c_depth
            self._phi_grid(psx, eta_c, self.gp4, active)
            self.xleft(t["Ae_p"], Nk, self.nrp, self.gp4, N, self.nrp * N, self.s3, N, n, N,
                       active=active)
            self.zrows(self.s3, t["Vi"], self.src, N, N, N, N, active=active)
            src = self.src
        a0, b0, b1 = self._set_step(dt)
        self.a0.copy_(torch.from_numpy(a0).to(self.dtype))
        self.b0.copy_(torch.from_numpy(b0).to(self.dtype))
        self.b1.copy_(torch.from_numpy(b1).to(self.dtype))
        sv = self.tsol
        hh = self.H if np.any(self.Hh != 0.0) else None
        L.theta_rhs(self.prec, B, Nk, N, self.p(t2), self.p(self.theta_prev), self.p(psx),
                    self.p(t["kx"]), self.p(self.b0), self.p(self.b1), self.p(hh), self.p(src),
                    self.p(sv.b), self.p(active), s)
        # Helmholtz factors (a0 + k^2 - D2) per run, refreshed when a0 changes
        refac = np.where(act & (self._a0_fact != a0))[0]
        if len(refac):
            mask = torch.zeros(B, dtype=torch.int32, device=self.dev)
            mask[torch.as_tensor(refac, device=self.dev)] = 1
            self.hlam.copy_(self.a0)
            L.tau_assemble(self.prec, 0, B, Nk, N, self.p(t["D2"]), self.p(t["kx"]),
                           self.p(self.hlam), None, None, None, None, None, self.p(self.helm),
                           self.p(mask), s)
            self._pc_factor(self.helm, self.helm_piv, mask)
            self._a0_fact[refac] = a0[refac]

        def precond(vin, zout, mask):
            self._pc_apply(self.helm, self.helm_piv, vin, zout, -1.0, mask)

        def op(zin, wout, mask):
            self.theta_matvec(zin, wout, mask)

        sv.solve(precond, op, active, key="theta")
        self.last_theta = sv.results()
        self._budget_means(self.theta_prev if np.any(self.has_hist) else None, th,
                           np.stack([a0, b0, b1], axis=1), active)
        self.project(sv.x, self.theta, active)           # output projected
        # history (theta^{n-1} = th, dt_prev, psi^{n-1} = psi_n)
        if active is None:
            self.theta_prev.copy_(th)
            self.psi_prev.copy_(psi_n)
        else:
            m = active.bool()
            self.theta_prev[m] = th[m]
            self.psi_prev[m] = psi_n[m]
        self.dt_prev[act] = dt[act]
        self.has_hist[act] = True

    def step(self, dt, active=None):
        """Advance every active run by its own dt (numpy array or scalar)."""
        dt = np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.B,)).copy()
        if self.kind == KIND_CONST:
            self._step_const(dt, active)
        else:
            self._step_implicit(dt, active)
        act = np.ones(self.B, bool) if active is None else active.cpu().numpy().astype(bool)
        self.n_steps[act] += 1
