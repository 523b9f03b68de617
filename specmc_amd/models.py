"""Device-backed model classes with the reference's model protocol.

``api.build_model`` returns one of

* ``RBC``              -- ``viscosity_model='const'``  (rbc_galerkin.py:258-497)
* ``RBCImplicitTheta`` -- ``viscosity_model='eta_z'``  (implicit_theta.py:538-638)
* ``RBCCoupled``       -- ``'eta_T'`` / ``'arrhenius'`` (coupled_model.py:123-821)

and every one of them offers what ``api`` / ``cfl_policy`` / ``bench_ic`` and
user scripts call (§ "drop-in boundary" of the design notes): ``step``,
``cfl_dt``, ``stokes``, ``velocity``, ``diagnostics``, ``seed``, ``coeffs``,
``grid``, ``describe``, ``reset_history`` and the attributes ``x, z, kx, Nk,
Nx, Nz, Lx, V, Vi, D, wq, n_steps``.

The state crosses the boundary as a host numpy array (the reference's
convention); the device copy is reused as long as the caller hands back the
array the model returned last, so a march makes no host-to-device copy per
step.  The stepper runs on the HIP device only.  The isoviscous model takes
any state, as RBC does: one with an x-antisymmetric part (imaginary
coefficients) is carried as two real channels.  The variable-viscosity
models carry mirror-symmetric states (real x-mode coefficients up to
roundoff); their ``step`` with ``symmetrize=True`` projects any other state
onto that subspace first, as the reference's does (convection_modes.py:198,
implicit_theta.py:527, 610), and their other calls refuse one.
``seed`` / ``coeffs`` / ``grid`` are host-side input/output helpers (the
rfft conventions of rbc_galerkin.py:322-330), not part of the step.
"""
import numpy as np
import torch

from . import _numerics as NUM
from .engine import (ANTISYM_TOL, Engine, KIND_CONST, KIND_DEPTH, KIND_COUPLED,
                     antisym_part)

SYM_TOL = ANTISYM_TOL


class _ThetaInfo:
    """Last implicit theta solve (the reference's ``model.theta`` fields)."""

    def __init__(self, rtol, restart):
        self.rtol, self.restart = rtol, restart
        self.last_iters = 0
        self.last_info = 0
        self.last_res = 0.0
        self.n_nonconv = 0
        self.n_solves = 0
        self.scheme = "bdf2"


class _DeviceModel:
    kind = None
    prec = 0

    def __init__(self, Ra=1e4, Nx=64, Nz=48, Lx=2.0, deta_T=1.0,
                 internal_heating=0.0, dissipation_number=0.0,
                 eta_min=1e-3, eta_max=1e3, eta_smooth=0.0, arrhenius_Ts=0.0,
                 arrhenius_V=0.0, reference_anchor="D",
                 precond="bfbt", symmetrize=True, stokes_maxiter=None, stokes_rtol=None,
                 defect_tol=1e-8, max_K=8, K=3, pc_apply=None, ortho=None,
                 precision="fp64", graph_rounds=0, **extra):
        self.Ra, self.Nx, self.Nz, self.Lx = float(Ra), int(Nx), int(Nz), float(Lx)
        self.internal_heating = float(internal_heating)
        self.dissipation_number = float(dissipation_number)
        self.deta_T = float(deta_T)
        self.symmetrize = bool(symmetrize)
        self.sym_removed = 0.0       # antisymmetric part removed by the last step's projection
        self.extra = dict(extra)
        params = dict(Ra=self.Ra, deta_T=self.deta_T, Ts=float(arrhenius_Ts),
                      arrhenius_V=float(arrhenius_V), reference_anchor=str(reference_anchor),
                      eta_min=float(eta_min), eta_max=float(eta_max),
                      eta_smooth=float(eta_smooth),
                      internal_heating=self.internal_heating,
                      dissipation_number=self.dissipation_number,
                      precond=str(precond).strip().lower(), lag=int(extra.get("lag", 16)))
        kw = dict(defect_tol=defect_tol, max_K=int(max_K), K=int(K), precision=precision,
                  graph_rounds=int(graph_rounds))
        if stokes_maxiter is not None:
            kw["stokes_maxiter"] = int(stokes_maxiter)
        if stokes_rtol is not None:
            kw["stokes_rtol"] = float(stokes_rtol)
        if pc_apply is not None:
            kw["pc_apply"] = pc_apply
        if ortho is not None:
            kw["ortho"] = ortho
        self.eng = Engine(self.kind, self.Nx, self.Nz, self.Lx, [params],
                          prec=self.prec, **kw)
        o = self.eng.ops
        self.kx, self.Nk, self.x, self.z = o.kx, o.Nk, o.x, o.z
        self.V, self.Vi, self.D, self.wq, self.Np = o.V, o.Vi, o.D, o.wq, 2 * self.Nx
        self._dev_key = None
        self.theta = _ThetaInfo(NUM.THETA_RTOL, NUM.THETA_RESTART)

    # ---- counters
    @property
    def n_steps(self):
        return int(self.eng.n_steps[0])

    @property
    def n_stokes(self):
        return int(self.eng.n_stokes[0])

    # ---- host transforms (rbc_galerkin.py:317-330)
    def grid(self, cf):
        return np.fft.irfft((self.V @ np.asarray(cf).T).T * self.Nx, n=self.Nx, axis=0)

    def coeffs(self, g):
        A = np.fft.rfft(np.asarray(g), axis=0) / self.Nx
        return (self.Vi @ A.T).T

    def vol_mean(self, g):
        return float(np.sum(self.wq * np.asarray(g).real.mean(axis=0)))

    def seed(self, amp=1e-3, m=1, n=1):
        X, Z = np.meshgrid(self.x, self.z, indexing="ij")
        return self.coeffs(amp * np.cos(m * np.pi * X) * np.sin(n * np.pi * Z))

    # ---- state transfer
    def _upload(self, that, project=False):
        """Hand ``that`` to the device.  ``project``: the caller is ``step``,
        which on a variable-viscosity model with ``symmetrize=True`` starts
        from project_mirror(that) (convection_modes.py:134-153, 198)."""
        a = np.asarray(that)
        if a.shape != (self.Nk, self.Nz):
            raise ValueError("state must have shape (Nk, Nz) = (%d, %d), got %r"
                             % (self.Nk, self.Nz, a.shape))
        key = a.tobytes()
        if project and self.kind != KIND_CONST and self.symmetrize:
            X = self.V @ a.T
            self.sym_removed = float(np.abs(np.imag(X)).max()
                                     / max(float(np.abs(X).max()), 1e-300))
        if key == self._dev_key:
            return
        anti = antisym_part(a)
        if anti <= SYM_TOL:
            a = np.real(a)
        elif self.kind != KIND_CONST:
            if not (project and self.symmetrize):
                raise ValueError(
                    "a %s model carries mirror-symmetric states (real x-mode "
                    "coefficients); this state has a relative antisymmetric part "
                    "of %.2e. Only step() with symmetrize=True accepts it, and "
                    "projects it as the reference does" % (self.kind, anti))
            a = np.real(a)       # the device step applies Vi Re(V .) to it
        self.eng.set_theta(a[None])
        self._dev_key = key

    def _resume_from_device(self):
        """The state held on the device (after a restore), as the array the
        next ``step`` call must be handed."""
        return self._download()

    def _download(self):
        out = self.eng.get_theta()[0]
        if self.kind == KIND_CONST:
            out = out.astype(complex)
        self._dev_key = out.tobytes()
        return out

    # ---- protocol
    def cfl_dt(self, that, safety=0.5, dtmax=5e-3, dtmin=1e-10):
        """dt = clip(safety min(Lx/Nx, dz_min) / max|u|, dtmin, dtmax)
        (rbc_galerkin.py:465-471)."""
        self._upload(that)
        return float(self.eng.cfl(safety, dtmax, dtmin)[0])

    def step(self, that, dt):
        self._upload(that, project=True)
        self.eng.step(float(dt))
        self._after_step()
        return self._download()

    def _after_step(self):
        if self.kind != KIND_CONST:
            it, info, res = self.eng.last_theta
            th = self.theta
            th.last_iters, th.last_info, th.last_res = int(it[0]), int(info[0]), float(res[0])
            th.n_solves += 1
            if info[0] != 0:
                th.n_nonconv += 1

    def stokes(self, that):
        """(psi, omega) coefficients of ``that`` (complex, (Nk, Nz))."""
        self._upload(that)
        e = self.eng
        e.stokes(e.theta)
        psi = 1j * e.get(e.psi)[0]
        if self.kind == KIND_CONST:
            om = 1j * e.get(e.omega)[0]
            if e.general:       # channel b: psi = -psi_o, omega = -omega_o
                psi = psi - e.get(e.psi_o)[0]
                om = om - e.get(e.omega_o)[0]
        else:
            om = -1j * e.get(e.vsol)[0]
        return psi, om

    def velocity(self, psi):
        """(ux, uz) on the (Nx, Nz) grid for stream-function coefficients."""
        e = self.eng
        p = np.asarray(psi)
        dev = lambda v: torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).reshape(
            1, -1).to(device=e.dev, dtype=e.dtype)
        e.psi.copy_(dev(np.imag(p)))
        # a real part of psi is the stream function of an x-antisymmetric state
        rel = (float(np.abs(np.real(p)).max()) / max(float(np.abs(p).max()), 1e-300)
               if np.iscomplexobj(p) and p.size else 1.0 if np.any(p) else 0.0)
        if rel <= SYM_TOL:
            e.velocity_unpadded(e.psi)
            return self._mirror(e.ux_u, odd=True), self._mirror(e.uz_u, odd=False)
        if self.kind != KIND_CONST:
            raise ValueError("a %s model carries mirror-symmetric stream functions "
                             "(imaginary coefficients); this one has a relative real part "
                             "of %.2e" % (self.kind, rel))
        e._ensure_general()
        e.psi_o.copy_(dev(-np.real(p)))
        e.velocity_unpadded(e.psi, psi_o=e.psi_o)
        return (self._full(e.ux_ue, e.ux_u), self._full(e.uz_u, e.uz_uo))

    def _full(self, even, odd):
        """The (Nx, Nz) grid of a field from its x-even and x-odd parts on
        the half grid: e + o at x_j, e - o at x_{-j}."""
        h = lambda t: t.detach().to("cpu", torch.float64).numpy().reshape(
            self.Nx // 2 + 1, self.Nz)
        e, o = h(even), h(odd)
        full = np.empty((self.Nx, self.Nz))
        full[: self.Nx // 2 + 1] = e + o
        full[self.Nx // 2 + 1:] = (e - o)[1: self.Nx // 2][::-1]
        return full

    def _mirror(self, half, odd):
        h = half.detach().to("cpu", torch.float64).numpy().reshape(self.Nx // 2 + 1, self.Nz)
        full = np.empty((self.Nx, self.Nz))
        full[: self.Nx // 2 + 1] = h
        tail = h[1: self.Nx // 2][::-1]
        full[self.Nx // 2 + 1:] = -tail if odd else tail
        return full

    def diagnostics(self, that):
        """Nu_top, Nu_bot, Nu_energy, urms, umax, amp (rbc_galerkin.py:439-453)."""
        self._upload(that)
        d = self.eng.diagnostics()[0]
        return dict(Nu_top=float(d[0]), Nu_bot=float(d[1]), Nu_energy=float(d[2]),
                    urms=float(d[3]), umax=float(d[4]), amp=float(d[5]))

    def reset_history(self):
        self.eng.reset_history()

    def energy_budget(self):
        """The discrete energy budget of the last step, evaluated on the
        device from the model's own states (heating.balance with the
        dissipation term of dissipation.balance, heating.py:279-306,
        dissipation.py:540-571), plus Ra <u_z theta> (dissipation.py:423-444).
        -> dict of floats: storage, nu_bot, nu_top, heating, advection,
        dissipation, residual, theta_mean, T_mean, pi_theta, phi_integral."""
        return {k: float(v[0]) for k, v in self.eng.energy_budget().items()}

    def project_mirror(self, that):
        X = (self.V @ np.asarray(that).T).T
        return (self.Vi @ X.real.T).T

    def describe(self):
        e = self.eng
        c = e.counters.cpu().numpy()[0]
        d = dict(device="hip", kind=self.kind, Nx=self.Nx, Nz=self.Nz, Lx=self.Lx,
                 Ra=self.Ra, deta_T=self.deta_T, internal_heating=self.internal_heating,
                 dissipation_number=self.dissipation_number, n_steps=self.n_steps,
                 n_stokes=self.n_stokes, precision=e.precision if self.prec == 0 else "fp32")
        if self.kind != KIND_CONST:
            th = self.theta
            d.update(theta_scheme=NUM.THETA_SCHEME, theta_rtol=NUM.THETA_RTOL,
                     n_theta_nonconv=th.n_nonconv, last_theta_iters=th.last_iters)
        if self.kind == KIND_COUPLED:
            d.update(precond="legacy" if e.legacy[0] else "bfbt",
                     stokes_mode="defect", lag=int(e.lag[0]), K=e.K, max_K=e.max_K,
                     defect_tol=e.defect_tol, eta_drift_tol=e.eta_drift_tol,
                     n_refresh=int(e.n_refresh[0]), n_defect=int(e.n_defect[0]),
                     last_K=int(e.last_K[0]),
                     n_defect_double_fail=int(e.n_defect_double_fail[0]),
                     n_T_out_of_range=int(c[0]), n_clip_T=int(c[1]),
                     n_clip_eta=int(c[2]), max_T_excursion=float(c[3]),
                     n_clip_eta_proj=int(c[4]), n_smooth_eta=int(c[5]),
                     n_smooth_eta_proj=int(c[6]), n_memo_hits=int(e.memo_hits[0]),
                     n_bfbt_iters=int(e.bfbt_iters[0]),
                     last_bfbt_iters=int(e.last_bfbt_iters[0]),
                     last_bfbt_info=int(e.last_bfbt_info[0]),
                     last_bfbt_true_res=float(e.last_bfbt_res[0]),
                     n_bfbt_rescue=int(e.n_rescue[0]),
                     n_bfbt_rescue_direct=int(e.n_rescue_direct[0]),
                     n_bfbt_rescue_weak=int(e.n_rescue_weak[0]),
                     n_defect_failed=int(e.n_defect_failed[0]),
                     n_rescue_fp32=int(e.n_rescue_fp32[0]),
                     n_rescue_fp32_redo=int(e.n_rescue_fp32_redo[0]),
                     n_rescue_fp32_iters=int(e.rescue_iters[0]),
                     n_rescue_carry=int(e.n_rescue_carry[0]),
                     n_bfbt_probe=int(e.n_bfbt_probe[0]),
                     # mixed precision: n_bfbt_rescue counts the rescues after
                     # a rejected full or probe BFBT attempt, n_rescue_carry
                     # the solves that went to the factor route without one;
                     # their sum is what the reference's n_bfbt_rescue counts
                     # when every solve attempts BFBT first
                     n_rescued_solves=int(e.n_rescue[0] + e.n_rescue_carry[0]),
                     n_singular_factor=int(e.n_singular_factor[0]),
                     last_defect_res=float(e.last_defect_res[0]),
                     last_stokes_route=e.last_route[0])
        return d


class RBC(_DeviceModel):
    """Isoviscous model: explicit dealiased advection + BE diffusion
    (rbc_galerkin.py:258-497)."""
    kind = KIND_CONST


class RBCImplicitTheta(_DeviceModel):
    """Depth-only eta(z) Stokes + BDF2 implicit theta (implicit_theta.py:538-638)."""
    kind = KIND_DEPTH

    @property
    def eta_grid(self):
        return self.eng.eta_grid_h[0].copy()


class RBCCoupled(_DeviceModel):
    """eta(T(x,z,t)) coupled Stokes through the BFBT-preconditioned FGMRES
    route or the legacy lagged-LU route + BDF2 implicit theta
    (coupled_model.py:123-821)."""
    kind = KIND_COUPLED

    @property
    def precond(self):
        return "legacy" if self.eng.legacy[0] else "bfbt"

    @property
    def lag(self):
        return int(self.eng.lag[0])

    @property
    def n_refresh(self):
        return int(self.eng.n_refresh[0])

    @property
    def n_since_refresh(self):
        return int(self.eng.n_since_refresh[0])

    @property
    def n_defect(self):
        return int(self.eng.n_defect[0])

    @property
    def n_defect_failed(self):
        return int(self.eng.n_defect_failed[0])

    @property
    def n_defect_double_fail(self):
        return int(self.eng.n_defect_double_fail[0])

    @property
    def last_K(self):
        return int(self.eng.last_K[0])

    @property
    def last_defect_res(self):
        return float(self.eng.last_defect_res[0])

    @property
    def last_bfbt_iters(self):
        return int(self.eng.last_bfbt_iters[0])

    @property
    def n_bfbt_iters(self):
        return int(self.eng.bfbt_iters[0])

    @property
    def n_memo_hits(self):
        return int(self.eng.memo_hits[0])

    @property
    def n_bfbt_rescue(self):
        return int(self.eng.n_rescue[0])

    @property
    def last_stokes_route(self):
        return self.eng.last_route[0]

    def eta_of(self, that):
        """eta on the (Nx, Nz) grid as the Stokes operator consumes it
        (coupled_model.py:401-496)."""
        self._upload(that)
        e = self.eng
        e._eta_of(e.theta, None)
        return self._mirror(e.eta_u, odd=False)
