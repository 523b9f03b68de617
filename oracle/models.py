"""The reference's steppers, restated (oracle copy).

TEST INFRASTRUCTURE: restatement used only as the checker (see ``oracle``).

* ``OracleRBC``        -- isoviscous explicit-advection / BE-diffusion step
  (rbc_galerkin.py:258-497), the ``viscosity_model='const'`` model;
* ``OracleDepth``      -- depth-only eta(z) Stokes + BDF2 implicit theta step
  (convection_modes.py:57-233, implicit_theta.py:129-620,
  vv_stokes.py:183-321), the ``'eta_z'`` model;
* ``OracleCoupled``    -- eta(T(x,z,t)) coupled Stokes (BFBT FGMRES route
  with the legacy / direct rescue, and the legacy lagged-LU route) + the same
  BDF2 theta step (coupled_model.py:123-765, stokes_coupled.py:132-325,
  lagged_stokes.py:86-343, bfbt.py:74-191), the ``'eta_T'`` / ``'arrhenius'``
  model.

All arithmetic follows the reference's operation order where it matters for
the last bits (matrix association order, complex arithmetic, the FGMRES
state machine).
"""
import numpy as np
import scipy.linalg as sla

from . import laws
from .krylov import fgmres
from .spectral import Basis, TauSolver

CFL_SAFETY_DEFAULT = 40.0        # cfl_policy.py:57
CFL_SAFETY_MAX = 80.0            # cfl_policy.py:59
DTMAX = 4.0e-5                   # _numerics.py:132
DTMIN = 1e-10                    # _numerics.py:134
DIRECT_RESCUE_MAX = 4096         # coupled_model.py:115


# --------------------------------------------------------------------------
# exact truncated Chebyshev product (vv_stokes.py:81-109, lagged_stokes.py:86-116)
# --------------------------------------------------------------------------
def cheb_mult_matrix(a, N):
    """``M[p, n] = (a_{p-n}[p>=n] + a_{p+n}[p+n<N] + a_{n-p}[p>0, n>=p]) / 2``."""
    a = np.asarray(a, dtype=float)
    P, Q = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    M = np.zeros((N, N))
    m1 = P >= Q
    M[m1] += 0.5 * a[(P - Q)[m1]]
    m2 = (P + Q) < N
    M[m2] += 0.5 * a[(P + Q)[m2]]
    m3 = (P > 0) & (Q >= P)
    M[m3] += 0.5 * a[(Q - P)[m3]]
    return M


def cheb_mult_complex(a, N):
    a = np.asarray(a)
    M = cheb_mult_matrix(a.real, N).astype(complex)
    if np.abs(a.imag).max() > 0.0:
        M = M + 1j * cheb_mult_matrix(a.imag, N)
    return M


def green_matrix(b, k):
    """``G_k = A^{-1} P``: Dirichlet Green operator of ``D^2 - k^2``
    (stokes_coupled.py:170-182, vv_stokes.py:281-296)."""
    N = b.Nz
    A = b.D2 - (k * k) * np.eye(N)
    A[-2, :] = b.V[0]
    A[-1, :] = b.V[-1]
    P = np.eye(N)
    P[-2:, :] = 0.0
    return sla.solve(A, P)


def reduced_operator(b, M, G, k):
    """``Lm M + 2 k^2 [D,[D,M]] G`` with the tau rows ``v(0), v(1)``
    (vv_stokes.py:266-307, bfbt.py:157-180)."""
    D, D2 = b.D, b.D2
    Comm = D2 @ M - 2.0 * (D @ M @ D) + M @ D2
    k2 = float(k) ** 2
    A = (D2 - k2 * np.eye(b.Nz)) @ M + 2.0 * k2 * (Comm @ G)
    A[-2:, :] = np.vstack([b.V[0], b.V[-1]])
    return A


# --------------------------------------------------------------------------
# viscous dissipation source (dissipation.py:234-487)
# --------------------------------------------------------------------------
def phi_padded(b, psi, eta_grid=None):
    k = b.kx[:, None]
    D = b.D
    psi = np.asarray(psi)
    cxz = (1j * k) * (D @ psi.T).T
    czz = (D @ (D @ psi.T)).T
    cxx = -(k ** 2) * psi
    exx, ezz, exz = cxz, -cxz, 0.5 * (czz - cxx)
    tr = exx + ezz
    gxx, gzz, gxz, gtr = (b.synth_pad(c) for c in (exx, ezz, exz, tr))
    dev = (gxx ** 2 + gzz ** 2 + 2.0 * gxz ** 2) - (1.0 / 3.0) * gtr ** 2
    if eta_grid is None:
        return 2.0 * dev
    e = np.asarray(eta_grid, dtype=float)
    if e.ndim == 1:
        e = np.tile(e[None, :], (b.Nx, 1))
    return 2.0 * b.synth_pad(b.coeffs(e)) * dev


def dissipation_source(b, psi, eta_grid, Di, Ra):
    c = (b.Vi @ b.trunc_pad(phi_padded(b, psi, eta_grid)).T).T
    c[:, -2:] = 0.0
    out = np.zeros((b.Nk, b.Nz), dtype=complex)
    out += (float(Di) / float(Ra)) * c
    return out


# --------------------------------------------------------------------------
# the isoviscous model (rbc_galerkin.py:258-497)
# --------------------------------------------------------------------------
class OracleRBC:
    kind = "const"

    def __init__(self, Ra=1e4, Nx=64, Nz=48, Lx=2.0, internal_heating=0.0,
                 dissipation_number=0.0):
        self.Ra = float(Ra)
        self.b = Basis(Nx, Nz, Lx)
        for name in ("Nx", "Nz", "Lx", "Nk", "kx", "x", "z", "V", "Vi", "D",
                     "wq", "Np"):
            setattr(self, name, getattr(self.b, name))
        self.Z = TauSolver(self.b)
        self.H = float(internal_heating)
        self.Di = float(dissipation_number)
        self.n_steps = 0
        self.n_stokes = 0

    # transforms
    def grid(self, cf):
        return self.b.grid(cf)

    def coeffs(self, g):
        return self.b.coeffs(g)

    def vol_mean(self, g):
        return self.b.vol_mean(g)

    def seed(self, amp=1e-3, m=1, n=1):
        X, Zg = np.meshgrid(self.x, self.z, indexing="ij")
        return self.coeffs(amp * np.cos(m * np.pi * X) * np.sin(n * np.pi * Zg))

    def stokes(self, that):
        psi = np.zeros_like(that, dtype=complex)
        om = np.zeros_like(that, dtype=complex)
        for i, k in enumerate(self.kx):
            if k == 0:
                continue
            om[i] = self.Z.solve(k, -1j * k * self.Ra * that[i], 0.0)
            psi[i] = self.Z.solve(k, -om[i], 0.0)
        self.n_stokes += 1
        return psi, om

    def velocity(self, psi):
        ux = self.grid(self.b.dz(psi))
        uz = self.grid(-1j * self.kx[:, None] * psi)
        return ux, uz

    def dissipation_eta(self):
        return None

    def explicit_rhs(self, that, psi):
        G = self.b.synth_pad
        thx = G(1j * self.kx[:, None] * that)
        thz = G(self.b.dz(that))
        ux = G(self.b.dz(psi))
        uz = G(-1j * self.kx[:, None] * psi)
        N = -(ux * thx + uz * thz) + uz
        if self.H != 0.0:
            N = N + self.H
        if self.Di != 0.0:
            N = N + (self.Di / self.Ra) * phi_padded(self.b, psi, None)
        return N

    def step(self, that, dt):
        psi, _ = self.stokes(that)
        N = self.explicit_rhs(that, psi)
        Nc = (self.Vi @ self.b.trunc_pad(N).T).T
        out = np.zeros_like(that, dtype=complex)
        for i, k in enumerate(self.kx):
            out[i] = self.Z.solve(k, -(that[i] / dt + Nc[i]), lam=1.0 / dt)
        self.n_steps += 1
        return out

    def cfl_dt(self, that, safety=0.5, dtmax=5e-3, dtmin=1e-10):
        psi, _ = self.stokes(that)
        ux, uz = self.velocity(psi)
        vmax = max(np.abs(ux).max(), np.abs(uz).max(), 1e-12)
        dzmin = float(np.diff(self.z).min())
        return float(np.clip(safety * min(self.Lx / self.Nx, dzmin) / vmax,
                             dtmin, dtmax))

    def diagnostics(self, that):
        psi, _ = self.stokes(that)
        ux, uz = self.velocity(psi)
        th = np.real(self.grid(that))
        dth = self.D @ that[0]
        urms = float(np.sqrt(max(self.vol_mean(ux ** 2 + uz ** 2), 0.0)))
        return dict(Nu_top=float(1.0 - (self.V[-1] @ dth).real),
                    Nu_bot=float(1.0 - (self.V[0] @ dth).real),
                    Nu_energy=float(1.0 + self.vol_mean(np.real(uz) * th)),
                    urms=urms,
                    umax=float(max(np.abs(ux).max(), np.abs(uz).max())),
                    amp=float(np.abs(th).max()))


# --------------------------------------------------------------------------
# depth-only Stokes + BDF2 implicit theta (convection_modes.py:57-233,
# implicit_theta.py:129-620, vv_stokes.py:183-321)
# --------------------------------------------------------------------------
class OracleDepth(OracleRBC):
    kind = "eta_z"

    def __init__(self, Ra=1e4, Nx=64, Nz=48, Lx=2.0, deta_T=1000.0,
                 internal_heating=0.0, dissipation_number=0.0,
                 eta_min=1e-3, eta_max=1e3, eta_smooth=0.0,
                 theta_rtol=1e-10, theta_maxiter=200, theta_restart=40,
                 theta_extrapolate=True, symmetrize=True):
        super().__init__(Ra=Ra, Nx=Nx, Nz=Nz, Lx=Lx,
                         internal_heating=internal_heating,
                         dissipation_number=dissipation_number)
        self.deta_T = float(deta_T)
        self.eta_clip_lo, self.eta_clip_hi = float(eta_min), float(eta_max)
        self.eta_smooth = float(eta_smooth)
        self.symmetrize = bool(symmetrize)
        eg = laws.eta_depth(self.z, self.deta_T)
        if self.eta_smooth > 0.0:
            eg, _ = laws.smooth_clip_eta(eg, self.eta_clip_lo,
                                         self.eta_clip_hi, self.eta_smooth)
        else:
            eg, _ = laws.clip_eta(eg, self.eta_clip_lo, self.eta_clip_hi)
        self.eta_grid = eg
        self.eta_c = self.Vi @ np.asarray(eg, dtype=float)
        self._G = [green_matrix(self.b, k) for k in self.kx]
        M = cheb_mult_matrix(self.eta_c, self.Nz)
        self._vv_lu = {}
        for i, k in enumerate(self.kx):
            if k != 0.0:
                self._vv_lu[i] = sla.lu_factor(
                    reduced_operator(self.b, M, self._G[i], k))
        # theta solver state (implicit_theta.py:137-212)
        self.rtol, self.maxiter, self.restart = (float(theta_rtol),
                                                 int(theta_maxiter),
                                                 int(theta_restart))
        self.extrapolate = bool(theta_extrapolate)
        self.k2 = (self.kx ** 2)[:, None]
        self.a0, self.b0, self.b1 = 1.0, 1.0, 0.0
        self.dt_cur = None
        self.psi = self.psi_n = None
        self.ux = self.uz = None
        self._phi = None
        self.that_prev = self.dt_prev = self.psi_prev = None
        self.last_iters = 0
        self.last_info = 0
        self.last_res = 0.0
        self.n_nonconv = 0
        self.sym_removed = 0.0

    def dissipation_eta(self):
        return self.eta_grid

    def project_mirror(self, that):
        """Keep the x-symmetric part (convection_modes.py:134-153)."""
        X = (self.V @ np.asarray(that).T).T
        self.sym_removed = float(np.abs(X.imag).max()
                                 / max(np.abs(X).max(), 1e-300))
        return (self.Vi @ X.real.T).T

    def stokes(self, that):
        psi = np.zeros_like(that, dtype=complex)
        om = np.zeros_like(that, dtype=complex)
        for i, k in enumerate(self.kx):
            if k == 0:
                continue
            bb = np.array(1j * float(k) * self.Ra
                          * np.asarray(that[i], dtype=complex))
            bb[-2:] = 0.0
            lu = self._vv_lu[i]
            u = sla.lu_solve(lu, bb.real) + 1j * sla.lu_solve(lu, bb.imag)
            psi[i] = self._G[i] @ u
            om[i] = -u
        self.n_stokes += 1
        return psi, om

    def reset_history(self):
        self.that_prev = self.dt_prev = self.psi_prev = None

    # -- implicit_theta.py:215-238 --
    def set_step(self, dt, dt_prev=None):
        self.dt_cur = float(dt)
        if dt_prev is None:
            self.a0 = self.b0 = 1.0 / self.dt_cur
            self.b1 = 0.0
            return
        r = self.dt_cur / float(dt_prev)
        self.a0 = (1.0 + 2.0 * r) / ((1.0 + r) * self.dt_cur)
        self.b0 = (1.0 + r) / self.dt_cur
        self.b1 = r * r / ((1.0 + r) * self.dt_cur)

    # -- implicit_theta.py:252-293 --
    def prepare(self, that, psi_prev=None, extrapolate=False):
        psi, _ = self.stokes(that)
        self.psi_n = psi
        if extrapolate and psi_prev is not None:
            psi = 2.0 * np.asarray(psi) - np.asarray(psi_prev)
        self.psi = psi
        self.ux = self.b.synth_pad(self.b.dz(psi))
        self.uz = self.b.synth_pad(-1j * self.kx[:, None] * psi)
        if self.Di != 0.0:
            self._phi = dissipation_source(self.b, psi, self.dissipation_eta(),
                                           self.Di, self.Ra)

    def lap(self, v):
        return (self.b.D2 @ np.asarray(v).T).T - self.k2 * v

    def advect(self, v):
        vx = self.b.synth_pad(1j * self.kx[:, None] * v)
        vz = self.b.synth_pad(self.b.dz(v))
        return (self.Vi @ self.b.trunc_pad(-(self.ux * vx + self.uz * vz)).T).T

    def matvec(self, v):
        v = np.asarray(v).reshape(self.Nk, self.Nz)
        out = self.a0 * v - self.lap(v) - self.advect(v)
        out[:, -2] = v @ self.V[0]
        out[:, -1] = v @ self.V[-1]
        return out.ravel()

    def rhs(self, that, dt, that_prev=None, dt_prev=None):
        self.set_step(dt, dt_prev)
        b = self.b0 * np.asarray(that)
        if self.b1 != 0.0:
            b = b - self.b1 * np.asarray(that_prev)
        b = b + (-1j * self.kx[:, None] * self.psi)
        b = b.copy()
        b[:, -2:] = 0.0
        if self.H != 0.0:
            b = np.array(b, copy=True)
            b[0, 0] += self.H
        if self.Di != 0.0:
            b = np.array(b, copy=True)
            b[:, :-2] += self._phi[:, :-2]
        return np.asarray(b).ravel()

    def theta_precond(self, r):
        R = np.asarray(r).reshape(self.Nk, self.Nz)
        out = np.zeros_like(R, dtype=complex)
        lam = float(self.a0)
        for i, k in enumerate(self.kx):
            out[i] = self.Z.solve(k, -R[i], lam=lam)
        return out.ravel()

    def theta_solve(self, that, dt, that_prev, dt_prev):
        b = self.rhs(that, dt, that_prev, dt_prev)
        x, info, it, _ = fgmres(self.matvec, b, M=self.theta_precond,
                                rtol=self.rtol, restart=self.restart,
                                maxiter=self.maxiter)
        bnorm = max(np.linalg.norm(b), 1e-300)
        self.last_iters, self.last_info = it, info
        self.last_res = float(np.linalg.norm(b - self.matvec(x.ravel()))) / bnorm
        if info != 0:
            self.n_nonconv += 1
        return x.reshape(self.Nk, self.Nz)

    # -- implicit_theta.py:523-535, 608-620 --
    def step(self, that, dt):
        th = self.project_mirror(that) if self.symmetrize else that
        use_x = self.extrapolate and self.psi_prev is not None
        t2 = self.project_mirror(th) if self.symmetrize else th
        self.prepare(t2, psi_prev=self.psi_prev, extrapolate=use_x)
        out = self.theta_solve(t2, dt, self.that_prev, self.dt_prev)
        self.n_steps += 1
        if self.symmetrize:
            out = self.project_mirror(out)
        self.that_prev = np.array(th, copy=True)
        self.dt_prev = float(dt)
        self.psi_prev = self.psi_n
        return out


# --------------------------------------------------------------------------
# coupled eta(T) model (coupled_model.py:123-765)
# --------------------------------------------------------------------------
class _SchurOps:
    """Per-eta pieces of the coupled Schur operator: eta coefficients, the
    Chebyshev multiplication matrices ``M_d`` and the matrix-free matvec
    (stokes_coupled.py:146-168, lagged_stokes.py:129-220)."""

    def __init__(self, model, eta_grid, partner_sign=-1.0):
        self.m = model
        self.b = model.b
        self.N, self.Nk = model.Nz, model.Nk
        self.kx = model.kx
        self.partner_sign = float(partner_sign)
        self.eta_c = model.coeffs(eta_grid)
        self._M = {}

    def M(self, d):
        if abs(d) >= self.Nk:
            return None
        Md = self._M.get(d)
        if Md is None:
            a = self.eta_c[d] if d >= 0 else np.conj(self.eta_c[-d])
            Md = cheb_mult_complex(a, self.N)
            self._M[d] = Md
        return Md

    def mult_eta(self, f, flip=False):
        """Both halves of the one-sided x-convolution (lagged_stokes.py:162-193)."""
        Nk, N = self.Nk, self.N
        out = np.zeros((Nk, N), dtype=complex)
        for d in range(-(Nk - 1), Nk):
            Md = self.M(d)
            if d >= 0:
                out[d:] += f[:Nk - d] @ Md.T
            else:
                out[:Nk + d] += f[-d:] @ Md.T
        sgn = -self.partner_sign if flip else self.partner_sign
        if sgn:
            for d in range(1, Nk):
                out[:d] += sgn * (f[1:d + 1][::-1] @ self.M(d).T)
        return out

    def green(self, w):
        out = np.zeros_like(w)
        for i in range(self.Nk):
            out[i] = self.m._G[i] @ w[i]
        return out

    def matvec(self, v):
        """Schur matvec (lagged_stokes.py:202-220)."""
        v = np.asarray(v).reshape(self.Nk, self.N)
        D, D2 = self.b.D, self.b.D2
        w = self.green(v)
        k2 = (self.kx ** 2)[:, None]
        t1 = self.mult_eta(v)
        t2 = self.mult_eta(2.0 * k2 * w)
        t3 = self.mult_eta((D2 @ w.T).T)
        t4 = self.mult_eta(self.kx[:, None] * (D @ w.T).T, flip=True)
        out = (D2 @ t1.T).T - k2 * t1
        out = out + (D2 @ t2.T).T
        out = out + 2.0 * k2 * t3
        out = out - 4.0 * self.kx[:, None] * (D @ t4.T).T
        out[:, -2] = v @ self.b.V[0]
        out[:, -1] = v @ self.b.V[-1]
        return out.ravel()

    def rhs_schur(self, that):
        N, Nk = self.N, self.Nk
        bb = np.zeros(Nk * N, dtype=complex)
        for i, k in enumerate(self.kx):
            if k == 0.0:
                continue
            bb[i * N:(i + 1) * N] = 1j * k * self.m.Ra * np.asarray(that[i])
            bb[i * N + N - 2] = 0.0
            bb[i * N + N - 1] = 0.0
        return bb

    def unpack(self, x):
        v = x.reshape(self.Nk, self.N)
        psi = np.zeros_like(v)
        for i in range(self.Nk):
            psi[i] = self.m._G[i] @ v[i]
        return psi, -v

    def assemble_schur(self):
        """Dense Schur matrix (stokes_coupled.py:242-289)."""
        N, Nk, kx = self.N, self.Nk, self.kx
        D, D2 = self.b.D, self.b.D2
        I = np.eye(N)
        cache = {}

        def mats(d):
            if d not in cache:
                Md = self.M(d)
                cache[d] = None if Md is None else (Md, D2 @ Md, Md @ D2,
                                                   D @ Md @ D)
            return cache[d]
        A = np.zeros((Nk * N, Nk * N), dtype=complex)
        for i in range(Nk):
            ki = kx[i]
            Lmi = D2 - ki * ki * I
            rows = slice(i * N, (i + 1) * N)
            for j in range(Nk):
                mm = mats(i - j)
                if mm is None:
                    continue
                Md, D2M, MD2, DMD = mm
                kj = kx[j]
                C = (2.0 * kj * kj * D2M + 2.0 * ki * ki * MD2
                     - 4.0 * ki * kj * DMD)
                A[rows, j * N:(j + 1) * N] = Lmi @ Md + C @ self.m._G[j]
            if self.partner_sign:
                for mm_ in range(1, Nk - i):
                    mm = mats(i + mm_)
                    if mm is None:
                        continue
                    Md, D2M, MD2, DMD = mm
                    kj = kx[mm_]
                    C = (2.0 * kj * kj * D2M + 2.0 * ki * ki * MD2
                         + 4.0 * ki * kj * DMD)
                    A[rows, mm_ * N:(mm_ + 1) * N] += self.partner_sign * (
                        Lmi @ Md + C @ self.m._G[mm_])
        for i in range(Nk):
            for r, vrow in ((N - 2, self.b.V[0]), (N - 1, self.b.V[-1])):
                row = i * N + r
                A[row, :] = 0.0
                A[row, i * N:(i + 1) * N] = vrow
        return A


def galerkin_profile(b, g):
    """``M_p^{-1} V^T W g`` with ``M_p = V^T W V`` (bfbt.py:74-96)."""
    V = np.asarray(b.V, dtype=float)
    w = np.asarray(b.wq, dtype=float)
    Mp = V.T @ (w[:, None] * V)
    return sla.solve(Mp, V.T @ (w * np.asarray(g, dtype=float)))


def profile_nodal(eta_grid, profile):
    """(bfbt.py:99-108)"""
    eta = np.asarray(eta_grid, dtype=float)
    if profile == "diag":
        return eta.mean(axis=0)
    if profile == "harm":
        return 1.0 / np.mean(1.0 / eta, axis=0)
    if profile == "geom":
        return np.exp(np.mean(np.log(eta), axis=0))
    raise ValueError(profile)


class OracleBFBT:
    """k-block-diagonal BFBT preconditioner (bfbt.py:129-191), partner block
    off (the reference default)."""

    def __init__(self, model, eta_grid, profile="harm"):
        b = model.b
        g = profile_nodal(eta_grid, profile)
        self.prof_c = np.real(galerkin_profile(b, g))
        M0 = cheb_mult_matrix(self.prof_c, b.Nz)
        D, D2 = b.D, b.D2
        Comm0 = D2 @ M0 - 2.0 * (D @ M0 @ D) + M0 @ D2
        self.N, self.Nk = b.Nz, b.Nk
        self.lus = []
        for i in range(b.Nk):
            k = float(b.kx[i])
            k2 = k * k
            Lm = D2 - k2 * np.eye(b.Nz)
            P = (Lm @ M0 + 2.0 * k2 * (Comm0 @ model._G[i])).astype(complex)
            P[-2:, :] = np.vstack([b.V[0], b.V[-1]])
            self.lus.append(sla.lu_factor(P))
        self._P = np.eye(b.Nz)
        self._P[-2:, :] = 0.0

    def apply(self, r):
        N = self.N
        out = np.zeros_like(r)
        for i in range(self.Nk):
            out[i * N:(i + 1) * N] = sla.lu_solve(self.lus[i],
                                                  self._P @ r[i * N:(i + 1) * N])
        return out


class OracleCoupled(OracleDepth):
    kind = "eta_T"

    def __init__(self, Ra=1e4, Nx=64, Nz=48, Lx=2.0, deta_T=1000.0,
                 internal_heating=0.0, dissipation_number=0.0,
                 eta_min=1e-3, eta_max=1e3, eta_smooth=0.0,
                 arrhenius_Ts=0.0, precond="legacy", precond_profile="harm",
                 lag=16, K=3, defect_tol=1e-8, max_K=8, stokes_rtol=1e-8,
                 eta_drift_tol=0.1, stokes_restart=80, stokes_maxiter=800,
                 theta_rtol=1e-10, symmetrize=True, stokes_mode="defect",
                 arrhenius_V=0.0, reference_anchor="D"):
        super().__init__(Ra=Ra, Nx=Nx, Nz=Nz, Lx=Lx, deta_T=deta_T,
                         internal_heating=internal_heating,
                         dissipation_number=dissipation_number,
                         eta_min=eta_min, eta_max=eta_max,
                         eta_smooth=eta_smooth, theta_rtol=theta_rtol,
                         symmetrize=symmetrize)
        self.kind = ("arrhenius" if float(arrhenius_Ts) > 0.0 or float(arrhenius_V) > 0.0
                     else "eta_T")
        self.arrhenius_Ts = float(arrhenius_Ts)
        # the pV law (coupled_model.py:186-229): Ts and E from the calibration
        self.pv = None
        if float(arrhenius_V) > 0.0:
            self.pv = laws.calibrate_pv(str(reference_anchor), float(arrhenius_V),
                                        deta_T=deta_T)
            self.arrhenius_Ts = self.pv[2]
        self.precond = str(precond).lower()
        self.precond_profile = str(precond_profile)
        self.stokes_mode = str(stokes_mode)
        self.lag, self.K, self.max_K = int(lag), int(K), int(max_K)
        self.defect_tol = defect_tol
        self.stokes_rtol = float(stokes_rtol)
        self.eta_drift_tol = float(eta_drift_tol)
        self.stokes_restart = int(stokes_restart)
        self.stokes_maxiter = int(stokes_maxiter)
        self.last_eta = None
        self._cache_key = self._cache_X = self._cache_val = None
        self.n_memo_hits = 0
        self.n_bfbt_solves = self.n_bfbt_iters = 0
        self.last_bfbt_iters = self.last_bfbt_info = 0
        self.last_bfbt_true_res = 0.0
        self.n_bfbt_rescue = self.n_bfbt_rescue_direct = 0
        self.n_bfbt_rescue_weak = 0
        self.last_stokes_route = "bfbt"
        self.n_clip_eta = self.n_clip_T = self.n_T_out_of_range = 0
        self.n_clip_eta_proj = 0
        self.n_smooth_eta = self.n_smooth_eta_proj = 0
        self.max_T_excursion = 0.0
        # legacy route state (lagged_stokes.py:229-300)
        self.lu_lag = None
        self.op_lag = None
        self.eta_lag = None
        self.n_since_refresh = 0
        self.n_refresh = 0
        self._force_refresh = False
        self._last_failed = False
        self.n_defect_failed = self.n_defect_double_fail = 0
        self.n_defect = 0
        self.last_K = 0
        self.last_corr = 0.0
        self.last_defect_res = 0.0

    def dissipation_eta(self):
        return self.last_eta if self.last_eta is not None else self.eta_grid

    # -- coupled_model.py:333-399 --
    def eta_from_T(self, T):
        T = np.asarray(T, dtype=float)
        lo, hi = float(T.min()), float(T.max())
        self.max_T_excursion = max(self.max_T_excursion, 0.0 - lo, hi - 1.0, 0.0)
        self.n_T_out_of_range += int(np.count_nonzero((T < 0.0) | (T > 1.0)))
        self.n_clip_T += int(np.count_nonzero((T < 0.0) | (T > 2.0)))
        Tc = np.clip(T, 0.0, 2.0)
        if self.pv is not None:
            # coupled_model.py:369-375: p0_norm = 1 - z on the rows of T
            E, V, Ts = self.pv
            eta = np.exp(laws.arrhenius_exponent_V(Tc, 1.0 - self.z[None, :], E, V, Ts))
        elif self.arrhenius_Ts > 0.0:
            eta = np.exp(laws.arrhenius_exponent(Tc, self.deta_T,
                                                 self.arrhenius_Ts))
        else:
            eta = laws.eta_exp_T(Tc, self.deta_T)
        if self.eta_smooth > 0.0:
            self.n_clip_eta += laws.n_clipped(eta, self.eta_clip_lo,
                                              self.eta_clip_hi)
            eta, n_sm = laws.smooth_clip_eta(eta, self.eta_clip_lo,
                                             self.eta_clip_hi, self.eta_smooth)
            self.n_smooth_eta += n_sm
            return eta
        eta, moved = laws.clip_eta(eta, self.eta_clip_lo, self.eta_clip_hi)
        self.n_clip_eta += moved
        return eta

    # -- coupled_model.py:401-496 --
    def eta_padded_truncated(self, that):
        Np, Nk = 2 * self.Nx, self.Nk
        sp = np.zeros((Np // 2 + 1, self.Nz), dtype=complex)
        sp[:Nk] = (self.V @ np.asarray(that).T).T
        th = np.fft.irfft(sp * Np, n=Np, axis=0)
        eta_p = self.eta_from_T(1.0 - self.z[None, :] + th)
        A = np.fft.rfft(eta_p, axis=0) / Np
        return np.fft.irfft(A[:Nk] * self.Nx, n=self.Nx, axis=0)

    def eta_of(self, that):
        eta = self.eta_padded_truncated(that)
        if float(np.asarray(eta).min()) > 0.0:
            return eta
        if self.eta_smooth > 0.0:
            self.n_clip_eta_proj += laws.n_clipped(eta, self.eta_clip_lo,
                                                   self.eta_clip_hi)
            eta, n_sm = laws.smooth_clip_eta(eta, self.eta_clip_lo,
                                             self.eta_clip_hi, self.eta_smooth)
            self.n_smooth_eta_proj += n_sm
            return eta
        eta, moved = laws.clip_eta(eta, self.eta_clip_lo, self.eta_clip_hi)
        self.n_clip_eta_proj += moved
        return eta

    # -- legacy route: lagged LU + Picard defect (coupled_model.py:509-581) --
    def _refresh(self, eta):
        op = _SchurOps(self, eta)
        self.lu_lag = sla.lu_factor(op.assemble_schur())
        self.op_lag = op
        self.eta_lag = np.asarray(eta, dtype=float)
        self.n_since_refresh = 0
        self.n_refresh += 1
        self._force_refresh = False

    def _eta_drift(self, eta):
        if self.eta_lag is None:
            return np.inf
        el = self.eta_lag
        return float((np.abs(eta - el) / np.maximum(np.abs(el), 1e-300)).max())

    def _defect_solve(self, that, eta):
        b = self.op_lag.rhs_schur(that)
        x = sla.lu_solve(self.lu_lag, b)
        Anew = _SchurOps(self, eta)
        bn = max(np.linalg.norm(b), 1e-300)
        xn = max(np.linalg.norm(x), 1e-300)
        tol = self.defect_tol
        kmax = self.max_K if tol is not None else self.K
        corr, res = [], []
        for _ in range(kmax):
            r = b - Anew.matvec(x)
            res.append(float(np.linalg.norm(r) / bn))
            dx = sla.lu_solve(self.lu_lag, r)
            c = float(np.linalg.norm(dx) / xn)
            corr.append(c)
            x = x + dx
            if tol is not None and c <= tol:
                break
        r = b - Anew.matvec(x)
        res.append(float(np.linalg.norm(r) / bn))
        self.last_K = len(corr)
        self.last_corr = corr[-1] if corr else 0.0
        self.last_defect_res = res[-1]
        self.n_defect += 1
        if tol is None:
            self._last_failed = False
        else:
            self._last_failed = bool(not np.isfinite(self.last_corr)
                                     or self.last_corr > tol)
            if self._last_failed:
                self.n_defect_failed += 1
        return x

    # -- BFBT route (coupled_model.py:588-654) --
    def _bfbt_solve(self, that, eta):
        op = _SchurOps(self, eta)
        b = op.rhs_schur(that)
        P = OracleBFBT(self, eta, profile=self.precond_profile)
        x, info, it, _ = fgmres(op.matvec, b, M=P.apply, rtol=self.stokes_rtol,
                                restart=self.stokes_restart,
                                maxiter=self.stokes_maxiter)
        self.n_bfbt_solves += 1
        self.n_bfbt_iters += int(it)
        self.last_bfbt_iters = int(it)
        self.last_bfbt_info = int(info)
        bn = max(float(np.linalg.norm(b)), 1e-300)
        res = float(np.linalg.norm(b - op.matvec(x)) / bn)
        self.last_bfbt_true_res = res
        self.last_K = 0
        self.last_defect_res = res
        return x, op, b, int(info), res

    def _rescue(self, that, eta, b, op):
        self._refresh(eta)
        x = self._defect_solve(that, eta)
        if self._last_failed:
            self._refresh(eta)
            x = self._defect_solve(that, eta)
            if self._last_failed:
                self.n_defect_double_fail += 1
        bn = max(float(np.linalg.norm(b)), 1e-300)
        res = float(np.linalg.norm(b - op.matvec(x)) / bn)
        self.last_defect_res = res
        if res <= self.stokes_rtol or not np.isfinite(res):
            return x, "bfbt->legacy"
        if self.Nk * self.Nz > DIRECT_RESCUE_MAX:
            self.n_bfbt_rescue_weak += 1
            return x, "bfbt->legacy(weak)"
        x2 = sla.lu_solve(sla.lu_factor(_SchurOps(self, eta).assemble_schur()), b)
        self.n_bfbt_rescue_direct += 1
        self.last_defect_res = float(np.linalg.norm(b - op.matvec(x2)) / bn)
        return x2, "bfbt->direct"

    def stokes_coupled(self, that):
        eta = self.eta_of(that)
        self.last_eta = eta
        if self.precond == "bfbt":
            x, op, b, info, res = self._bfbt_solve(that, eta)
            if info != 0 and not (res <= self.stokes_rtol):
                self.n_bfbt_rescue += 1
                x, route = self._rescue(that, eta, b, op)
            else:
                route = "bfbt"
            self.last_stokes_route = route
            psi, om = op.unpack(x)
        elif self.stokes_mode == "direct":
            op = _SchurOps(self, eta)
            x = sla.lu_solve(sla.lu_factor(op.assemble_schur()),
                             op.rhs_schur(that))
            psi, om = op.unpack(x)
            self.last_stokes_route = "direct"
        else:
            drift = self._eta_drift(eta)
            if (self.op_lag is None or self._force_refresh
                    or self.n_since_refresh >= self.lag
                    or drift > self.eta_drift_tol):
                self._refresh(eta)
            self.n_since_refresh += 1
            x = self._defect_solve(that, eta)
            if self._last_failed:
                self._refresh(eta)
                self.n_since_refresh += 1
                x = self._defect_solve(that, eta)
                if self._last_failed:
                    self.n_defect_double_fail += 1
            psi, om = self.op_lag.unpack(x)
            self.last_stokes_route = "legacy"
        return psi, om

    # -- coupled_model.py:725-765 --
    def stokes(self, that):
        that = np.asarray(that)
        key = that.tobytes()
        if key == self._cache_key:
            return self._cache_val
        X = (self.V @ that.T).T
        cX = self._cache_X
        if cX is not None and cX.shape == X.shape:
            den = max(float(np.abs(cX).max()), 1e-300)
            if float(np.abs(X - cX).max()) <= 1e-12 * den:
                self.n_memo_hits += 1
                return self._cache_val
        val = self.stokes_coupled(that)
        self.n_stokes += 1
        self._cache_key, self._cache_X, self._cache_val = key, X, val
        return val


# --------------------------------------------------------------------------
# config -> model recipes (models/const.py:37-57, eta_z.py:32-59, eta_T.py:44-102)
# --------------------------------------------------------------------------
def build_oracle(cfg, **model_kw):
    """``cfg``: a mapping with ConvectionConfig field names (or an object
    with ``as_dict()``); returns the oracle model the reference would build.
    ``model_kw``: constructor arguments of the coupled model beyond the
    config's (``stokes_maxiter``, ``stokes_rtol``), as the reference's
    self-tests pass them (coupled_model.py:955-972)."""
    if hasattr(cfg, "as_dict"):
        cfg = cfg.as_dict()
    c = dict(cfg)
    vm = str(c.get("viscosity_model", "eta_T")).strip().lower()
    vm = {"constant": "const", "isoviscous": "const", "etaz": "eta_z",
          "eta_t": "eta_T", "etat": "eta_T"}.get(vm, vm)
    common = dict(Ra=float(c.get("Ra", 1e4)), Nx=int(c.get("Nx", 64)),
                  Nz=int(c.get("Nz", 96)), Lx=float(c.get("Lx", 2.0)),
                  internal_heating=float(c.get("internal_heating", 0.0)),
                  dissipation_number=float(c.get("dissipation_number", 0.0)))
    if vm == "const":
        return OracleRBC(**common)
    win = dict(eta_min=float(c.get("eta_min", 1e-3)),
               eta_max=float(c.get("eta_max", 1e3)),
               eta_smooth=float(c.get("eta_smooth", 0.0)),
               deta_T=float(c.get("deta_T", 1000.0)),
               symmetrize=bool(c.get("symmetrize", True)))
    if vm == "eta_z":
        return OracleDepth(**common, **win)
    dtol = float(c.get("defect_tol", 1e-8))
    return OracleCoupled(**common, **win,
                         arrhenius_Ts=float(c.get("arrhenius_Ts", 0.0)),
                         arrhenius_V=float(c.get("arrhenius_V", 0.0)),
                         reference_anchor=str(c.get("reference_anchor", "D")),
                         precond=str(c.get("precond", "legacy")).strip().lower(),
                         lag=int(c.get("lag", 16)), K=int(c.get("K", 3)),
                         defect_tol=(None if dtol <= 0.0 else dtol),
                         max_K=int(c.get("max_K", 8)), **model_kw)


def clamp_safety(s):
    return float(np.clip(float(s), 0.0, CFL_SAFETY_MAX))


def policy_dt(model, that, safety=CFL_SAFETY_DEFAULT, dtmax=DTMAX, dtmin=DTMIN):
    """(cfl_policy.py:64-74)"""
    fac = clamp_safety(safety)
    return float(min(dtmax, model.cfl_dt(that, safety=fac, dtmax=dtmax,
                                         dtmin=dtmin)))


def march_adaptive(model, that, tstop, safety=None, dtmax=None, on_step=None):
    """(api.py:169-190)"""
    safety = CFL_SAFETY_DEFAULT if safety is None else float(safety)
    dtmax = DTMAX if dtmax is None else float(dtmax)
    s = np.array(that, copy=True)
    t, n, dt = 0.0, 0, float("nan")
    while t < float(tstop):
        dt = float(policy_dt(model, s, safety=safety, dtmax=dtmax))
        s = model.step(s, dt)
        t += dt
        n += 1
        if on_step is not None:
            on_step(n, t, s)
    return s, dict(steps=n, t=t, dt=dt, safety=safety, dtmax=dtmax)


def march(model, that, dt, nsteps, on_step=None):
    """(api.py:147-166)"""
    s = np.array(that, copy=True)
    for n in range(int(nsteps)):
        s = model.step(s, float(dt))
        if on_step is not None:
            on_step(n + 1, (n + 1) * float(dt), s)
    return s, dict(steps=int(nsteps), t=int(nsteps) * float(dt), dt=float(dt))
