"""Per-shape operators, computed once on the host in FP64 and kept in HBM.

Everything here is shape data (it depends on ``Nx, Nz, Lx`` only) and is built
at model construction, never inside a step (rbc_galerkin.py:259-315 builds the
same pieces in ``RBC.__init__``).  The half-grid DFT matrices replace
``np.fft.rfft/irfft`` on the mirror-symmetric fields:

* padded grid, ``P = 2 Nx`` points, mirror half ``j = 0..Nx`` (``nrp`` rows):
  ``Ce`` / ``So`` synthesise real (x-even) / imaginary (x-odd) mode
  coefficients exactly as ``irfft(sp * P, n=P)`` of ``_synth_pad`` does
  (rbc_galerkin.py:332-336); ``Ae`` / ``Ao`` are ``rfft(g)/P`` truncated to
  ``Nk`` modes (``_trunc_pad``, rbc_galerkin.py:338-339);
* unpadded grid, ``Nx`` points, half ``j = 0..Nx/2`` (``nru`` rows): ``grid`` /
  ``coeffs`` (rbc_galerkin.py:322-330), the Nyquist bin counted once;
* z-padded product grid with ``Q = ceil((3 Nz - 2)/2) + 1`` CGL nodes: ``Vp``
  evaluates a Chebyshev series there and ``Tp`` returns the first ``Nz``
  coefficients of a degree ``< 2 Nz - 1`` product exactly, so that
  ``Tp (Vp a * Vp f)`` is the truncated Chebyshev product ``M(a) f`` of
  vv_stokes.py:81-109.
"""
import numpy as np
import scipy.linalg as sla
from numpy.polynomial import chebyshev as _C


def cgl(N):
    return 0.5 * (1.0 - np.cos(np.pi * np.arange(N) / (N - 1)))


def cheb_V(z, N):
    return _C.chebvander(2.0 * np.asarray(z, float) - 1.0, N - 1)


def cheb_D(N):
    D = np.zeros((N, N))
    for k in range(1, N):
        c = np.zeros(N)
        c[k] = 1.0
        d = 2.0 * _C.chebder(c)
        D[:len(d), k] = d
    return D


def _cos_sin(m, j, P):
    """cos / sin of 2 pi m j / P with the argument reduced mod P (exact phase)."""
    ph = (np.outer(j, m) % P) * (2.0 * np.pi / P)
    return np.cos(ph), np.sin(ph)


class ShapeOps:
    """Host FP64 operators of one grid shape."""

    def __init__(self, Nx, Nz, Lx):
        Nx, Nz = int(Nx), int(Nz)
        if Nx % 2:
            raise ValueError("Nx must be even (got %d)" % Nx)
        self.Nx, self.Nz, self.Lx = Nx, Nz, float(Lx)
        N = Nz
        self.kx = 2.0 * np.pi * np.fft.rfftfreq(Nx, d=self.Lx / Nx)
        self.Nk = Nk = len(self.kx)
        self.x = np.arange(Nx) * self.Lx / Nx
        self.z = cgl(N)
        self.V = cheb_V(self.z, N)
        self.Vi = np.linalg.inv(self.V)
        self.D = cheb_D(N)
        self.D2 = self.D @ self.D
        self.VD = self.V @ self.D
        self.V0 = self.V[0].copy()
        self.V1 = self.V[-1].copy()
        gk = np.array([0.0 if (k % 2) else 1.0 / (1.0 - k * k) for k in range(N)])
        gk[0] = 1.0
        self.wq = self.Vi.T @ gk
        self.dzmin = float(np.diff(self.z).min())
        self.hcfl = min(self.Lx / Nx, self.dzmin)
        # Dirichlet Green operators G_k = A_k^{-1} P (stokes_coupled.py:170-182)
        G = np.empty((Nk, N, N))
        Pm = np.eye(N)
        Pm[-2:, :] = 0.0
        for i, k in enumerate(self.kx):
            A = self.D2 - (k * k) * np.eye(N)
            A[-2, :] = self.V0
            A[-1, :] = self.V1
            G[i] = sla.solve(A, Pm)
        self.G = G
        # Galerkin profile projection M_p^{-1} V^T W (bfbt.py:74-96)
        Mp = self.V.T @ (self.wq[:, None] * self.V)
        self.Qprof = sla.solve(Mp, self.V.T * self.wq[None, :])
        # padded x grid (P = 2 Nx), mirror half j = 0..Nx
        P = 2 * Nx
        self.nrp = Nx + 1
        m = np.arange(Nk)
        jp = np.arange(self.nrp)
        c, s = _cos_sin(m, jp, P)
        w = np.where(m == 0, 1.0, 2.0)
        self.Ce_p = c * w[None, :]
        self.So_p = -2.0 * s
        self.So_p[:, 0] = 0.0
        u = np.where((jp == 0) | (jp == Nx), 1.0, 2.0)
        self.Ae_p = (c * u[:, None]).T / P
        self.Ao_p = -(2.0 / P) * s.T
        self.Ao_p[:, 0] = 0.0
        self.Ao_p[:, -1] = 0.0
        # unpadded grid (Nx points), half j = 0..Nx/2, Nyquist bin once
        self.nru = Nx // 2 + 1
        ju = np.arange(self.nru)
        c, s = _cos_sin(m, ju, Nx)
        w = np.where((m == 0) | (m == Nx // 2), 1.0, 2.0)
        self.Ce_u = c * w[None, :]
        self.So_u = -2.0 * s
        self.So_u[:, 0] = 0.0
        self.So_u[:, Nx // 2] = 0.0
        u = np.where((ju == 0) | (ju == Nx // 2), 1.0, 2.0)
        self.Ae_u = (c * u[:, None]).T / Nx
        # z-padded CGL product grid
        M = int(np.ceil((3 * N - 2) / 2.0))
        self.Q = Q = M + 1
        self.zp = cgl(Q)
        self.Vp = cheb_V(self.zp, N)
        xq = 2.0 * self.zp - 1.0
        Tfull = _C.chebvander(xq, M)                 # (Q, Q): T_n(x_q)
        wts = np.ones(Q)
        wts[0] = wts[-1] = 0.5
        gam = np.ones(N)
        gam[0] = 2.0
        self.Tp = (2.0 / (M * gam))[:, None] * (Tfull[:, :N].T * wts[None, :])
        self.VpD = self.Vp @ self.D
        self.VpD2 = self.Vp @ self.D2
        self._schur = None

    def schur_apply_ops(self):
        """Operators of the three-launch Schur apply (FreeSchurOperator.matvec,
        lagged_stokes.py:202-220), built once per shape in FP64:

        * ``Wz`` (Nk, Nz, 3Q): per mode the transposed stack
          ``[Vp + 2k^2 Vp G_k; 2 Vp D2 G_k - Vp; k Vp D G_k]``, so that one
          GEMM gives, on the Q product nodes, the three z-fields whose
          x-products the operator combines.  The four fields of v and
          w = G_k v (v, 2k^2 w, D2 w, k D w) enter the operator only as
          (t0 + t1, 2 t2 - t0, t3) of their products with eta, and the first
          three have one x-parity, so the sums are taken before the
          (linear) x-product instead of after it;
        * ``Mc`` (3Q, Nz): ``[(D2 Tp)^T; Tp^T; (D Tp)^T]``, applied to the
          products ``(u0, k^2 u1, -4k u2)`` of those three fields.
        """
        if self._schur is None:
            N, Q, Nk = self.Nz, self.Q, self.Nk
            W = np.empty((Nk, N, 3 * Q))
            for i, k in enumerate(self.kx):
                Gk = self.G[i]
                W[i] = np.concatenate([self.Vp + 2.0 * k * k * (self.Vp @ Gk),
                                       2.0 * (self.VpD2 @ Gk) - self.Vp,
                                       k * (self.VpD @ Gk)], 0).T
            Mc = np.concatenate([(self.D2 @ self.Tp).T, self.Tp.T, (self.D @ self.Tp).T], 0)
            self._schur = (W, np.ascontiguousarray(Mc))
        return self._schur

    def device_tensors(self, dtype, device, schur=True):
        """-> dict of torch tensors of every operator in ``dtype`` on ``device``
        (with ``schur``, also the composite operators of the Schur apply)."""
        import torch
        names = ["kx", "z", "zp", "V", "Vi", "D", "D2", "VD", "V0", "V1", "wq",
                 "G", "Qprof", "Ce_p", "So_p", "Ae_p", "Ao_p", "Ce_u", "So_u",
                 "Ae_u", "Vp", "Tp", "VpD", "VpD2"]
        out = {}
        for nm in names:
            a = np.ascontiguousarray(getattr(self, nm), dtype=np.float64)
            out[nm] = torch.from_numpy(a).to(device=device, dtype=dtype)
        # stacked z operators used by single GEMM launches
        out["V_VD"] = torch.cat([out["V"], out["VD"]], 0).contiguous()
        out["VpD2_VpD"] = torch.cat([out["Vp"], out["VpD2"], out["VpD"]], 0).contiguous()
        if schur:
            W, Mc = self.schur_apply_ops()
            out["Wz"] = torch.from_numpy(W).to(device=device, dtype=dtype)
            out["Mc"] = torch.from_numpy(Mc).to(device=device, dtype=dtype)
        return out


_CACHE = {}


def shape_ops(Nx, Nz, Lx):
    key = (int(Nx), int(Nz), float(Lx))
    op = _CACHE.get(key)
    if op is None:
        op = ShapeOps(*key)
        _CACHE[key] = op
    return op
