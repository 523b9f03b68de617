"""Chebyshev-in-z / Fourier-in-x discretisation of the reference (oracle copy).

TEST INFRASTRUCTURE: restatement used only as the checker (see ``oracle``).

State convention (rbc_galerkin.py:63-65, 272, 322-330): a field is a complex
``(Nk, Nz)`` array, row ``m`` the rfft x-mode ``kx = 2 pi m / Lx`` (rfft
normalised by 1/Nx), column ``j`` the coefficient of ``T_j(2z - 1)``; ``z`` is
the CGL grid ascending from the hot bottom ``z = 0`` to the cold top ``z = 1``.
"""
import numpy as np
import scipy.linalg as sla

__all__ = ["cgl", "cheb_vander", "cheb_deriv", "Basis", "TauSolver"]


def cgl(N):
    """CGL nodes on [0, 1], ascending (rbc_galerkin.py:63-65)."""
    return 0.5 * (1.0 - np.cos(np.pi * np.arange(N) / (N - 1)))


def cheb_vander(x, N):
    """``T_k(x_j)`` for k < N by the three-term recurrence.

    Same values as ``chebvander(x, N-1)`` that rbc_galerkin.py:79-81 calls
    (``V[j, k] = T_k(2 z_j - 1)``); the recurrence is written out here and
    pinned against numpy in the oracle tests.
    """
    x = np.asarray(x, dtype=float)
    out = np.empty(x.shape + (N,))
    out[..., 0] = 1.0
    if N > 1:
        out[..., 1] = x
        two_x = 2.0 * x
        for k in range(2, N):
            out[..., k] = out[..., k - 1] * two_x - out[..., k - 2]
    return out


def cheb_deriv(N):
    """Coefficient-space d/dz on [0, 1] (rbc_galerkin.py:68-76).

    ``T_k'(x) = 2k sum_{j<k, k-j odd} T_j / c_j`` with ``c_0 = 2``, times 2 for
    the map ``x = 2z - 1``.  Strictly upper triangular.
    """
    D = np.zeros((N, N))
    for k in range(1, N):
        for j in range(k - 1, -1, -2):
            D[j, k] = 2.0 * (2.0 * k) / (2.0 if j == 0 else 1.0)
    return D


class Basis:
    """The per-shape geometry of ``RBC.__init__`` (rbc_galerkin.py:155-162,
    259-287): nodes, V / Vi / D / D2, wavenumbers and Clenshaw-Curtis weights.
    """

    def __init__(self, Nx, Nz, Lx):
        self.Nx, self.Nz, self.Lx = int(Nx), int(Nz), float(Lx)
        self.kx = 2.0 * np.pi * np.fft.rfftfreq(self.Nx, d=self.Lx / self.Nx)
        self.Nk = len(self.kx)
        self.x = np.arange(self.Nx) * self.Lx / self.Nx
        self.z = cgl(self.Nz)
        self.V = cheb_vander(2.0 * self.z - 1.0, self.Nz)
        self.Vi = np.linalg.inv(self.V)
        self.D = cheb_deriv(self.Nz)
        self.D2 = self.D @ self.D
        self.Np = 2 * self.Nx
        gk = np.array([0.0 if (k % 2) else 1.0 / (1.0 - k * k)
                       for k in range(self.Nz)])
        gk[0] = 1.0
        self.wq = self.Vi.T @ gk

    # -- transforms (rbc_galerkin.py:317-339) ----------------------------
    def vol_mean(self, g):
        return float(np.sum(self.wq * np.asarray(g).real.mean(axis=0)))

    def grid(self, cf):
        return np.fft.irfft((self.V @ np.asarray(cf).T).T * self.Nx,
                            n=self.Nx, axis=0)

    def coeffs(self, g):
        A = np.fft.rfft(np.asarray(g), axis=0) / self.Nx
        return (self.Vi @ A.T).T

    def synth_pad(self, cf):
        sp = np.zeros((self.Np // 2 + 1, np.shape(cf)[1]), dtype=complex)
        sp[:self.Nk] = (self.V @ np.asarray(cf).T).T
        return np.fft.irfft(sp * self.Np, n=self.Np, axis=0)

    def trunc_pad(self, g):
        return (np.fft.rfft(g, axis=0) / self.Np)[:self.Nk]

    def dz(self, cf):
        return (self.D @ np.asarray(cf).T).T

    def dz2(self, cf):
        return (self.D2 @ np.asarray(cf).T).T


class TauSolver:
    """Chebyshev-tau Dirichlet solver ``u'' - (k^2 + lam) u = rhs``
    (rbc_galerkin.py:147-185).

    Rows 0..N-3 carry the equation, rows N-2 / N-1 the values at z = 0 / 1;
    the two matching rhs entries are zeroed and real and imaginary parts are
    solved separately with one real LU.  Factors are kept per ``(k, lam)``:
    refactorising gives the same factors, so the reference's bounded cache
    (rbc_galerkin.py:85-144) has no numerical effect and is not modelled.
    """

    def __init__(self, basis):
        self.b = basis
        self.N = basis.Nz
        self._lu = {}

    def assemble(self, k, lam):
        N = self.N
        A = self.b.D2 - (k * k + lam) * np.eye(N)
        A[-2, :] = (-1.0) ** np.arange(N)
        A[-1, :] = 1.0
        return A

    def factors(self, k, lam):
        key = (float(k), float(lam))
        lu = self._lu.get(key)
        if lu is None:
            lu = sla.lu_factor(self.assemble(k, lam))
            if len(self._lu) > 4 * (self.b.Nk + 1):
                self._lu.clear()
            self._lu[key] = lu
        return lu

    def solve(self, k, rhs, lam=0.0):
        rhs = np.asarray(rhs)
        b = np.array(rhs, dtype=complex, copy=True)
        b[-2] = 0.0
        b[-1] = 0.0
        lu = self.factors(k, lam)
        if np.iscomplexobj(rhs):
            return sla.lu_solve(lu, b.real) + 1j * sla.lu_solve(lu, b.imag)
        return sla.lu_solve(lu, b.real)
