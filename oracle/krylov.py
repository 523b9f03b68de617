"""Right-preconditioned restarted FGMRES (oracle copy of krylov_fgmres.py:42-179).

TEST INFRASTRUCTURE: restatement used only as the checker (see ``oracle``).

The semantics that decide which iterate is returned are kept exactly:
modified Gram-Schmidt Arnoldi, complex Givens rotations with real ``c >= 0``,
the stop test on the quasi-residual ``|g[j+1]| <= tol`` with
``tol = max(atol, rtol ||b||)``, the true residual recomputed at the end of a
cycle that neither converged nor broke down, ``maxiter`` counting inner
iterations over all restarts, and the info codes 0 (converged), 1 (budget
exhausted), 2 (lucky breakdown).
"""
import numpy as np

__all__ = ["givens", "fgmres"]


def givens(a, b):
    """Rotation ``[[c, s], [-conj(s), c]]`` with ``G [a; b] = [r; 0]``
    (krylov_fgmres.py:42-57)."""
    aa, bb = abs(a), abs(b)
    if bb == 0.0:
        return 1.0, 0.0, a
    if aa == 0.0:
        return 0.0, np.conj(b) / bb, bb
    den = np.sqrt(aa * aa + bb * bb)
    c = aa / den
    s = c * np.conj(b) / np.conj(a)
    return c, s, c * (aa * aa + bb * bb) / np.conj(a)


def fgmres(A, b, M=None, x0=None, rtol=1e-8, atol=0.0, restart=None,
           maxiter=None):
    """-> ``(x, info, iters, res_hist)`` (krylov_fgmres.py:60-179)."""
    b = np.asarray(b)
    n = b.size
    dt = complex if np.iscomplexobj(b) else float
    x = np.zeros(n, dtype=dt) if x0 is None else np.array(x0, dtype=dt)
    m = max(1, int(min(n, 30) if restart is None else restart))
    maxiter = int(10 * n if maxiter is None else maxiter)
    tol = max(float(atol), float(rtol) * float(np.linalg.norm(b)))
    r = b - A(x)
    beta = float(np.linalg.norm(r))
    hist = [beta]
    if beta <= tol:
        return x, 0, 0, hist
    total, info = 0, 1
    while total < maxiter:
        V = np.zeros((m + 1, n), dtype=complex)
        Z = np.zeros((m, n), dtype=complex)
        H = np.zeros((m + 1, m), dtype=complex)
        cs = np.zeros(m)
        sn = np.zeros(m, dtype=complex)
        g = np.zeros(m + 1, dtype=complex)
        V[0] = r / beta
        g[0] = beta
        k = 0
        for j in range(m):
            if total >= maxiter:
                info = 1
                break
            Z[j] = M(V[j]) if M is not None else V[j].copy()
            w = A(Z[j])
            for i in range(j + 1):
                h = np.vdot(V[i], w)
                H[i, j] = h
                w = w - h * V[i]
            hn = float(np.linalg.norm(w))
            H[j + 1, j] = hn
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -np.conj(sn[i]) * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            c, s, rr = givens(H[j, j], H[j + 1, j])
            cs[j], sn[j] = c, s
            H[j, j] = rr
            H[j + 1, j] = 0.0
            t = c * g[j] + s * g[j + 1]
            g[j + 1] = -np.conj(s) * g[j] + c * g[j + 1]
            g[j] = t
            total += 1
            k = j + 1
            rn = float(abs(g[j + 1]))
            hist.append(rn)
            if rn <= tol:
                info = 0
                break
            if hn == 0.0:
                info = 2
                break
            V[j + 1] = w / hn
        if k == 0:
            break
        y = np.linalg.solve(H[:k, :k], g[:k])
        x = x + Z[:k].T @ y
        if info in (0, 2):
            break
        r = b - A(x)
        beta = float(np.linalg.norm(r))
        hist[-1] = beta
        if beta <= tol:
            info = 0
            break
    return x, info, total, hist
