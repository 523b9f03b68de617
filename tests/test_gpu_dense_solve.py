"""The multi-workgroup dense LU solve (smx_dense_lu_solve_mw) against the
one-workgroup-per-matrix solve and LAPACK (MI355X).

The two solves run the same triangular sweeps in the factor's type with
different summation orders: FP64 work agrees to 1e-12, FP32 work (prec 1
and 2) to the FP32 rounding of a well-conditioned system.
"""
import numpy as np
import pytest
import torch

from golden_io import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    from specmc_amd.engine import Engine
    return Engine("const", 16, 24, 2.0, [dict(Ra=1e4)])


def _factor(eng, M, prec):
    nm, n = M.shape[0], M.shape[1]
    dt = torch.float64 if prec == 0 else torch.float32
    A = torch.tensor(M, device=eng.dev).to(dt)
    piv = torch.zeros(nm, n, dtype=torch.int32, device=eng.dev)
    info = torch.ones(nm, dtype=torch.int32, device=eng.dev)
    eng.L.dense_lu(min(prec, 1), nm, n, A.data_ptr(), n * n, piv.data_ptr(), info.data_ptr(),
                   eng.stream())
    perm = torch.zeros(nm, n, dtype=torch.int32, device=eng.dev)
    eng.L.dense_lu_perm(nm, n, piv.data_ptr(), perm.data_ptr(), None, eng.stream())
    return A, piv, perm, info


def _ws(eng, prec, nm, n):
    nbytes = eng.L.dense_lu_solve_mw_bytes(prec, nm, n)
    assert nbytes > 0
    return torch.full((nbytes,), 0xAB, dtype=torch.uint8, device=eng.dev)


def test_perm_is_the_pivot_order(eng):
    """perm[i] = the row of b that row i of P b takes (LAPACK swaps in order)."""
    n, nm = 700, 3
    rng = np.random.default_rng(1)
    piv = np.stack([np.array([rng.integers(i, n) for i in range(n)]) for _ in range(nm)])
    pt = torch.tensor(piv, dtype=torch.int32, device=eng.dev)
    perm = torch.full((nm, n), -1, dtype=torch.int32, device=eng.dev)
    act = torch.tensor([1, 0, 1], dtype=torch.int32, device=eng.dev)
    eng.L.dense_lu_perm(nm, n, pt.data_ptr(), perm.data_ptr(), act.data_ptr(), eng.stream())
    got = perm.cpu().numpy()
    for q in (0, 2):
        idx = np.arange(n)
        for i in range(n):
            p = piv[q, i]
            idx[i], idx[p] = idx[p], idx[i]
        assert np.array_equal(got[q], idx)
    assert np.all(got[1] == -1)


@pytest.mark.parametrize("prec", [0, 1, 2])
@pytest.mark.parametrize("n,nm", [(40, 3), (64, 2), (200, 5), (2112, 6), (3168, 3)])
def test_mw_solve_matches_batched_solve(eng, prec, n, nm):
    """Same solutions as smx_dense_lu_solve (masked matrices untouched), a
    small residual against the FP64 matrix, and a clear error word."""
    rng = np.random.default_rng(n + 7 * prec)
    M = rng.normal(size=(nm, n, n)) + 2.0 * np.sqrt(n) * np.eye(n)
    A, piv, perm, info = _factor(eng, M, prec)
    vdt = torch.float32 if prec == 1 else torch.float64
    b = torch.tensor(rng.normal(size=(nm, n)), device=eng.dev).to(vdt)
    mask = torch.ones(nm, dtype=torch.int32, device=eng.dev)
    mask[1] = 0
    xb = torch.full_like(b, 7.0)
    xm = torch.full_like(b, 7.0)
    ws = _ws(eng, prec, nm, n)
    print("mw solve prec %d n=%d nmat=%d" % (prec, n, nm), flush=True)
    eng.L.dense_lu_solve(prec, nm, n, A.data_ptr(), n * n, piv.data_ptr(), b.data_ptr(),
                         xb.data_ptr(), mask.data_ptr(), eng.stream())
    eng.L.dense_lu_solve_mw(prec, nm, n, A.data_ptr(), n * n, perm.data_ptr(), b.data_ptr(),
                            xm.data_ptr(), mask.data_ptr(), ws.data_ptr(), ws.numel(),
                            eng.stream())
    torch.cuda.synchronize()
    assert info.cpu().numpy().tolist() == [0] * nm
    err_word = int(ws[8:12].view(torch.int32).cpu()[0])
    assert err_word == 0
    xb, xm, bh = xb.double().cpu().numpy(), xm.double().cpu().numpy(), b.double().cpu().numpy()
    tol = 1e-12 if prec == 0 else 1e-5
    for q in range(nm):
        if q == 1:
            assert np.all(xm[q] == 7.0)
            continue
        res = np.linalg.norm(M[q] @ xm[q] - bh[q]) / np.linalg.norm(bh[q])
        print("  q=%d mw vs batched %.2e, residual %.2e" % (q, rel(xm[q], xb[q]), res),
              flush=True)
        assert rel(xm[q], xb[q]) < tol
        assert res < (1e-12 if prec == 0 else 1e-5)


def test_mw_solve_repeats_and_skips_inactive(eng):
    """Two solves in a row on one workspace (the second with fewer active
    matrices) and a solve with none active."""
    n, nm = 1000, 8
    rng = np.random.default_rng(3)
    M = rng.normal(size=(nm, n, n)) + 2.0 * np.sqrt(n) * np.eye(n)
    A, piv, perm, _ = _factor(eng, M, 2)
    b = torch.tensor(rng.normal(size=(nm, n)), device=eng.dev)
    ws = _ws(eng, 2, nm, n)
    ref = torch.zeros_like(b)
    eng.L.dense_lu_solve(2, nm, n, A.data_ptr(), n * n, piv.data_ptr(), b.data_ptr(),
                         ref.data_ptr(), None, eng.stream())
    for act in ([1] * nm, [0, 0, 1, 0, 0, 0, 0, 1], [0] * nm):
        m = torch.tensor(act, dtype=torch.int32, device=eng.dev)
        x = torch.zeros_like(b)
        eng.L.dense_lu_solve_mw(2, nm, n, A.data_ptr(), n * n, perm.data_ptr(), b.data_ptr(),
                                x.data_ptr(), m.data_ptr(), ws.data_ptr(), ws.numel(),
                                eng.stream())
        torch.cuda.synchronize()
        on = np.array(act, dtype=bool)
        xh, rh = x.cpu().numpy(), ref.cpu().numpy()
        assert np.all(xh[~on] == 0.0)
        for q in np.where(on)[0]:
            assert rel(xh[q], rh[q]) < 1e-5
        assert int(ws[8:12].view(torch.int32).cpu()[0]) == 0


def test_mw_solve_large_fp32_factor(eng):
    """One FP32 factor at the 128x192 Schur size refined in FP64 with the
    multi-workgroup solve: the residual reaches 1e-12."""
    n = 12480
    g = torch.Generator(device="cpu").manual_seed(n)
    print("mw solve fp32 n=%d" % n, flush=True)
    M = torch.randn(n, n, generator=g, dtype=torch.float64).to(eng.dev)
    M.diagonal().add_(4.0 * np.sqrt(n))
    A = M.to(torch.float32)
    piv = torch.zeros(1, n, dtype=torch.int32, device=eng.dev)
    info = torch.ones(1, dtype=torch.int32, device=eng.dev)
    eng.L.dense_lu(1, 1, n, A.data_ptr(), n * n, piv.data_ptr(), info.data_ptr(), eng.stream())
    perm = torch.zeros(1, n, dtype=torch.int32, device=eng.dev)
    eng.L.dense_lu_perm(1, n, piv.data_ptr(), perm.data_ptr(), None, eng.stream())
    ws = _ws(eng, 2, 1, n)
    b = torch.randn(1, n, generator=g, dtype=torch.float64).to(eng.dev)
    x, r, dx = torch.zeros_like(b), b.clone(), torch.zeros_like(b)
    hist = []
    for _ in range(4):
        eng.L.dense_lu_solve_mw(2, 1, n, A.data_ptr(), n * n, perm.data_ptr(), r.data_ptr(),
                                dx.data_ptr(), None, ws.data_ptr(), ws.numel(), eng.stream())
        x += dx
        r = b - (M @ x[0])[None, :]
        hist.append(float(torch.linalg.norm(r) / torch.linalg.norm(b)))
    torch.cuda.synchronize()
    print("  residuals %s" % ["%.1e" % h for h in hist], flush=True)
    assert int(info.cpu()[0]) == 0
    assert hist[0] < 1e-4 and hist[-1] < 1e-12


@pytest.mark.parametrize("prec", [0, 2])
def test_mw_solve_is_backward_stable(eng, prec):
    """On matrices of condition ~1e9 (the Schur operators of the rescue reach
    1e8-1e9) the multi-workgroup solve's normwise backward error
    |A x - b| / (|A| |x| + |b|) stays within a small factor of the batched
    solve's and LAPACK's (FP64 factor), or of the FP32 rounding level (FP32
    factor); its forward error, cond * backward error, is then as good."""
    import scipy.linalg as sla
    n, nm = 600, 4
    rng = np.random.default_rng(11 + prec)
    Ms = []
    for _ in range(nm):
        U, _ = np.linalg.qr(rng.normal(size=(n, n)))
        V, _ = np.linalg.qr(rng.normal(size=(n, n)))
        Ms.append((U * np.logspace(0, -9, n)) @ V.T)
    M = np.stack(Ms)
    A, piv, perm, _ = _factor(eng, M, prec)
    b = torch.tensor(rng.normal(size=(nm, n)), device=eng.dev)
    xb, xm = torch.zeros_like(b), torch.zeros_like(b)
    ws = _ws(eng, prec, nm, n)
    eng.L.dense_lu_solve(prec, nm, n, A.data_ptr(), n * n, piv.data_ptr(), b.data_ptr(),
                         xb.data_ptr(), None, eng.stream())
    eng.L.dense_lu_solve_mw(prec, nm, n, A.data_ptr(), n * n, perm.data_ptr(), b.data_ptr(),
                            xm.data_ptr(), None, ws.data_ptr(), ws.numel(), eng.stream())
    torch.cuda.synchronize()
    bh, xb, xm = b.cpu().numpy(), xb.cpu().numpy(), xm.cpu().numpy()
    Mf = M if prec == 0 else A.double().cpu().numpy()

    def berr(q, x):
        return (np.linalg.norm(Mf[q] @ x - bh[q])
                / (np.linalg.norm(Mf[q], 2) * np.linalg.norm(x) + np.linalg.norm(bh[q])))
    for q in range(nm):
        xl = sla.lu_solve(sla.lu_factor(Mf[q]), bh[q])
        e_mw, e_b, e_l = berr(q, xm[q]), berr(q, xb[q]), berr(q, xl)
        print("  q=%d backward error mw %.1e batched %.1e lapack %.1e" % (q, e_mw, e_b, e_l),
              flush=True)
        floor = 1e-15 if prec == 0 else 1e-6
        assert e_mw <= 4.0 * max(e_b, e_l, floor)


@pytest.mark.parametrize("n", [300, 1100, 3168])
def test_fp32_fused_factor_matches_block_update(eng, n):
    """FP32 factorisation with the block's swaps and U rows in the sub-panel
    kernel and the rank-16 update on the matrix cores, against the separate
    block-update kernel: the same pivots, factors equal to FP32 roundoff, and
    P A = L U to FP32 accuracy (orders that end in partial 128-column blocks
    and partial 16-column sub-panels included)."""
    rng = np.random.default_rng(n)
    nm = 2
    M = rng.normal(size=(nm, n, n))
    got = {}
    try:
        for fuse in (1, 0):
            eng.L.set_lud_fuse(fuse)
            A, piv, perm, info = _factor(eng, M, 1)
            torch.cuda.synchronize()
            got[fuse] = (A.cpu().numpy().astype(np.float64), piv.cpu().numpy(),
                         perm.cpu().numpy(), info.cpu().numpy())
    finally:
        eng.L.set_lud_fuse(1)
    LU1, piv1, perm1, info1 = got[1]
    LU0, piv0, _, info0 = got[0]
    assert np.all(info1 == 0) and np.all(info0 == 0)
    # the two orders of the rank-16 sums could split a near-tie of two pivot
    # candidates differently; the factors compare only where the pivots agree
    same = np.array_equal(piv1, piv0)
    d = np.abs(LU1 - LU0).max() / np.abs(LU0).max() if same else 0.0
    res = []
    for q in range(nm):
        L = np.tril(LU1[q], -1) + np.eye(n)
        U = np.triu(LU1[q])
        res.append(np.abs(M[q][perm1[q]] - L @ U).max() / np.abs(M[q]).max())
    print("fp32 LU n=%d: same pivots %s, fused vs block kernel %.2e, |PA - LU| / |A| %s" % (
        n, same, d, ["%.2e" % r for r in res]), flush=True)
    assert same or n == 3168
    assert d < 1e-3
    assert max(res) < 2e-4
