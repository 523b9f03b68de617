"""Mixed precision on the device (MI355X): FP32 kernels against float64 host
products, and the FP32-factored dense rescue under the FP64 defect
correction against the oracle.

Tolerances: FP32 kernels on well-conditioned data to 1e-5 relative; the
mixed rescue must meet the reference's own acceptance test (true relative
residual <= 1e-8), so its Stokes solution agrees with the oracle's to
1e-8 relative at these small grids, and a trajectory through rescued
solves keeps theta within 1e-8 and Nu / Vrms within 1e-6 relative of the
oracle (the mixed-precision bar).
"""
import numpy as np
import pytest
import torch

import oracle
from golden_io import rel, seeded_theta

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    from specmc_amd.engine import Engine
    return Engine("const", 16, 24, 2.0, [dict(Ra=1e4)])


@pytest.mark.parametrize("mfma", [1, 2, 0])
def test_gemm_fp32(eng, mfma):
    """smx_gemm with prec=1 (FP32 matrix cores / vector FMA) against float64."""
    lib = eng.L
    lib.set_gemm_mfma(mfma)
    try:
        rng = np.random.default_rng(5)
        for (M, N, K, nz1) in ((37, 71, 29, 1), (130, 96, 145, 3), (2112, 145, 96, 1),
                               (300, 260, 64, 2)):
            A = rng.normal(size=(nz1, M, K)).astype(np.float32)
            Bm = rng.normal(size=(nz1, K, N)).astype(np.float32)
            C0 = rng.normal(size=(nz1, M, N)).astype(np.float32)
            dA, dB, dC = (torch.tensor(x, device=eng.dev) for x in (A, Bm, C0))
            lib.gemm(1, M, N, K, 0.5, -1.5, dA.data_ptr(), K, 1, M * K, 0,
                     dB.data_ptr(), N, 1, K * N, 0, dC.data_ptr(), N, 1, M * N, 0, nz1, 1,
                     None, 1, None, 0, 1, eng.stream())
            torch.cuda.synchronize()
            want = 0.5 * np.einsum("aik,akj->aij", A.astype(float), Bm.astype(float)) \
                - 1.5 * C0.astype(float)
            err = rel(dC.cpu().numpy().astype(float), want)
            print("fp32 gemm mfma=%d %s: rel %.2e" % (mfma, (M, N, K, nz1), err), flush=True)
            assert err < 1e-5, (M, N, K)
    finally:
        lib.set_gemm_mfma(1)


@pytest.mark.parametrize("n", [70, 1100, 3168])
def test_dense_lu_fp32(eng, n):
    """FP32 batched LU: its FP32 solve and its FP32-factor / FP64-vector solve
    against the float64 solution; defect correction in float64 with the FP32
    factor reaches the float64 residual floor."""
    lib = eng.L
    rng = np.random.default_rng(100 + n)
    nm = 2
    M = rng.normal(size=(nm, n, n)) + 4.0 * np.sqrt(n) * np.eye(n)
    print("fp32 dense lu n=%d" % n, flush=True)
    A = torch.tensor(M.astype(np.float32), device=eng.dev)
    piv = torch.zeros(nm, n, dtype=torch.int32, device=eng.dev)
    info = torch.ones(nm, dtype=torch.int32, device=eng.dev)
    lib.dense_lu(1, nm, n, A.data_ptr(), n * n, piv.data_ptr(), info.data_ptr(), eng.stream())
    b = rng.normal(size=(nm, n))
    b32 = torch.tensor(b.astype(np.float32), device=eng.dev)
    x32 = torch.zeros_like(b32)
    lib.dense_lu_solve(1, nm, n, A.data_ptr(), n * n, piv.data_ptr(), b32.data_ptr(),
                       x32.data_ptr(), None, eng.stream())
    bt = torch.tensor(b, device=eng.dev)
    x = torch.zeros_like(bt)
    lib.dense_lu_solve(2, nm, n, A.data_ptr(), n * n, piv.data_ptr(), bt.data_ptr(),
                       x.data_ptr(), None, eng.stream())
    torch.cuda.synchronize()
    assert info.cpu().numpy().tolist() == [0] * nm
    xs = np.stack([np.linalg.solve(M[q], b[q]) for q in range(nm)])
    e32 = rel(x32.cpu().numpy().astype(float), xs)
    emx = rel(x.cpu().numpy(), xs)
    # iterative refinement: x += LU32^{-1} (b - M x) in float64
    xr = x.cpu().numpy()
    for _ in range(4):
        r = b - np.einsum("qij,qj->qi", M, xr)
        rt = torch.tensor(r, device=eng.dev)
        dx = torch.zeros_like(rt)
        lib.dense_lu_solve(2, nm, n, A.data_ptr(), n * n, piv.data_ptr(), rt.data_ptr(),
                           dx.data_ptr(), None, eng.stream())
        torch.cuda.synchronize()
        xr = xr + dx.cpu().numpy()
    eref = rel(xr, xs)
    print("  fp32 solve %.2e, mixed solve %.2e, refined %.2e" % (e32, emx, eref), flush=True)
    assert e32 < 1e-4 and emx < 1e-4
    assert eref < 1e-12


def test_mixed_forced_rescue_matches_oracle(eng):
    """stokes_maxiter=1 forces the rescue; the FP32-factored rescue must be
    accepted by the reference's test and agree with the oracle's FP64 rescue."""
    from specmc_amd.models import RBCCoupled
    from oracle.models import OracleCoupled
    m = RBCCoupled(Ra=1e4, Nx=16, Nz=24, Lx=2.0, deta_T=1e3, stokes_maxiter=1,
                   precision="mixed")
    ref = OracleCoupled(Ra=1e4, Nx=16, Nz=24, Lx=2.0, deta_T=1e3, precond="bfbt",
                        stokes_maxiter=1)
    th = seeded_theta(m, 0.0, 7, sigma=5e-2).real
    psi, _ = m.stokes(th)
    psi_r, _ = ref.stokes(th)
    d = m.describe()
    print("mixed rescue: psi rel %.2e, defect res %.2e, fp32 %d redo %d" % (
        rel(psi, psi_r), d["last_defect_res"], d["n_rescue_fp32"], d["n_rescue_fp32_redo"]),
        flush=True)
    assert m.n_bfbt_rescue == 1 and ref.n_bfbt_rescue == 1
    assert d["n_rescue_fp32"] == 1 and d["n_rescue_fp32_redo"] == 0
    assert d["last_defect_res"] <= 1e-8
    # modes k >= 1 to the FP64 bar; the k = 0 row (the x-mean stream function,
    # zero by the mirror symmetry) lies along a near-null direction of the
    # Schur operator, which the residual test cannot see: the FP32 factor
    # leaves it at the FP32 level there (it enters theta, Nu and Vrms only at
    # second order)
    e1 = rel(psi[1:], psi_r[1:])
    e0 = float(np.abs(psi[0]).max() / np.abs(psi).max())
    print("  modes k>=1 rel %.2e, k=0 row %.2e of max |psi|" % (e1, e0), flush=True)
    assert e1 < 1e-10
    assert e0 < 1e-6


@pytest.mark.parametrize("shape", [(16, 24), (32, 48)])
def test_mixed_trajectory_through_rescues(eng, shape):
    """A march whose every Stokes solve goes through the rescue
    (stokes_maxiter=2): mixed precision vs the oracle on theta, Nu and Vrms."""
    from specmc_amd import ConvectionConfig, api
    Nx, Nz = shape
    cfg = ConvectionConfig(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=Nx, Nz=Nz,
                           Lx=2.0, precond="bfbt")
    kw = api.model_kwargs(cfg)
    from specmc_amd.models import RBCCoupled
    from oracle.models import OracleCoupled
    m = RBCCoupled(stokes_maxiter=2, precision="mixed", **kw)
    ref = OracleCoupled(stokes_maxiter=2, **kw)
    th0 = seeded_theta(m, 0.0, 3, sigma=3e-2).real
    s, info = api.march(m, th0, 2e-5, 6)
    s_r, _ = oracle.march(ref, th0.astype(complex), 2e-5, 6)
    d, dr = m.diagnostics(s), ref.diagnostics(s_r)
    dsc = m.describe()
    eth = rel(s, s_r.real)
    enu = abs(d["Nu_top"] - dr["Nu_top"]) / abs(dr["Nu_top"])
    evr = abs(d["urms"] - dr["urms"]) / abs(dr["urms"])
    print("mixed %dx%d: rescues %d + carried %d (fp32 %d, redo %d) theta %.2e Nu %.2e "
          "Vrms %.2e" % (Nx, Nz, dsc["n_bfbt_rescue"], dsc["n_rescue_carry"],
                         dsc["n_rescue_fp32"], dsc["n_rescue_fp32_redo"], eth, enu, evr),
          flush=True)
    assert dsc["n_bfbt_rescue"] + dsc["n_rescue_carry"] >= 6
    assert dsc["n_rescue_fp32"] >= 1
    assert eth < 1e-8
    assert enu < 1e-6 and evr < 1e-6


def test_rescue_carry_equals_attempt_then_rescue(eng):
    """Mixed precision carries a run whose BFBT solve was rejected straight to
    the rescue on its reused FP32 factor.  The rescue starts from zero, so
    when every BFBT attempt fails (stokes_maxiter=2) the carried march must
    equal the march that makes every attempt, bit for bit, with one attempt
    per carry period."""
    from specmc_amd import ConvectionConfig, api
    from specmc_amd.models import RBCCoupled
    cfg = ConvectionConfig(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=16, Nz=24,
                           Lx=2.0, precond="bfbt")
    kw = api.model_kwargs(cfg)
    runs = {}
    for carry in (0, 4):
        m = RBCCoupled(stokes_maxiter=2, precision="mixed", **kw)
        m.eng.rescue_carry = carry
        th0 = seeded_theta(m, 0.0, 3, sigma=3e-2).real
        s, _ = api.march(m, th0, 2e-5, 12)
        runs[carry] = (s, m.describe())
    (s0, d0), (s4, d4) = runs[0], runs[4]
    nsol = d0["n_bfbt_rescue"]
    print("carry: %d solves, %d attempts + %d carried; max |diff| %.1e" % (
        nsol, d4["n_bfbt_rescue"], d4["n_rescue_carry"], np.abs(s0 - s4).max()), flush=True)
    assert d0["n_rescue_carry"] == 0 and nsol >= 12
    assert d4["n_bfbt_rescue"] + d4["n_rescue_carry"] == nsol
    assert d4["n_bfbt_rescue"] == -(-nsol // 5)
    # every attempt after the first is a probe (its budget, 100, is above 2)
    assert d4["n_bfbt_probe"] == d4["n_bfbt_rescue"] - 1 and d0["n_bfbt_probe"] == 0
    assert np.array_equal(s0, s4)


@pytest.mark.parametrize("n", [12480, 24768])
def test_dense_lu_fp32_large(eng, n):
    """FP32 factor beyond the FP64 limit (the 128x192 and 256x192 Schur
    sizes): FP64 refinement with it reaches a 1e-12 residual."""
    lib = eng.L
    g = torch.Generator(device="cpu").manual_seed(n)
    print("fp32 dense lu n=%d" % n, flush=True)
    M = torch.randn(n, n, generator=g, dtype=torch.float64).to(eng.dev)
    M.diagonal().add_(4.0 * np.sqrt(n))
    A = M.to(torch.float32)
    piv = torch.zeros(1, n, dtype=torch.int32, device=eng.dev)
    info = torch.ones(1, dtype=torch.int32, device=eng.dev)
    lib.dense_lu(1, 1, n, A.data_ptr(), n * n, piv.data_ptr(), info.data_ptr(), eng.stream())
    b = torch.randn(1, n, generator=g, dtype=torch.float64).to(eng.dev)
    x = torch.zeros_like(b)
    r = b.clone()
    dx = torch.zeros_like(b)
    hist = []
    for _ in range(4):
        lib.dense_lu_solve(2, 1, n, A.data_ptr(), n * n, piv.data_ptr(), r.data_ptr(),
                           dx.data_ptr(), None, eng.stream())
        x += dx
        r = b - (M @ x[0])[None, :]
        hist.append(float(torch.linalg.norm(r) / torch.linalg.norm(b)))
    torch.cuda.synchronize()
    print("  residuals %s" % ["%.1e" % h for h in hist], flush=True)
    assert int(info.cpu()[0]) == 0
    assert hist[0] < 1e-4
    assert hist[-1] < 1e-12
