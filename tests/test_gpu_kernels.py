"""Kernel variants on the device (MI355X) against each other and float64 host
arithmetic: the matrix-core pseudo-spectral product against the vector
kernel and a numpy restatement; the row-owner Gauss-Jordan inversion against
numpy.linalg.inv and the first kernel; the matrix-core Gauss-Jordan
inversion against both.

Tolerance: 1e-13 relative (the variants reorder floating-point sums only).
"""
import numpy as np
import pytest
import torch

from golden_io import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    from specmc_amd import runtime
    return runtime.load()


@pytest.mark.parametrize("shape", [(16, 24), (64, 96), (128, 96), (256, 192)])
def test_xprod_mfma_matches_vector_kernel_and_numpy(lib, shape):
    from specmc_amd.basis import shape_ops
    Nx, Nz = shape
    o = shape_ops(Nx, Nz, 2.0)
    dev = torch.device("cuda")
    t = o.device_tensors(torch.float64, dev)
    Nk, nrow, Q, nf, nodd, B = o.Nk, o.nrp, o.Q, 4, 3, 3
    rng = np.random.default_rng(Nx + Nz)
    y = rng.normal(size=(B, Nk, nf, Q))
    fs = rng.normal(size=(nf, Nk))
    eta = rng.uniform(0.1, 2.0, size=(B, nrow, Q))
    dy, dfs, de = (torch.tensor(a, device=dev) for a in (y, fs, eta))
    outs = []
    s = torch.cuda.current_stream().cuda_stream
    for on in (1, 0):
        lib.set_xprod_mfma(on)
        out = torch.zeros(B, Nk, nf, Q, dtype=torch.float64, device=dev)
        lib.xprod(0, B, nf, nodd, Nk, nrow, Q, dy.data_ptr(), dfs.data_ptr(), de.data_ptr(),
                  nrow * Q, out.data_ptr(), t["So_p"].data_ptr(), t["Ce_p"].data_ptr(),
                  t["Ao_p"].data_ptr(), t["Ae_p"].data_ptr(), None, s)
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
    lib.set_xprod_mfma(1)
    So, Ce = t["So_p"].cpu().numpy(), t["Ce_p"].cpu().numpy()
    Ao, Ae = t["Ao_p"].cpu().numpy(), t["Ae_p"].cpu().numpy()
    want = np.zeros_like(y)
    for b in range(B):
        for f in range(nf):
            S, An = (So, Ao) if f < nodd else (Ce, Ae)
            g = S @ (fs[f][:, None] * y[b, :, f, :])
            want[b, :, f, :] = An @ (g * eta[b])
    e_mfma, e_vec = rel(outs[0], want), rel(outs[1], want)
    print("xprod %s: |want| %.3e |mfma| %.3e |vector| %.3e; mfma %.2e, vector %.2e, "
          "max |mfma - vector| %.2e" % (shape, np.linalg.norm(want), np.linalg.norm(outs[0]),
                                        np.linalg.norm(outs[1]), e_mfma, e_vec,
                                        np.abs(outs[0] - outs[1]).max()), flush=True)
    assert np.linalg.norm(want) > 0.0
    assert e_mfma < 1e-13 and e_vec < 1e-13


@pytest.mark.parametrize("N", [16, 24, 40, 48, 96, 128])
def test_gj_inverse_variants(lib, N):
    """The register-resident, row-owner and first Gauss-Jordan kernels give
    the same inverse (the same arithmetic per entry) and match numpy; the
    matrix-core kernel (16-column panels, K = 16 products) matches them to
    roundoff."""
    rng = np.random.default_rng(N)
    nm = 7
    M = rng.normal(size=(nm, N, N)) + np.sqrt(N) * np.eye(N)
    M[0] = rng.normal(size=(N, N))           # pivoting on every column
    M[1] = M[1][rng.permutation(N)]          # a row permutation of a dominant matrix
    dev = torch.device("cuda")
    s = torch.cuda.current_stream().cuda_stream
    got = {}
    try:
        for name, mf, reg, rows in (("mfma", 1, 1, 0), ("reg", 0, 1, 0), ("rows", 0, 0, 1),
                                    ("first", 0, 0, 0)):
            lib.set_gj_mfma(mf)
            lib.set_gj_reg(reg)
            lib.set_gj_rows(rows)
            A = torch.tensor(np.ascontiguousarray(M.transpose(0, 2, 1)), device=dev)
            lib.invert(0, nm, N, A.data_ptr(), 1, None, s)
            torch.cuda.synchronize()
            got[name] = A.cpu().numpy().transpose(0, 2, 1)
    finally:
        lib.set_gj_mfma(1)
        lib.set_gj_reg(1)
        lib.set_gj_rows(0)
    want = np.linalg.inv(M)
    print("gj N=%d: mfma %.2e reg %.2e rows %.2e first %.2e; reg vs first %.2e, "
          "mfma vs reg %.2e" % (
              N, rel(got["mfma"], want), rel(got["reg"], want), rel(got["rows"], want),
              rel(got["first"], want), rel(got["reg"], got["first"]),
              rel(got["mfma"], got["reg"])), flush=True)
    for k in got:
        assert rel(got[k], want) < 1e-11, k
    assert rel(got["reg"], got["first"]) < 1e-14
    assert rel(got["rows"], got["first"]) < 1e-14
    assert rel(got["mfma"], got["reg"]) < 1e-12


@pytest.mark.parametrize("N", [24, 96])
def test_gj_mfma_active_mask(lib, N):
    """The matrix-core inversion leaves the blocks of inactive runs untouched
    and inverts the others (per_run blocks per run)."""
    rng = np.random.default_rng(100 + N)
    per_run, nrun = 3, 4
    M = rng.normal(size=(per_run * nrun, N, N)) + np.sqrt(N) * np.eye(N)
    act = np.array([1, 0, 1, 0], dtype=np.int32)
    dev = torch.device("cuda")
    s = torch.cuda.current_stream().cuda_stream
    A = torch.tensor(np.ascontiguousarray(M.transpose(0, 2, 1)), device=dev)
    da = torch.tensor(act, device=dev)
    lib.invert(0, per_run * nrun, N, A.data_ptr(), per_run, da.data_ptr(), s)
    torch.cuda.synchronize()
    got = A.cpu().numpy().transpose(0, 2, 1)
    for q in range(per_run * nrun):
        if act[q // per_run]:
            assert rel(got[q], np.linalg.inv(M[q])) < 1e-11, q
        else:
            assert np.array_equal(got[q], M[q]), q


@pytest.mark.parametrize("N", [24, 48, 96, 192])
def test_inv_apply_pairs(lib, N):
    """Row-pair inverse application against the first kernel and numpy, with
    the tau zeroing, per-mode and per-run scales."""
    rng = np.random.default_rng(7 + N)
    nm, per_run = 12, 4
    Inv = rng.normal(size=(nm, N, N))
    b = rng.normal(size=(nm, N))
    ks = rng.normal(size=per_run)
    bsc = rng.normal(size=nm // per_run)
    dev = torch.device("cuda")
    s = torch.cuda.current_stream().cuda_stream
    dI = torch.tensor(np.ascontiguousarray(Inv.transpose(0, 2, 1)), device=dev)
    db, dk, dbs = (torch.tensor(a, device=dev) for a in (b, ks, bsc))
    got = []
    for on in (1, 0, 2):
        lib.set_inv_pairs(on)
        x = torch.zeros(nm, N, dtype=torch.float64, device=dev)
        lib.inv_apply(0, nm, N, dI.data_ptr(), db.data_ptr(), x.data_ptr(), 1, -0.5,
                      dk.data_ptr(), dbs.data_ptr(), per_run, 0, None, s)
        torch.cuda.synchronize()
        got.append(x.cpu().numpy())
    lib.set_inv_pairs(1)
    sc = -0.5 * ks[np.arange(nm) % per_run] * bsc[np.arange(nm) // per_run]
    bz = b * sc[:, None]
    bz[:, N - 2:] = 0.0
    want = np.einsum("mij,mj->mi", Inv, bz)
    print("inv_apply N=%d: pairs %.2e, first %.2e" % (N, rel(got[0], want), rel(got[1], want)),
          flush=True)
    assert rel(got[0], want) < 1e-13 and rel(got[1], want) < 1e-13
    # the prefetching row-pair kernel accumulates in the same order
    assert np.array_equal(got[0], got[2])


@pytest.mark.parametrize("shape,nact", [((32, 48), 5), ((64, 96), 70)])
def test_schur_apply_over_listed_active_runs(lib, shape, nact):
    """The Schur apply with a run mask runs its z-synthesis over the active
    runs listed on the device: the active runs' outputs equal those of the
    unmasked apply and of the apply without the list, bit for bit, and the
    other runs' outputs are left as they were."""
    from specmc_amd.engine import Engine
    Nx, Nz = shape
    B = 96 if Nx == 64 else 12
    e = Engine("coupled", Nx, Nz, 2.0, [dict(Ra=1e4 * (1 + 0.01 * i), deta_T=1e3,
                                              precond="bfbt") for i in range(B)])
    rng = np.random.default_rng(B + nact)
    e.set_theta(rng.normal(size=(B, e.Nk, e.Nz)) * 1e-2 * np.exp(-np.arange(e.Nz) / 8.0))
    e._eta_of(e.theta, None)
    v = torch.tensor(rng.normal(size=(B, e.n)), dtype=torch.float64, device=e.dev)
    act = np.zeros(B, dtype=np.int32)
    act[rng.choice(B, min(nact, B), replace=False)] = 1
    mask = torch.tensor(act, device=e.dev)
    ref = torch.empty_like(v)
    e.schur_matvec(v, ref, None)
    outs = []
    try:
        for compact in (1, 0):
            lib.set_schur_compact(compact)
            out = torch.full_like(v, np.nan)
            e.schur_matvec(v, out, mask)
            torch.cuda.synchronize()
            outs.append(out.cpu().numpy())
    finally:
        lib.set_schur_compact(1)
    on = act.astype(bool)
    r = ref.cpu().numpy()
    assert np.isfinite(r[on]).all()
    assert np.array_equal(outs[0][on], r[on])
    assert np.array_equal(outs[1][on], r[on])
    assert np.isnan(outs[0][~on]).all() and np.isnan(outs[1][~on]).all()
