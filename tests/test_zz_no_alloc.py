"""smx_schur_apply_static (the Schur operator apply on the library's own
device workspace) against smx_schur_apply with a caller's workspace, through
the C ABI: equal outputs on two calls in a row, the second with fewer runs,
and a clear error for a batch whose workspace exceeds the library's."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _engine(B):
    from specmc_amd.engine import Engine
    params = [dict(Ra=1e4 * (1 + 0.3 * i), deta_T=1e3, precond="bfbt") for i in range(B)]
    e = Engine("coupled", 32, 48, 2.0, params)
    rng = np.random.default_rng(7)
    th = rng.normal(size=(B, e.Nk, e.Nz)) * 1e-2 * np.exp(-np.arange(e.Nz) / 8.0)
    e.set_theta(th)
    e._eta_of(e.theta, None)
    return e


def _args(e, B, v, out):
    t = e.t
    return (e.prec, B, e.Nk, e.Nz, e.nrp, e.Q, e.p(v), e.p(e.eta_S), e.nrp * e.Q, e.p(out),
            e.p(t["Wz"]), e.p(t["Mc"]), e.p(t["kx"]), e.p(t["So_p"]), e.p(t["Ce_p"]),
            e.p(t["Ao_p"]), e.p(t["Ae_p"]), e.p(t["V0"]), e.p(t["V1"]))


def test_static_workspace_apply_equals_workspace_apply():
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    from specmc_amd import runtime
    L = runtime.load()
    B = 5
    e = _engine(B)
    print("engine", e.Nx, e.Nz, "B", B, flush=True)
    rng = np.random.default_rng(11)
    v = torch.tensor(rng.normal(size=(B, e.n)), dtype=torch.float64, device=e.dev)
    for nb in (B, 2):                         # two calls, the second smaller
        out_ws = torch.full((B, e.n), np.nan, dtype=torch.float64, device=e.dev)
        out_st = torch.full((B, e.n), np.nan, dtype=torch.float64, device=e.dev)
        work = torch.zeros(int(L.schur_apply_work(nb, e.Nk, e.Q)), dtype=torch.float64,
                           device=e.dev)
        L.schur_apply(*_args(e, nb, v, out_ws), e.p(work), work.numel(), None, e.stream())
        L.schur_apply_static(*_args(e, nb, v, out_st), None, e.stream())
        torch.cuda.synchronize()
        a, b = out_ws[:nb].cpu().numpy(), out_st[:nb].cpu().numpy()
        print("call with", nb, "runs: max |diff|", float(np.abs(a - b).max()), flush=True)
        assert np.isfinite(a).all()
        assert np.array_equal(a, b)
        # runs beyond nb are untouched by both
        assert torch.isnan(out_st[nb:]).all() and torch.isnan(out_ws[nb:]).all()
    # the engine's own apply (caller workspace) is the same operator
    ref = torch.empty_like(v)
    e.schur_matvec(v, ref, None)
    got = torch.empty_like(v)
    L.schur_apply_static(*_args(e, B, v, got), None, e.stream())
    torch.cuda.synchronize()
    assert torch.equal(ref, got)


def test_static_workspace_rejects_oversized_batches():
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    from specmc_amd import runtime
    L = runtime.load()
    e = _engine(1)
    cap = int(L.schur_apply_static_bytes())
    assert cap == 64 << 20
    nb = cap // (8 * int(L.schur_apply_work(1, e.Nk, e.Q))) + 1
    # the call below must be rejected before any launch (its buffers hold
    # one run): check on the host that this batch exceeds the workspace
    assert 8 * int(L.schur_apply_work(nb, e.Nk, e.Q)) > cap
    v = torch.zeros(1, e.n, dtype=torch.float64, device=e.dev)
    out = torch.zeros_like(v)
    # rejected on the host before any launch, so the small buffers are never read
    with pytest.raises(RuntimeError, match="workspace"):
        L.schur_apply_static(*_args(e, nb, v, out), None, e.stream())
