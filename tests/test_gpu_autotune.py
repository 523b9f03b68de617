"""Autotuner on the device (MI355X): measured GEMM choices and the
preconditioner-block schedule leave the results unchanged to roundoff.

Tolerance: 1e-12 relative on theta against the untuned batch (the kernel
choices reorder floating-point sums only), identical step counts.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    from specmc_amd import runtime
    return runtime.require_device()


def _ensemble(**kw):
    from specmc_amd import ConvectionConfig
    from specmc_amd.batch import Ensemble
    from specmc_amd.core import bench_ic
    from specmc_amd.workloads import _Grid
    cfgs = [ConvectionConfig(viscosity_model="eta_T", deta_T=1e3, Ra=1e4 * (1 + q), Nx=32,
                             Nz=48, Lx=2.0, precond="bfbt") for q in range(4)]
    g = _Grid(32, 48, 2.0)
    th = [bench_ic.seeded_theta(g, H=0.0, seed=q, sigma=2e-2).real for q in range(4)]
    return Ensemble(cfgs, th, np.full(4, 2e-4), **kw)


def test_autotuned_batch_matches_untuned(dev):
    from specmc_amd import autotune, runtime
    lib = runtime.load()
    lib.gemm_tune_clear()
    ref = _ensemble()
    ref.march()
    ens = _ensemble()
    try:
        rep = autotune.tune(ens.eng, ens.step)
        n_shapes = len(rep["gemm"])
        print("tuned %d GEMM shapes; pc %s" % (n_shapes, rep["pc"]["choice"]), flush=True)
        for k, v in rep["gemm"].items():
            print("  %s role %d: %s %.1f us (x%d per step)" % (k, v["role"], v["kernel"],
                                                              v["us"], v["calls_per_step"]))
        assert n_shapes >= 5
        assert rep["pc"]["choice"] in ("lu", "inv")
        ens.march()
    finally:
        lib.gemm_tune_clear()
    got, want = ens.states(), ref.states()
    err = max(np.linalg.norm(got[q] - want[q]) / np.linalg.norm(want[q]) for q in range(4))
    print("tuned vs untuned theta rel %.2e" % err, flush=True)
    assert np.array_equal(ens.steps, ref.steps)
    assert err < 1e-12
