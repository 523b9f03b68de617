"""Host-side measurement code (no device): the rocprof summariser's name
parsing and counter scaling, and the bench's phase weights."""
import csv
import types

import numpy as np
import pytest

import bench
from specmc_amd import profsum


def test_kernel_names_and_gemm_roles():
    n = "void k_gemm_mfma<double, 32, 32, 16, 1>(GemmArgs)"
    assert profsum._short(n) == "k_gemm_mfma<double, 32, 32, 16, 1>"
    assert profsum._role(profsum._short(n)) == 1
    k = profsum._short("void k_fg_post<double, 4, 0>(int, long, int, double const*)")
    assert k == "k_fg_post<double, 4, 0>" and profsum._role(k) == -1
    assert profsum._role("k_gemm<float, 64, 64, 3>") == 3
    groups = {g for g, _, f in profsum.GROUPS if f("k_gemm_mfma<double, 32, 32, 16, 0>")}
    assert groups == {"b"}
    groups = {g for g, _, f in profsum.GROUPS if f("k_lud_solve<float, double>")}
    assert groups == {"rescue"}


def _counter_csv(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Value"])
        w.writeheader()
        for i, (name, v) in enumerate(rows):
            w.writerow(dict(Dispatch_Id=i, Kernel_Name=name, Counter_Value=v))


def test_traffic_scales_each_counter_by_the_copy_probe(tmp_path):
    # FETCH_SIZE reads the 1 MiB copy as 2048 KiB (scale 1/2), WRITE_SIZE as
    # 1024 KiB (scale 1); the kernel's mean per dispatch is scaled to bytes
    cp = "__amd_rocclr_copyBuffer"
    f, w = str(tmp_path / "f.csv"), str(tmp_path / "w.csv")
    _counter_csv(f, [(cp, 12.0), (cp, 2048.0), ("void k_fg_post<double, 4, 0>(int)", 100.0),
                     ("void k_fg_post<double, 4, 0>(int)", 300.0), ("void other(int)", 9e9)])
    _counter_csv(w, [(cp, 1024.0), ("void k_fg_post<double, 4, 0>(int)", 50.0),
                     ("void k_fg_post<double, 4, 0>(int)", 50.0)])
    r = profsum.traffic(f, w, "void k_fg_post", 1.0)
    assert r["fetch_scale"] == pytest.approx(0.5) and r["write_scale"] == pytest.approx(1.0)
    assert r["fetch_bytes_per_launch"] == pytest.approx(200.0 * 1024 * 0.5)
    assert r["write_bytes_per_launch"] == pytest.approx(50.0 * 1024)
    assert r["hbm_bytes_per_launch"] == pytest.approx(102400.0 + 51200.0)
    assert r["fetch_dispatches"] == 2


def test_counter_means_sum_the_counters_of_one_dispatch(tmp_path):
    # an FP64 and an FP32 MFMA counter of the same dispatch add up
    p = str(tmp_path / "m.csv")
    with open(p, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Value"])
        w.writeheader()
        for d, (a, b) in enumerate(((10.0, 2.0), (20.0, 4.0))):
            for v in (a, b):
                w.writerow(dict(Dispatch_Id=d, Kernel_Name="void k_x<1>(int)", Counter_Value=v))
    means, scale = profsum._counter_means(p)
    assert scale == 1.0 and means["k_x<1>"] == pytest.approx((12.0 + 24.0) / 2)


def test_phase_weights_split_the_recorded_run():
    # columns used: 2 = Stokes iterations, 4 = rescue flag; 10 early steps,
    # 5 rescued, 2 at the iteration budget after them, 3 late
    log = np.zeros((20, 5))
    log[:10, 2] = 20
    log[10:15, 2], log[10:15, 4] = 800, 1
    log[15:17, 2] = 750
    log[17:, 2] = 400
    w = bench.phase_weights(log)
    np.testing.assert_allclose(w, np.array([10, 7, 3]) / 20.0)



def test_algorithmic_bytes_per_unit_of_the_clocked_kernels():
    eng = types.SimpleNamespace(n=3168, esize=8, Nz=96)
    # Arnoldi step: one FP64 vector of n = Nk * Nz per unit
    assert bench.unit_bytes(eng, "fg_post") == 25344
    # preconditioner apply: a 96 x 96 FP64 inverse block, its right-hand side
    # and its solution per system
    assert bench.unit_bytes(eng, "pc_apply") == (96 * 96 + 2 * 96) * 8 == 75264
    # dense solve: the factor once, pivots, FP64 right-hand side and solution
    assert bench.unit_bytes(eng, "dense_solve", 4) == 3168 * 3168 * 4 + 3168 * 4 + 2 * 3168 * 8
