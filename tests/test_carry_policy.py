"""Host-side policy of the mixed-precision rescue carry (no device): which runs
skip the BFBT attempt and go straight to their reused FP32 factor."""
import numpy as np

from specmc_amd import engine as E
from specmc_amd.engine import Engine


class _Runs:
    """The per-run state the policy reads, with the engine's own methods."""
    _carry_select = Engine._carry_select
    _carry_ready = Engine._carry_ready
    _carries = Engine._carries
    _lag_bytes = Engine._lag_bytes

    def __init__(self, B, n, carry):
        self.B, self.n, self.lag_prec = B, n, 1
        self.rescue_carry = carry
        self.rescue_streak = np.zeros(B, dtype=np.int64)
        self.carry_run = np.zeros(B, dtype=np.int64)
        self.lag_have = np.zeros(B, dtype=bool)


def test_carry_truth_table():
    P = E.RESCUE_CARRY
    r = _Runs(6, 3168, P)
    act = np.ones(6, dtype=bool)
    # run 0: rejected, holds a factor, none carried yet -> carried
    # run 1: rejected, factor, P carried already -> probes BFBT instead
    # run 2: last attempt accepted -> BFBT
    # run 3: rejected but no factor (the rescue went to FP64 only) -> BFBT
    # run 4: rejected, factor, P - 1 carried -> carried (the last one)
    # run 5: as run 0 but not active this step
    r.rescue_streak[:] = [1, 1, 0, 1, 1, 1]
    r.lag_have[:] = [True, True, True, False, True, True]
    r.carry_run[:] = [0, P, 0, 0, P - 1, 0]
    act[5] = False
    got = r._carry_select(act)
    assert got.tolist() == [True, False, False, False, True, False]


def test_costly_bfbt_runs_are_carried_without_a_factor():
    # a run whose last BFBT solve converged after more than RESCUE_SWITCH_ITERS
    # iterations is carried onto the factor route; its first carried solve
    # factors, so it needs no factor yet
    P = E.RESCUE_CARRY
    r = _Runs(4, 3168, P)
    r.rescue_streak[:] = [E.CARRY_COSTLY, E.CARRY_COSTLY, E.CARRY_REJECTED, 0]
    r.lag_have[:] = [False, True, False, True]
    r.carry_run[:] = [0, P, 0, 0]
    got = r._carry_select(np.ones(4, dtype=bool))
    assert got.tolist() == [True, False, False, False]
    assert 0 < E.RESCUE_PROBE_ITERS <= E.RESCUE_SWITCH_ITERS


def test_fp64_never_carries():
    r = _Runs(4, 3168, 0)
    r.rescue_streak[:] = 1
    r.lag_have[:] = True
    assert not r._carries()
    assert not r._carry_select(np.ones(4, dtype=bool)).any()


def test_no_carry_when_factors_do_not_fit_the_store():
    # 256 x 192 (n = 24768): one FP32 factor is 2.45 GB, so 64 slots exceed
    # the lagged-factor store and every rescue refactors at this eta
    n = 129 * 192
    B = int(E.LAG_STORE_BYTES // (n * n * 4)) + 1
    r = _Runs(B, n, E.RESCUE_CARRY)
    r.rescue_streak[:] = 1
    r.lag_have[:] = True
    assert r._lag_bytes() > E.LAG_STORE_BYTES
    assert not r._carries()
    assert not r._carry_select(np.ones(B, dtype=bool)).any()
    r2 = _Runs(B - 1, n, E.RESCUE_CARRY)
    assert r2._carries()


def test_probe_budget_is_below_the_stokes_budget():
    # a probe is only worth its cost if it is far shorter than a full attempt
    from specmc_amd import _numerics as NUM
    assert 0 < E.RESCUE_PROBE_ITERS < NUM.STOKES_MAXITER
    assert E.RESCUE_CARRY >= 1
