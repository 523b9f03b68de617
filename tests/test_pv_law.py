"""The pressure-dependent (pV) Arrhenius law: calibration and configuration
(CPU).  The calibration is pinned by the reference's documented numbers
(V-tilde = 22.504874770248204 and Ts = 0.091 for the default constants,
reference_state.py:233-246) and by the identities it rests on: the hot end
sits at -ln D, the V = 0 limit is the T-only Arrhenius law, and along the
conductive profile the pV term is an activation shift."""
import math

import numpy as np
import pytest

import oracle
from oracle import laws
from specmc_amd import ConvectionConfig, api
from specmc_amd import reference_state as rs


def test_documented_constants():
    assert rs.V_tilde() == pytest.approx(22.504874770248204, rel=1e-15)
    assert rs.Ts_tilde() == pytest.approx(0.091, rel=1e-15)
    assert rs.infer_V_a(rs.V_tilde()) == pytest.approx(6e-6, rel=1e-14)


@pytest.mark.parametrize("anchor,V", [("D", 5.0), ("D", 22.5), ("Ea", 10.0), ("Ea", 0.0)])
def test_calibration_identities(anchor, V):
    rec = rs.calibrate_pv(anchor=anchor, V=V, deta_T=1e3 if anchor == "D" else None)
    E, Ts, C = rec["E"], rec["Ts"], rec["C"]
    lneta = lambda T, z: -E * T / (Ts * (T + Ts)) + V * (1.0 - z) / (T + Ts)
    # eta = 1 at the cold top, eta = 1/D at the hot bottom
    assert lneta(0.0, 1.0) == 0.0
    assert abs(lneta(1.0, 0.0) + rec["ln_D"]) < 1e-12 * max(1.0, rec["ln_D"])
    assert rec["hot_end_residual"] < 1e-12
    if anchor == "D":
        assert rec["D"] == 1e3
    # the contract form (E + V p0)/(T + Ts) + C equals the evaluated form
    for T, z in ((0.3, 0.7), (0.9, 0.05), (1.7, 0.4)):
        assert abs((E + V * (1 - z)) / (T + Ts) + C - lneta(T, z)) < 1e-11
    # along T = 1 - z: E_eff / (T + Ts) + C + V
    for z in (0.1, 0.5, 0.95):
        T = 1.0 - z
        assert abs(lneta(T, z) - (rec["E_eff"] / (T + Ts) + C + V)) < 1e-11
    # the oracle's restatement agrees
    Eo, Vo, To = laws.calibrate_pv(anchor, V, deta_T=1e3)
    assert (Eo, Vo, To) == pytest.approx((E, V, Ts), rel=1e-15)


def test_V0_limit_is_the_T_only_law():
    rec = rs.calibrate_pv(anchor="D", V=0.0, deta_T=1e3)
    T = np.linspace(0.0, 2.0, 11)
    got = laws.arrhenius_exponent_V(T, 0.5, rec["E"], 0.0, rec["Ts"])
    want = laws.arrhenius_exponent(T, 1e3, rec["Ts"])
    assert np.allclose(got, want, rtol=1e-14, atol=1e-14)


def test_config_rules():
    base = dict(viscosity_model="arrhenius", deta_T=1e3, Nx=16, Nz=24)
    ConvectionConfig(arrhenius_V=5.0, **base).validate()
    ConvectionConfig(arrhenius_V=5.0, reference_anchor="Ea", **base).validate()
    rec = rs.calibrate_pv(anchor="D", V=5.0, deta_T=1e3)
    ConvectionConfig(arrhenius_V=5.0, arrhenius_Ts=rec["Ts"], arrhenius_E=rec["E"],
                     **base).validate()
    with pytest.raises(ValueError, match="needs viscosity_model='arrhenius'"):
        ConvectionConfig(viscosity_model="eta_T", arrhenius_V=5.0).validate()
    with pytest.raises(ValueError, match="unknown reference_anchor"):
        ConvectionConfig(arrhenius_V=5.0, reference_anchor="X", **base).validate()
    with pytest.raises(ValueError, match="contradicts the reference-state value"):
        ConvectionConfig(arrhenius_V=5.0, arrhenius_Ts=0.1, **base).validate()
    with pytest.raises(ValueError, match="contradicts the calibrated pV activation"):
        ConvectionConfig(arrhenius_V=5.0, arrhenius_E=1.0, **base).validate()
    with pytest.raises(ValueError, match="must stay at its default 'D'"):
        ConvectionConfig(reference_anchor="Ea", **base).validate()
    cfg = ConvectionConfig(arrhenius_V=5.0, **base)
    kw = api.model_kwargs(cfg)
    assert kw["arrhenius_V"] == 5.0 and kw["reference_anchor"] == "D"
    from specmc_amd.config import arrhenius_effective_E
    assert arrhenius_effective_E(cfg) == pytest.approx(rec["E"], rel=1e-15)


def test_oracle_pv_eta_field():
    """The oracle's eta for a pV run: exp of the law on the padded grid, with
    p0 = 1 - z of each row (checked pointwise before the window)."""
    m = oracle.build_oracle(dict(viscosity_model="arrhenius", deta_T=1e2, Ra=1e4, Nx=16,
                                 Nz=24, arrhenius_V=3.0, eta_min=1e-300, eta_max=1e300))
    E, V, Ts = m.pv
    T = np.clip(np.random.default_rng(0).uniform(-0.1, 1.1, size=(32, 24)), 0, 2)
    got = m.eta_from_T(T)
    want = np.exp(-E * T / (Ts * (T + Ts)) + V * (1.0 - m.z[None, :]) / (T + Ts))
    assert np.allclose(got, want, rtol=1e-14)
    assert math.isclose(Ts, 0.091, rel_tol=1e-15)
