"""The drop-in surface on the CPU: ConvectionConfig validation and the api
introspection helpers behave like the reference's (config.py:426-647,
api.py:68-144)."""
import json
import warnings

import numpy as np
import pytest

from specmc_amd import ConvectionConfig, _numerics
from specmc_amd import api
from specmc_amd.config import normalize_viscosity_model

REFERENCE_FIELDS = (
    "viscosity_model", "deta_T", "eta_z_coeff", "depth_dependence", "arrhenius_E",
    "arrhenius_V", "reference_anchor", "arrhenius_Ts", "eta_min", "eta_max",
    "eta_smooth", "precond", "internal_heating", "dissipation_number",
    "compressibility", "Ra", "Nx", "Nz", "Lx", "lag", "K", "defect_tol", "max_K",
    "symmetrize")


def test_fields_and_defaults():
    assert ConvectionConfig.field_names() == REFERENCE_FIELDS
    c = ConvectionConfig()
    assert (c.viscosity_model, c.deta_T, c.Ra, c.Nx, c.Nz, c.Lx) == ("eta_T", 1000.0,
                                                                     1e4, 64, 96, 2.0)
    assert (c.eta_min, c.eta_max, c.precond, c.lag, c.K, c.max_K) == (1e-3, 1e3,
                                                                      "legacy", 16, 3, 8)


@pytest.mark.parametrize("kw,exc", [
    (dict(viscosity_model="nope"), ValueError),
    (dict(viscosity_model="eta_Tz"), NotImplementedError),
    (dict(compressibility="ala"), NotImplementedError),
    (dict(compressibility="nope"), ValueError),
    (dict(Ra=-1.0), ValueError),
    (dict(Nx=3), ValueError),
    (dict(Nz=2.5), ValueError),
    (dict(Lx=0.0), ValueError),
    (dict(deta_T=0.0), ValueError),
    (dict(eta_min=0.0), ValueError),
    (dict(eta_min=2.0, eta_max=1.0), ValueError),
    (dict(eta_smooth=-0.1), ValueError),
    (dict(eta_smooth=10.0), ValueError),
    (dict(precond="ilu"), ValueError),
    (dict(dissipation_number=-1.0), ValueError),
    (dict(internal_heating=float("inf")), ValueError),
    (dict(arrhenius_Ts=-1.0), ValueError),
    (dict(arrhenius_Ts=1.0), ValueError),                 # needs 'arrhenius'
    (dict(arrhenius_E=5.0), ValueError),                  # Ts = 0: law off
    (dict(viscosity_model="arrhenius", arrhenius_Ts=1.0, arrhenius_E=1.0), ValueError),
    (dict(reference_anchor="Ea"), ValueError),
    (dict(eta_z_coeff=1.0), NotImplementedError),
    (dict(depth_dependence=0.5), NotImplementedError),
    (dict(viscosity_model="const", eta_min=2.0, eta_max=3.0), ValueError),
    (dict(K=9), ValueError),
    (dict(lag=0), ValueError),
    (dict(defect_tol=0.0, K=0), ValueError),
])
def test_validation_errors(kw, exc):
    with pytest.raises(exc):
        ConvectionConfig(**kw)


def test_not_implemented_messages():
    with pytest.raises(NotImplementedError, match="not implemented"):
        ConvectionConfig(viscosity_model="eta_Tz")


def test_symmetrize_warning():
    with pytest.warns(UserWarning):
        ConvectionConfig(symmetrize=False)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        ConvectionConfig(viscosity_model="const", symmetrize=False)


def test_aliases_and_arrhenius_calibration():
    assert normalize_viscosity_model(" Isoviscous ") == "const"
    assert normalize_viscosity_model("ETAT") == "eta_T"
    E = np.log(1000.0) * 1.0 * 2.0
    c = ConvectionConfig(viscosity_model="arrhenius", arrhenius_Ts=1.0, arrhenius_E=E)
    assert c.canonical()["viscosity_model"] == "arrhenius"
    # 'arrhenius' with Ts = 0 runs the exponential law (reference behaviour)
    ConvectionConfig(viscosity_model="arrhenius")


def test_api_introspection():
    with pytest.raises(TypeError):
        api.model_spec(dict(Nx=16))
    c = ConvectionConfig(viscosity_model="eta_T", Nx=16, Nz=24, precond="bfbt")
    d = api.describe(c)
    json.dumps(d)
    assert d["model"] == "eta_T" and d["model_class"] == "RBCCoupled"
    assert api.describe(ConvectionConfig(viscosity_model="const"))["model_class"] == "RBC"
    assert api.describe(ConvectionConfig(viscosity_model="eta_z"))["model_class"] == \
        "RBCImplicitTheta"
    n = api.numerics()
    assert n["THETA_RTOL"] == 1e-10 and n["STOKES_RTOL"] == 1e-8 and n["DTMAX"] == 4e-5
    assert n is not _numerics.as_dict() and n == _numerics.as_dict()
    kw = api.model_kwargs(c)
    assert kw["Nx"] == 16 and kw["precond"] == "bfbt"


def test_load_checkpoint(tmp_path):
    p = tmp_path / "ck_0.001000.npz"
    th = np.arange(12.0).reshape(3, 4) + 0j
    np.savez_compressed(p, that=th, t=0.001, n=25, fac=40.0)
    got, meta = api.load_checkpoint(str(p))
    assert np.array_equal(got, th) and meta["n"] == 25 and meta["t"] == 0.001
