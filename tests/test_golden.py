"""Golden fixtures: the reference regression test's two profiles
(tests/golden/regression_inputs.json, parsed from the reference's
Load_Atm_Data.inc / Load_Sfc_Data.inc) through the oracle, and the
AddLayers extension they need (their top level is 0.714 hPa)."""
from __future__ import annotations

import ctypes
import json
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, GOLD)
import make_golden  # noqa: E402


@pytest.fixture(scope="module")
def gold():
    doc = json.load(open(os.path.join(GOLD, "regression_inputs.json")))
    exp = np.load(os.path.join(GOLD, "regression_expected.npz"), allow_pickle=False)
    return doc, exp, make_golden.profile_set(doc["profiles"])


@pytest.fixture(scope="module")
def golden_coeffs(gold):
    from crtmv3_amd import synthetic
    doc, _, _ = gold
    d = synthetic.default_coeff_dir()
    for name in make_golden.SENSORS:
        synthetic.write_sensor(d, name)
        if make_golden.fingerprint(d, name) != doc["coefficients"][name]:
            pytest.fail(f"{name}: synthetic coefficient files differ from the fixture's; "
                        "regenerate with tests/golden/make_golden.py")
    return d


def test_inputs_are_the_reference_profiles(gold):
    doc, _, ps = gold
    assert [len(p["temperature"]) for p in doc["profiles"]] == [92, 92]
    assert [p["climatology"] for p in doc["profiles"]] == [6, 1]        # US standard, tropical
    assert ps.level_pressure[0] == pytest.approx(0.714) and ps.level_pressure[92] == pytest.approx(1100.0)
    np.testing.assert_allclose([ps.land_coverage[0], ps.water_coverage[0], ps.snow_coverage[0], ps.ice_coverage[0]],
                               [0.1, 0.5, 0.25, 0.15])


def test_add_layers_extension(gold, oracle_mod):
    """AddLayers on the fixture: model levels above the profile top, the
    top itself kept, layer means from the climatology shifted so the lowest
    added layer continues the profile's top two layers."""
    doc, _, ps = gold
    L = oracle_mod.lib()
    clim = json.load(open(os.path.join(ROOT, "crtmv3_amd", "data", "model_profiles.json")))
    cp = np.array(clim["level_pressure"])
    for m in range(2):
        keep = []
        p = oracle_mod.make_profile(ps, m, keep)
        n = L.oracle_add_layers(ctypes.byref(p), None, None, None, None)
        top = ps.profile(m)["level_pressure"][0]
        assert n == int(np.argmax(cp[1:] >= top)) + 1 and cp[n - 1] < top <= cp[n]
        K = 92 + n
        lp, pr, t, a = np.zeros(K + 1), np.zeros(K), np.zeros(K), np.zeros(2 * K)
        L.oracle_add_layers(ctypes.byref(p), lp.ctypes.data_as(oracle_mod._dp), pr.ctypes.data_as(oracle_mod._dp),
                            t.ctypes.data_as(oracle_mod._dp), a.ctypes.data_as(oracle_mod._dp))
        np.testing.assert_array_equal(lp[:n], cp[:n])
        assert lp[n] == top
        np.testing.assert_array_equal(lp[n:], ps.profile(m)["level_pressure"])
        np.testing.assert_array_equal(t[n:], ps.profile(m)["temperature"])
        assert np.all(np.diff(lp) > 0)
        np.testing.assert_allclose(pr[:n], (lp[1:n + 1] - lp[:n]) / np.log(lp[1:n + 1] / lp[:n]), rtol=1e-14)
        # Shift_Profile: lowest added layer on the log-pressure line through the top two layers
        pu, tu = ps.profile(m)["pressure"], ps.profile(m)["temperature"]
        x = (np.log(pr[n - 1]) - np.log(pu[0])) / (np.log(pu[1]) - np.log(pu[0]))
        assert t[n - 1] == pytest.approx(tu[0] + (tu[1] - tu[0]) * x, rel=1e-13)
        # the added layers keep the climatology's layer-to-layer differences
        mt = np.array(clim["temperature"][doc["profiles"][m]["climatology"] - 1])
        tl = mt[:n + 1].copy()
        tl[n] = mt[n - 1] + (mt[n] - mt[n - 1]) * (np.log(top) - np.log(cp[n - 1])) / (np.log(cp[n]) - np.log(cp[n - 1]))
        np.testing.assert_allclose(np.diff(t[:n]), np.diff(0.5 * (tl[:-1] + tl[1:])), atol=1e-11)


def test_oracle_reproduces_expected(gold, golden_coeffs, oracle_mod):
    from crtmv3_amd import coeffio
    _, exp, ps = gold
    for name in make_golden.SENSORS:
        tau = coeffio.read_odps(os.path.join(golden_coeffs, f"{name}.TauCoeff.bin"))
        spc = coeffio.read_spccoeff(os.path.join(golden_coeffs, f"{name}.SpcCoeff.bin"))
        s = oracle_mod.Sensor(tau, spc)
        eps = make_golden.emissivity(spc.sensor_type, tau.n_channels)
        for m in range(2):
            r = oracle_mod.run_profile(s, ps, m, eps, "k")
            for key in ("bt", "t_k", "abs_k", "sfc_t_k", "emis_k"):
                np.testing.assert_allclose(r[key], exp[f"{name}/{m}/{key}"], rtol=1e-12, atol=1e-300,
                                           err_msg=f"{name} profile {m} {key}")


def test_added_layers_adjoint_is_consistent(gold, golden_coeffs, oracle_mod):
    """With AddLayers active the reverse-mode K-matrix still equals the
    complex-step derivative of the forward model (both treat the added
    layers as fixed, as CRTM_Atmosphere_AddLayers_AD does)."""
    from crtmv3_amd import coeffio
    _, _, ps = gold
    name = "hirs_syn"
    tau = coeffio.read_odps(os.path.join(golden_coeffs, f"{name}.TauCoeff.bin"))
    spc = coeffio.read_spccoeff(os.path.join(golden_coeffs, f"{name}.SpcCoeff.bin"))
    s = oracle_mod.Sensor(tau, spc)
    eps = make_golden.emissivity(spc.sensor_type, tau.n_channels)
    r = oracle_mod.run_profile(s, ps, 1, eps, "k")
    c = oracle_mod.run_profile(s, ps, 1, eps, "cstep")
    for key in ("t_k", "abs_k", "sfc_t_k", "emis_k"):
        np.testing.assert_allclose(r[key], c[key], rtol=0, atol=1e-11 * np.abs(c[key]).max(), err_msg=key)
