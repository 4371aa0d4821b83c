"""GPU parity: the HIP forward and K-matrix kernels, called through the
C-ABI (libcrtm_mi355x.so), against the FP64 oracle on the same seeded
inputs.

Tolerances (brightness temperature in K, Jacobians relative to the largest
element of the same Jacobian block, and element-wise as
conftest.assert_jacobian_close states: rtol 1e-7 FP64, 1e-3 FP32-split, plus
a small absolute allowance):
  FP64 path        BT 1e-8 K,  Jacobians 1e-9
  FP32-split path  BT 1e-2 K,  Jacobians 1e-3
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import SMALL_SENSORS, assert_jacobian_close, load_sensor

pytestmark = pytest.mark.gpu

TOL = {0: dict(bt=1e-8, jac=1e-9), 1: dict(bt=1e-2, jac=1e-3)}


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def contexts(coeff_dir, gpu):
    from crtmv3_amd import api
    ctxs = {name: api.crtm_init([name], coeff_dir) for name in SMALL_SENSORS}
    yield ctxs
    for c in ctxs.values():
        c.destroy()


def _oracle_batch(orc, coeff_dir, name, ps, eps, bt_k=None):
    tau, spc = load_sensor(coeff_dir, name)
    s = orc.Sensor(tau, spc)
    return [orc.run_profile(s, ps, m, eps[m], "k", bt_k=None if bt_k is None else bt_k[m])
            for m in range(ps.n_profiles)]


def _check(ps, rts, kd, ref, prec, L, check_k=True):
    tol = TOL[prec]
    bt_ref = np.stack([r["bt"] for r in ref])
    assert np.abs(rts["brightness_temperature"] - bt_ref).max() <= tol["bt"]
    rad_ref = np.stack([r["radiance"] for r in ref])
    np.testing.assert_allclose(rts["radiance"], rad_ref, rtol=1e-3 if prec else 1e-10)
    if not check_k:
        return
    for m, r in enumerate(ref):
        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        tk = np.asarray(kd["temperature_k"][a:b], np.float64).T
        ak = np.asarray(kd["absorber_k"][:, a:b, :], np.float64).transpose(2, 0, 1)
        for got, want, what in ((tk, r["t_k"], "T_K"), (ak, r["abs_k"], "absorber_K"),
                                (kd["water_temperature_k"][m], r["sfc_t_k"][:, 1], "Ts_K"),
                                (kd["surface_emissivity_k"][m], r["emis_k"], "emis_K")):
            assert_jacobian_close(got, want, tol["jac"], f"profile {m} {what}", prec)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "split"])
@pytest.mark.parametrize("name", SMALL_SENSORS)
def test_k_matrix_matches_oracle_uneven_layers(contexts, coeff_dir, oracle_mod, name, prec):
    from crtmv3_amd import synthetic
    ctx = contexts[name]
    L = ctx.n_channels
    ks = [100, 60, 137, 83, 100, 71, 30]     # 30 layers: wide averaging supports
    ps = synthetic.make_profiles(len(ks), ks, seed=101)
    rng = np.random.default_rng(5)
    eps = rng.uniform(0.5, 0.99, (len(ks), L))
    btk = rng.uniform(0.5, 1.5, (len(ks), L))
    print(f"case {name} prec={prec} M={len(ks)} layers={ks}", flush=True)
    ref = _oracle_batch(oracle_mod, coeff_dir, name, ps, eps, btk)
    rts, kd = ctx.k_matrix(ps, eps, bt_k=btk, precision=prec, jacobian_f32=bool(prec))
    _check(ps, rts, kd, ref, prec, L)
    fwd = ctx.forward(ps, eps, precision=prec, extras=True)
    _check(ps, fwd, None, ref, prec, L, check_k=False)
    # forward and K-matrix share the forward half: identical BT
    np.testing.assert_array_equal(fwd["brightness_temperature"], rts["brightness_temperature"])


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "split"])
def test_device_memory_equals_host_memory(contexts, gpu, prec):
    import torch
    from crtmv3_amd import api, synthetic
    name = "cris_small_syn"
    ctx = contexts[name]
    L = ctx.n_channels
    ps = synthetic.make_profiles(5, [90, 100, 64, 120, 100], seed=7)
    eps = np.full((5, L), 0.97)
    print(f"case device-vs-host {name} prec={prec}", flush=True)
    h_rts, h_k = ctx.k_matrix(ps, eps, precision=prec, jacobian_f32=False)
    dps = api.device_profiles(ps, gpu)
    d_rts, d_k = ctx.k_matrix(dps, torch.as_tensor(eps, device=gpu), precision=prec, jacobian_f32=False)
    ctx.synchronize()
    np.testing.assert_array_equal(d_rts["brightness_temperature"].cpu().numpy(), h_rts["brightness_temperature"])
    np.testing.assert_array_equal(d_k["temperature_k"].cpu().numpy(), h_k["temperature_k"])
    np.testing.assert_array_equal(d_k["absorber_k"].cpu().numpy(), h_k["absorber_k"])


def test_high_zenith_and_scan(contexts, coeff_dir, oracle_mod):
    """Zenith beyond the 63.6122 deg transmittance cap."""
    from crtmv3_amd import synthetic
    name = "hirs_syn"
    ctx = contexts[name]
    ps = synthetic.make_profiles(3, 100, seed=77)
    ps.sensor_zenith[:] = [63.0, 64.0, 72.0]
    eps = np.full((3, ctx.n_channels), 0.95)
    print("case high zenith", flush=True)
    ref = _oracle_batch(oracle_mod, coeff_dir, name, ps, eps)
    rts, kd = ctx.k_matrix(ps, eps, precision=0)
    _check(ps, rts, kd, ref, 0, ctx.n_channels)


def test_empty_batch_is_a_noop(contexts):
    from crtmv3_amd import synthetic
    ctx = contexts["amsua_syn"]
    ps = synthetic.make_profiles(0, 100, seed=1)
    out = ctx.forward(ps, np.zeros((0, ctx.n_channels)))
    assert out["brightness_temperature"].shape == (0, ctx.n_channels)


def test_invalid_inputs_are_reported(contexts):
    from crtmv3_amd import api, synthetic
    ctx = contexts["amsua_syn"]
    eps = np.full((2, ctx.n_channels), 0.6)
    ps = synthetic.make_profiles(2, [100, 201], seed=1)      # CRTM_MAX_N_LAYERS = 200
    with pytest.raises(api.CRTMError, match="n_layers"):
        ctx.forward(ps, eps)
    ps = synthetic.make_profiles(2, 100, seed=1)
    ps.temperature[150] = -1.0
    with pytest.raises(api.CRTMError, match="profile #2"):
        ctx.forward(ps, eps, check_input=True)
    # the context stays usable after an error
    ok = ctx.forward(synthetic.make_profiles(2, 100, seed=1), eps)
    assert np.all(np.isfinite(ok["brightness_temperature"]))


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "split"])
def test_reference_regression_profiles_with_added_layers(gpu, prec):
    """The reference regression test's two 92-layer profiles (top at
    0.714 hPa, so AddLayers extends them with climatology; mixed surface
    types) against the golden fixture."""
    import json
    import os
    import sys
    from conftest import ROOT
    from crtmv3_amd import api, synthetic
    gold = os.path.join(ROOT, "tests", "golden")
    sys.path.insert(0, gold)
    import make_golden
    doc = json.load(open(os.path.join(gold, "regression_inputs.json")))
    exp = np.load(os.path.join(gold, "regression_expected.npz"), allow_pickle=False)
    ps = make_golden.profile_set(doc["profiles"])
    d = synthetic.default_coeff_dir()
    tol = TOL[prec]
    for name in make_golden.SENSORS:
        synthetic.write_sensor(d, name)
        assert make_golden.fingerprint(d, name) == doc["coefficients"][name], "regenerate the golden fixture"
        ctx = api.crtm_init([name], d)
        L = ctx.n_channels
        eps = np.tile(make_golden.emissivity(ctx.channel_info[0].sensor_type, L), (2, 1))
        print(f"case golden {name} prec={prec}", flush=True)
        rts, kd = ctx.k_matrix(ps, eps, precision=prec, jacobian_f32=bool(prec))
        ctx.destroy()
        for m in range(2):
            bt_err = np.abs(rts["brightness_temperature"][m] - exp[f"{name}/{m}/bt"]).max()
            assert bt_err <= tol["bt"], (name, m, bt_err)
            a, b = 92 * m, 92 * (m + 1)
            for got, key in ((np.asarray(kd["temperature_k"][a:b], np.float64).T, "t_k"),
                             (np.asarray(kd["absorber_k"][:, a:b, :], np.float64).transpose(2, 0, 1), "abs_k")):
                want = exp[f"{name}/{m}/{key}"]
                assert_jacobian_close(got, want, tol["jac"], f"golden {name} profile {m} {key}", prec)
            sfc = np.stack([kd[f"{c}_temperature_k"][m] for c in ("land", "water", "snow", "ice")], axis=1)
            want = exp[f"{name}/{m}/sfc_t_k"]
            assert np.abs(sfc - want).max() <= tol["jac"] * np.abs(want).max()


def test_three_absorbers(contexts, coeff_dir, oracle_mod):
    """H2O, CO2 and O3 as user absorbers (CO2 Jacobian through the ODPS CO2
    component) on the group-1 sensor."""
    from crtmv3_amd import synthetic
    name = "cris_small_syn"
    ctx = contexts[name]
    ps = synthetic.make_profiles(3, [100, 64, 120], seed=41)
    co2 = np.full((1, ps.absorber.shape[1]), 395.0) * np.random.default_rng(2).uniform(0.97, 1.03, (1, ps.absorber.shape[1]))
    ps.absorber = np.concatenate([ps.absorber[:1], co2, ps.absorber[1:]])
    ps.absorber_id = np.array([1, 2, 3], np.int32)
    ps.absorber_units = np.array([3, 1, 1], np.int32)
    eps = np.full((3, ctx.n_channels), 0.97)
    print("case three absorbers", flush=True)
    ref = _oracle_batch(oracle_mod, coeff_dir, name, ps, eps)
    for prec in (0, 1):
        rts, kd = ctx.k_matrix(ps, eps, precision=prec, jacobian_f32=bool(prec))
        _check(ps, rts, kd, ref, prec, ctx.n_channels)
