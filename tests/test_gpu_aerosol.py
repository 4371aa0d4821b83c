"""Aerosols on the GPU against the oracle (through the C-ABI).

Infrared (HIRS-like) columns with dust, sea salt and sulfate in the lowest 20
layers, with and without clouds: forward radiances and the K-matrix including
the aerosol concentration and effective-radius Jacobians; microwave columns
ignore aerosols; a two-stream infrared call with aerosols fails as the
reference's CRTM_Compute_AerosolScatter does.  Tolerances as for clouds:
FP64 BT 1e-6 K / Jacobians 1e-8 relative, FP32-split 1e-2 K / 1e-3.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from conftest import assert_jacobian_close, load_sensor
from crtmv3_amd import api, coeffio, synthetic

pytestmark = pytest.mark.gpu

TOL = {api.PREC_FP64: dict(bt=1e-6, jac=1e-8), api.PREC_FP32_SPLIT: dict(bt=1e-2, jac=1e-3)}


@pytest.fixture(scope="module")
def ir(coeff_dir, oracle_mod):
    tau, spc = load_sensor(coeff_dir, "hirs_syn")
    cc = coeffio.read_cloudcoeff(synthetic.write_cloud_coeff(coeff_dir))
    ac = coeffio.read_aerosolcoeff(synthetic.write_aerosol_coeff(coeff_dir))
    ctx = api.crtm_init(["hirs_syn"], coeff_dir)
    yield ctx, oracle_mod.Sensor(tau, spc), oracle_mod.CloudCoeff(cc, ac)
    ctx.destroy()


def _profiles(seed, clouds=False, M=5):
    ps = synthetic.make_profiles(M, [100, 80, 100, 91, 100][:M], seed=seed)
    if clouds:
        ps = synthetic.add_clouds(ps, seed=seed + 1, median_gm3=0.005, cover="random")
    return synthetic.add_aerosols(ps, seed=seed + 2)


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
@pytest.mark.parametrize("clouds", [False, True], ids=["aerosol", "aerosol+cloud"])
@pytest.mark.parametrize("n_streams", [0, 8])
def test_aerosol_forward_matches_oracle(ir, oracle_mod, precision, clouds, n_streams):
    ctx, S, C = ir
    print(f"aerosol forward clouds={clouds} n_streams={n_streams} precision={precision}", flush=True)
    ps = _profiles(301, clouds)
    eps = np.full((ps.n_profiles, S.n_channels), 0.97)
    out = ctx.forward(ps, eps, precision=precision, n_streams=n_streams)
    ref = np.stack([oracle_mod.run_allsky(S, C, ps, m, eps[m], n_streams)["bt"] for m in range(ps.n_profiles)])
    clear = ctx.forward(dataclasses.replace(ps, aerosol_type=None, aerosol_concentration=None,
                                            aerosol_effective_radius=None), eps, precision=precision,
                        n_streams=n_streams)
    err = np.abs(out["brightness_temperature"] - ref).max()
    print(f"  max |dBT| {err:.2e} K; aerosol effect up to "
          f"{np.abs(out['brightness_temperature'] - clear['brightness_temperature']).max():.2f} K", flush=True)
    assert err < TOL[precision]["bt"]


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
@pytest.mark.parametrize("clouds", [False, True], ids=["aerosol", "aerosol+cloud"])
def test_aerosol_k_matrix_matches_oracle(ir, oracle_mod, precision, clouds):
    ctx, S, C = ir
    print(f"aerosol K clouds={clouds} precision={precision}", flush=True)
    ps = _profiles(311, clouds, M=4)
    L = S.n_channels
    eps = np.full((ps.n_profiles, L), 0.97)
    btk = np.random.default_rng(4).uniform(0.5, 1.5, (ps.n_profiles, L))
    rts, kd = ctx.k_matrix(ps, eps, bt_k=btk, precision=precision, jacobian_f32=bool(precision), n_streams=4)
    tol = TOL[precision]
    worst = {}
    for m in range(ps.n_profiles):
        f, k = oracle_mod.k_matrix_allsky(S, C, ps, m, eps[m], 4, 0, btk[m], None, method="adjoint")
        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        assert np.abs(rts["brightness_temperature"][m] - f["bt"]).max() < tol["bt"]
        blocks = [("T", kd["temperature_k"][a:b].T, k["t_k"]),
                  ("absorber", np.asarray(kd["absorber_k"][:, a:b, :]).transpose(2, 0, 1), k["abs_k"]),
                  ("aerosol concentration",
                   np.asarray(kd["aerosol_concentration_k"][:, a:b, :]).transpose(2, 0, 1), k["aconc_k"]),
                  ("aerosol radius",
                   np.asarray(kd["aerosol_effective_radius_k"][:, a:b, :]).transpose(2, 0, 1), k["areff_k"]),
                  ("Ts", kd["water_temperature_k"][m], k["sfc_t_k"][:, 1]),
                  ("emissivity", kd["surface_emissivity_k"][m], k["emis_k"])]
        if clouds:
            blocks += [("water content", np.asarray(kd["cloud_water_content_k"][:, a:b, :]).transpose(2, 0, 1),
                        k["wc_k"]), ("fraction", kd["cloud_fraction_k"][a:b].T, k["cf_k"])]
        for what, got, want in blocks:
            err, elem = assert_jacobian_close(got, want, tol["jac"], f"profile {m} {what}_K", precision)
            worst[what] = max(worst.get(what, (0.0, 0.0))[0], err), max(worst.get(what, (0.0, 0.0))[1], elem)
    print("  worst Jacobian errors (block-relative / element-wise): "
          + ", ".join(f"{k} {v[0]:.1e}/{v[1]:.1e}" for k, v in worst.items()), flush=True)
    assert np.abs(kd["aerosol_concentration_k"]).max() > 0


def test_microwave_ignores_aerosols(coeff_dir):
    print("aerosols in a microwave call", flush=True)
    synthetic.write_sensor(coeff_dir, "atms_syn")
    ctx = api.crtm_init(["atms_syn"], coeff_dir)
    try:
        ps = synthetic.make_profiles(4, 100, seed=321)
        eps = np.full((4, 22), 0.6)
        a = ctx.forward(synthetic.add_aerosols(ps, seed=322), eps, precision=api.PREC_FP64, n_streams=4)
        b = ctx.forward(ps, eps, precision=api.PREC_FP64)
        np.testing.assert_allclose(a["brightness_temperature"], b["brightness_temperature"], rtol=0, atol=1e-9)
    finally:
        ctx.destroy()


def test_aerosol_failures(ir):
    ctx, S, _ = ir
    print("aerosol failure cases", flush=True)
    ps = _profiles(331, M=3)
    eps = np.full((3, S.n_channels), 0.97)
    with pytest.raises(api.CRTMError, match="failed"):       # two streams: no aerosol phase function
        ctx.forward(ps, eps, precision=api.PREC_FP64, n_streams=2)
    bad = dataclasses.replace(ps, aerosol_type=ps.aerosol_type.copy())
    bad.aerosol_type[1, 0] = 42
    with pytest.raises(api.CRTMError, match="aerosol type"):
        ctx.forward(bad, eps, precision=api.PREC_FP64, n_streams=4)
    out = ctx.forward(ps, eps, precision=api.PREC_FP64, n_streams=4)   # the context is still usable
    assert np.isfinite(out["brightness_temperature"]).all()


def test_aerosol_failure_is_sticky_across_sensors(coeff_dir):
    """A profile that fails in an infrared sensor's aerosol step fails for the
    whole call, whatever the sensor order: FAILURE, status set, NaN
    brightness temperatures on every channel of every sensor, and NaN
    Jacobian rows; profiles without aerosols are computed."""
    import torch
    print("aerosol failure across sensors (hirs before atms), two streams", flush=True)
    synthetic.write_sensor(coeff_dir, "atms_syn")
    synthetic.write_sensor(coeff_dir, "hirs_syn")
    ctx = api.crtm_init(["hirs_syn", "atms_syn"], coeff_dir)
    try:
        ps = synthetic.add_aerosols(synthetic.make_profiles(6, 80, seed=351), seed=352, aerosol_fraction=0.5)
        fails = (ps.aerosol_type != 0).any(axis=1)
        assert fails.any() and (~fails).any()
        L = ctx.n_channels
        eps = np.full((6, L), 0.9)
        with pytest.raises(api.CRTMError, match="failed"):
            ctx.forward(ps, eps, precision=api.PREC_FP64, n_streams=2)
        dev = torch.device("cuda:0")
        dps = api.device_profiles(ps, dev)
        demis = torch.full((6, L), 0.9, dtype=torch.float64, device=dev)
        for prec in (api.PREC_FP64, api.PREC_FP32_SPLIT):
            rts = ctx.forward(dps, demis, precision=prec, n_streams=2)
            rk, kd = ctx.k_matrix(dps, demis, precision=prec, n_streams=2)
            torch.cuda.synchronize()
            for r in (rts, rk):
                st = r["status"].cpu().numpy()
                bt = r["brightness_temperature"].cpu().numpy()
                np.testing.assert_array_equal(st != 0, fails)
                assert np.isnan(bt[fails]).all()
                assert np.isfinite(bt[~fails]).all()
            tk = kd["temperature_k"].cpu().numpy()
            ak = kd["absorber_k"].cpu().numpy()
            for m in range(6):
                a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
                if fails[m]:
                    assert np.isnan(tk[a:b]).all() and np.isnan(ak[:, a:b]).all(), m
                else:
                    assert np.isfinite(tk[a:b]).all() and np.isfinite(ak[:, a:b]).all(), m
            with pytest.raises(api.CRTMError, match="failed"):
                ctx.adjoint(dps, demis, torch.ones((6, L), dtype=torch.float64, device=dev), precision=prec,
                            n_streams=2)
    finally:
        ctx.destroy()


def test_empty_aerosol_slots_two_streams(coeff_dir, oracle_mod):
    """Profiles whose aerosol slots are all unused (type 0) run at two
    streams next to profiles with aerosols at four, against the oracle."""
    print("aerosol empty slots, per-profile streams 2 / 4", flush=True)
    tau, spc = load_sensor(coeff_dir, "hirs_syn")
    cc = coeffio.read_cloudcoeff(synthetic.write_cloud_coeff(coeff_dir))
    ac = coeffio.read_aerosolcoeff(synthetic.write_aerosol_coeff(coeff_dir))
    S, C = oracle_mod.Sensor(tau, spc), oracle_mod.CloudCoeff(cc, ac)
    ctx = api.crtm_init(["hirs_syn"], coeff_dir)
    try:
        ps = synthetic.add_clouds(synthetic.make_profiles(6, 80, seed=241), seed=242, median_gm3=0.005)
        ps = synthetic.add_aerosols(ps, seed=243, aerosol_fraction=0.5)
        ns = np.where((ps.aerosol_type != 0).any(axis=1), 4, 2).astype(np.int32)
        assert (ns == 2).any() and (ns == 4).any()
        eps = np.full((6, 19), 0.97)
        out = ctx.forward(ps, eps, precision=api.PREC_FP64, n_streams=ns)
        for m in range(6):
            ref = oracle_mod.run_allsky(S, C, ps, m, eps[m], n_streams=int(ns[m]))
            assert np.abs(out["brightness_temperature"][m] - ref["bt"]).max() < 1e-6, m
    finally:
        ctx.destroy()
