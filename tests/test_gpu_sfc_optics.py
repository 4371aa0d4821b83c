"""Computed surface optics (Options%Use_Emissivity = .FALSE.) on the GPU
against the oracle: FASTEM6 for microwave sensors (every polarisation flag,
the azimuth model and the transmittance-dependent reflection correction) and
the IR sea-surface emissivity table (with a temperature axis), forward,
K-matrix (including Surface_K salinity / wind speed / wind direction and the
transmittance path into every layer's Jacobian) and CRTM_Adjoint, FP64 and
FP32-split, host and device memory.  The oracle's optics are pinned against
the compiled reference routines in test_oracle_sfc.py.

Tolerances: BT 1e-8 K (FP64) / 1e-2 K (split); Jacobians per variable and
profile relative to the block's largest element 1e-9 (FP64) / 1e-3 (split).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_sensor
from crtmv3_amd import api, coeffio, synthetic

pytestmark = pytest.mark.gpu

TOL = {api.PREC_FP64: 1e-9, api.PREC_FP32_SPLIT: 1e-3}
BT_TOL = {api.PREC_FP64: 1e-8, api.PREC_FP32_SPLIT: 1e-2}


def _rel(got, want):
    scale = np.abs(want).max()
    return np.abs(np.asarray(got) - want).max() / scale if scale > 0 else np.abs(got).max()


@pytest.fixture(scope="module")
def sfc_setup(coeff_dir, oracle_mod):
    synthetic.write_mwwater_coeff(coeff_dir)
    synthetic.write_irwater_coeff(coeff_dir)
    mw = oracle_mod.MWwater(coeffio.read_mwwatercoeff(f"{coeff_dir}/{synthetic.MWWATERCOEFF_FILE}"))
    ir = oracle_mod.IRwater(coeffio.read_irwatercoeff(f"{coeff_dir}/{synthetic.IRWATERCOEFF_FILE}"))
    out = {}
    for name in ("amsua_syn", "atms_syn", "hirs_syn", "cris_small_syn"):
        synthetic.write_sensor(coeff_dir, name)
        tau, spc = load_sensor(coeff_dir, name)
        out[name] = (api.crtm_init([name], coeff_dir), oracle_mod.Sensor(tau, spc), spc)
    yield out, mw, ir
    for c, _, _ in out.values():
        c.destroy()


def _profiles(M, seed):
    ps = synthetic.make_profiles(M, [100, 80, 100, 91, 77, 100][:M], seed=seed)
    ps = synthetic.ocean_surface(ps, seed=seed + 1)
    ps.sensor_zenith[:] = np.linspace(3.0, 58.0, M)
    ps.sensor_scan[:] = np.degrees(np.arcsin(np.sin(np.radians(ps.sensor_zenith)) * 6370.0 / 7203.0)) * \
        np.where(np.arange(M) % 2, -1.0, 1.0)
    ps.sensor_azimuth[1] = 999.9      # azimuth model off for one profile
    ps.wind_speed[2] = 19.0           # past the 15 / 18 m/s clamps of the azimuth model
    return ps


def _so(oracle_mod, spc, mw, ir, ps, m):
    return oracle_mod.SurfaceOptics(spc, mw=mw if spc.sensor_type == 1 else None,
                                    ir=ir if spc.sensor_type != 1 else None, salinity=ps.salinity[m],
                                    wind_speed=ps.wind_speed[m], wind_direction=ps.wind_direction[m],
                                    sensor_azimuth=ps.sensor_azimuth[m])


@pytest.mark.parametrize("memory", ["host", "device"])
@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
@pytest.mark.parametrize("name", ["amsua_syn", "atms_syn", "hirs_syn", "cris_small_syn"])
def test_k_matrix_with_computed_optics_matches_oracle(sfc_setup, oracle_mod, name, precision, memory):
    import torch
    sensors, mw, ir = sfc_setup
    ctx, S, spc = sensors[name]
    print(f"computed optics K {name} precision={precision} memory={memory}", flush=True)
    M, L = 5, S.n_channels
    ps = _profiles(M, 701)
    rng = np.random.default_rng(71)
    btk = rng.normal(size=(M, L))
    if memory == "device":
        dps = api.device_profiles(ps)
        rts, k = ctx.k_matrix(dps, None, bt_k=torch.as_tensor(btk, device="cuda"), precision=precision)
        torch.cuda.synchronize()
        rts = {a: v.cpu().numpy() for a, v in rts.items()}
        k = {a: v.cpu().numpy() for a, v in k.items()}
        fw = ctx.forward(dps, None, precision=precision, extras=True)
        fw = {a: v.cpu().numpy() for a, v in fw.items()}
    else:
        rts, k = ctx.k_matrix(ps, None, bt_k=btk, precision=precision)
        fw = ctx.forward(ps, None, precision=precision, extras=True)
    assert np.all(rts["status"] == 0)
    worst = {}
    for m in range(M):
        so = _so(oracle_mod, spc, mw, ir, ps, m)
        f = oracle_mod.run_profile_sfc(S, ps, m, so, "forward")
        r = oracle_mod.run_profile_sfc(S, ps, m, so, "k", bt_k=btk[m])
        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        assert np.abs(rts["brightness_temperature"][m] - f["bt"]).max() < BT_TOL[precision], m
        assert np.abs(fw["brightness_temperature"][m] - f["bt"]).max() < BT_TOL[precision], m
        etol = 1e-12 if precision == api.PREC_FP64 else 1e-6
        for key in ("emis", "refl"):
            got = fw["surface_emissivity" if key == "emis" else "surface_reflectivity"][m]
            assert np.all(np.abs(got - f[key]) <= etol * np.maximum(1.0, np.abs(f[key]))), (m, key)
        assert np.all(np.abs(f["refl"]) < 2.0), f["refl"]    # physically sized synthetic coefficients
        blocks = [("T", k["temperature_k"][a:b].T, r["t_k"]),
                  ("absorber", np.moveaxis(k["absorber_k"][:, a:b], 2, 0), r["abs_k"]),
                  ("Ts", k["water_temperature_k"][m], r["sfc_t_k"][:, 1]),
                  ("wind speed", k["wind_speed_k"][m], r["sfc_k"][:, 1])]
        if spc.sensor_type == 1:
            blocks += [("salinity", k["salinity_k"][m], r["sfc_k"][:, 0]),
                       ("wind direction", k["wind_direction_k"][m], r["sfc_k"][:, 2])]
        else:
            assert np.all(k["salinity_k"][m] == 0) and np.all(k["wind_direction_k"][m] == 0)
        for key, got, want in blocks:
            if np.abs(want).max() == 0:
                assert np.abs(got).max() == 0, (m, key)
                continue
            err = _rel(got, want)
            worst[key] = max(worst.get(key, 0.0), err)
            assert err <= TOL[precision], (m, key, err)
        assert np.all(k["surface_emissivity_k"][m] == 0)
    print("  worst: " + ", ".join(f"{a} {v:.1e}" for a, v in worst.items()), flush=True)


@pytest.mark.parametrize("precision", [api.PREC_FP64, api.PREC_FP32_SPLIT], ids=["fp64", "split"])
@pytest.mark.parametrize("name", ["atms_syn", "hirs_syn"])
def test_adjoint_with_computed_optics_matches_oracle(sfc_setup, oracle_mod, name, precision):
    sensors, mw, ir = sfc_setup
    ctx, S, spc = sensors[name]
    print(f"computed optics adjoint {name} precision={precision}", flush=True)
    M, L = 4, S.n_channels
    ps = _profiles(M, 711)
    bt_ad = np.random.default_rng(72).normal(size=(M, L))
    bt, ad = ctx.adjoint(ps, None, bt_ad, precision=precision)
    tol = TOL[precision]
    for m in range(M):
        so = _so(oracle_mod, spc, mw, ir, ps, m)
        r = oracle_mod.run_profile_sfc(S, ps, m, so, "k", bt_k=bt_ad[m])
        a, b = int(ps.layer_offset[m]), int(ps.layer_offset[m + 1])
        assert np.abs(bt[m] - r["bt"]).max() < BT_TOL[precision]
        for key, got, want in (("T", ad["temperature_ad"][a:b], r["t_k"].sum(0)),
                               ("absorber", ad["absorber_ad"][:, a:b], r["abs_k"].sum(0)),
                               ("Ts", ad["water_temperature_ad"][m], r["sfc_t_k"][:, 1].sum()),
                               ("salinity", ad["salinity_ad"][m], r["sfc_k"][:, 0].sum()),
                               ("wind speed", ad["wind_speed_ad"][m], r["sfc_k"][:, 1].sum()),
                               ("wind direction", ad["wind_direction_ad"][m], r["sfc_k"][:, 2].sum())):
            want = np.asarray(want)
            if np.abs(want).max() == 0:
                assert np.abs(got).max() == 0, (m, key)
                continue
            assert _rel(got, want) <= tol, (m, key, _rel(got, want))


def test_computed_optics_skip_and_failures(sfc_setup, coeff_dir, tmp_path):
    """Skip_Profile; a land surface fails its profile (host: the call, with a
    message); clouds with computed optics and a missing coefficient file are
    refused."""
    import shutil
    sensors, _, _ = sfc_setup
    ctx, S, _ = sensors["atms_syn"]
    print("computed optics: skip / land / clouds / missing file", flush=True)
    ps = _profiles(4, 721)
    rts, k = ctx.k_matrix(ps, None, skip_profile=np.array([0, 1, 0, 0], np.int32))
    assert np.all(rts["brightness_temperature"][1] == 0) and np.all(k["wind_speed_k"][1] == 0)
    full, kf = ctx.k_matrix(ps, None)
    np.testing.assert_array_equal(rts["brightness_temperature"][[0, 2, 3]], full["brightness_temperature"][[0, 2, 3]])
    land = synthetic.ProfileSet(**{f: getattr(ps, f) for f in ps.__dataclass_fields__})
    land.land_coverage = np.array([0.0, 0.3, 0.0, 0.0])
    land.water_coverage = np.array([1.0, 0.7, 1.0, 1.0])
    with pytest.raises(api.CRTMError, match="water"):
        ctx.forward(land, None)
    cloudy = synthetic.add_clouds(_profiles(4, 722), seed=723)
    with pytest.raises(api.CRTMError, match="clear-sky"):
        ctx.forward(cloudy, None)
    # device memory: the land profile fails, the others are computed
    import torch
    dl = api.device_profiles(land)
    out = ctx.forward(dl, None)
    torch.cuda.synchronize()
    st = out["status"].cpu().numpy()
    assert st[1] != 0 and np.all(st[[0, 2, 3]] == 0)
    bt = out["brightness_temperature"].cpu().numpy()
    np.testing.assert_allclose(bt[[0, 2, 3]], full["brightness_temperature"][[0, 2, 3]], rtol=0, atol=1e-10)
    d = tmp_path / "nofastem"
    d.mkdir()
    for suffix in ("SpcCoeff.bin", "TauCoeff.bin"):
        shutil.copy(f"{coeff_dir}/atms_syn.{suffix}", d / f"atms_syn.{suffix}")
    c2 = api.crtm_init(["atms_syn"], str(d))
    try:
        with pytest.raises(api.CRTMError, match="FASTEM6"):
            c2.forward(ps, None)
        c2.forward(ps, np.full((4, c2.n_channels), 0.6))     # user emissivity still works
    finally:
        c2.destroy()
